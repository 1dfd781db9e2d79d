"""Nelder-Mead as CAFE5's optimizer runs it (src/optimizer.cpp:61-330, the
Standard strategy of :365-388), restated in Python.

TEST INFRASTRUCTURE ONLY: the checker the C++ optimizer mirror is compared
against.  Only tests/ may import it; the product never does.

Each method follows the step of the reference it names (file:line); the
arithmetic is in the reference's order so that the same scorer gives the same
sequence of evaluated points.
"""
from __future__ import annotations

import math
from typing import Callable, List, Sequence


class Candidate:
    """optimizer.h:34-39: a point and its score."""

    def __init__(self, n: int):
        self.values = [0.0] * n
        self.score = 0.0


class NelderMead:
    """FMinSearch (optimizer.h:44-63) with the defaults of fminsearch_new
    (optimizer.cpp:61-74): reflection 1, expansion 2, contraction 0.5,
    shrink 0.5, tolerances 1e-6, initial steps 5% (0.00025 from zero)."""

    def __init__(self, scorer: Callable[[Sequence[float]], float], n: int):
        self.scorer = scorer
        self.n = n
        self.rho, self.chi, self.psi, self.sigma = 1.0, 2.0, 0.5, 0.5
        self.tolx = self.tolf = 1e-6
        self.delta, self.zero_delta = 0.05, 0.00025
        self.maxiters = 250
        self.iters = 0
        self.candidates: List[Candidate] = [Candidate(n) for _ in range(n + 1)]
        self.x_mean = [0.0] * n
        self.x_r = [0.0] * n
        self.x_tmp = [0.0] * n

    def score(self, x: Sequence[float]) -> float:
        return self.scorer(list(x[: self.n]))

    # optimizer.cpp:119-124 (std::sort of n+1 <= 16 candidates: insertion
    # sort, so equal scores keep their order, as Python's sort does)
    def sort(self):
        self.candidates.sort(key=lambda c: c.score)

    # :127-145
    def check_values(self) -> bool:
        big = -math.inf
        for i in range(self.n):
            for j in range(self.n):
                big = max(big, abs(self.candidates[i + 1].values[j] - self.candidates[i].values[j]))
        return big <= self.tolx

    # :147-160
    def check_scores(self) -> bool:
        big = -math.inf
        for i in range(1, self.n + 1):
            big = max(big, abs(self.candidates[i].score - self.candidates[0].score))
        return big <= self.tolf

    def threshold_achieved(self) -> bool:   # :603-606
        return self.check_values() and self.check_scores()

    # :162-196
    def min_init(self, x0: Sequence[float]):
        for i in range(self.n + 1):
            c = self.candidates[i]
            for j in range(self.n):
                if i - 1 == j:
                    if i > 1 and math.isinf(self.candidates[i - 1].score):
                        c.values[j] = (1 + self.delta * 100) * x0[j] if x0[j] else self.zero_delta
                    else:
                        c.values[j] = (1 + self.delta) * x0[j] if x0[j] else self.zero_delta
                else:
                    c.values[j] = x0[j]
            c.score = self.score(c.values)
        self.sort()

    # :198-210: mean of the n best points
    def mean(self):
        for i in range(self.n):
            s = 0.0
            for j in range(self.n):
                s += self.candidates[j].values[i]
            self.x_mean[i] = s / self.n

    def reflection(self) -> float:           # :212-220
        worst = self.candidates[self.n].values
        for i in range(self.n):
            self.x_r[i] = self.x_mean[i] + self.rho * (self.x_mean[i] - worst[i])
        return self.score(self.x_r)

    def expansion(self) -> float:            # :223-231
        for i in range(self.n):
            self.x_tmp[i] = self.x_mean[i] + self.chi * (self.x_r[i] - self.x_mean[i])
        return self.score(self.x_tmp)

    def contract_outside(self) -> float:     # :233-241
        for i in range(self.n):
            self.x_tmp[i] = self.x_mean[i] + self.psi * (self.x_r[i] - self.x_mean[i])
        return self.score(self.x_tmp)

    def contract_inside(self) -> float:      # :243-251
        worst = self.candidates[self.n].values
        for i in range(self.n):
            self.x_tmp[i] = self.x_mean[i] + self.psi * (self.x_mean[i] - worst[i])
        return self.score(self.x_tmp)

    def shrink(self):                         # :253-265
        best = self.candidates[0].values
        for i in range(1, self.n + 1):
            c = self.candidates[i]
            for j in range(self.n):
                c.values[j] = best[j] + self.sigma * (c.values[j] - best[j])
            c.score = self.score(c.values)
        self.sort()

    def set_last(self, x: Sequence[float], f: float):   # :268-274
        c = self.candidates[-1]
        c.values[: self.n] = list(x[: self.n])
        c.score = f
        self.sort()

    # :276-322
    def minimize(self, x0: Sequence[float]) -> bool:
        self.min_init(x0)
        i = 0
        while i < self.maxiters:
            if self.threshold_achieved():
                break
            self.mean()
            fr = self.reflection()
            last = self.candidates[self.n].score
            if fr < self.candidates[0].score:
                fe = self.expansion()
                if fe < fr:
                    self.set_last(self.x_tmp, fe)
                else:
                    self.set_last(self.x_r, fr)
            elif fr >= last:
                if fr > last:
                    fc = self.contract_inside()
                    if fc < self.candidates[self.n].score:
                        self.set_last(self.x_tmp, fc)
                    else:
                        self.shrink()
                else:
                    fc = self.contract_outside()
                    if fc <= fr:
                        self.set_last(self.x_tmp, fc)
                    else:
                        self.shrink()
            else:
                self.set_last(self.x_r, fr)
            i += 1
        self.iters = i
        return i == self.maxiters


def optimize(scorer: Callable[[Sequence[float]], float], x0: Sequence[float], iterations: int = 300,
             expansion: float = 2.0, reflection: float = 1.0):
    """optimizer::optimize with the Standard strategy (optimizer.cpp:365-388,
    :546-573, :608-614): the parameters' expansion, reflection and iteration
    cap, tolerances 1e-6, from x0.  Returns (values, score, iterations)."""
    nm = NelderMead(scorer, len(x0))
    nm.chi, nm.rho, nm.maxiters = expansion, reflection, iterations
    nm.tolx = nm.tolf = 1e-6
    nm.minimize(x0)
    best = nm.candidates[0]
    return list(best.values), best.score, nm.iters
