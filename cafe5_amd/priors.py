"""Root-size priors as root_equilibrium_distribution builds them
(src/root_equilibrium_distribution.cpp:15-87).

compute(j) returns a float in the reference (.h:40); the models index it with j
for root size j+1 (base_model.cpp:84, gamma_core.cpp:156).  Each function
returns double(float(compute(j))) for j < R, the array the engine takes.
"""
from __future__ import annotations

import math
from typing import Dict

import numpy as np


def uniform(max_size: int, R: int) -> np.ndarray:
    """root_equilibrium_distribution(size_t max_size): sizes 0..max_size-1
    each with frequency float(1)/float(max_size) (.cpp:30-35, 70-79)."""
    pct = np.float32(1.0) / np.float32(max_size)
    out = np.zeros(R)
    out[: min(R, max_size)] = float(pct)
    return out


def _poisspdf(x: int, lam: float) -> float:
    # poisson.cpp:21-24
    return math.exp(x * math.log(lam) - math.lgamma(x + 1) - lam)


def poisson(poisson_lambda: float, num_values: int, R: int) -> np.ndarray:
    """create_from_poisson (.cpp:56-68): shifted Poisson percentages until
    ceil(pct * num_values) copies fill num_values, plus five more."""
    pct = []
    filled = 0
    i = 0
    while filled < num_values:
        p = _poisspdf(i, poisson_lambda)
        filled += math.ceil(p * num_values) if p * num_values > 0 else 0
        pct.append(p)
        i += 1
    for _ in range(5):
        pct.append(_poisspdf(len(pct), poisson_lambda))
    out = np.zeros(R)
    n = min(R, len(pct))
    out[:n] = np.asarray(pct[:n], dtype=np.float32).astype(np.float64)
    return out


def from_counts(root_distribution: Dict[int, int], R: int) -> np.ndarray:
    """root_equilibrium_distribution(map) (.cpp:15-28 + build_percentages)."""
    total = sum(root_distribution.values())
    sizes = [s for s, c in root_distribution.items() if c > 0]
    top = max(sizes) if sizes else -1
    out = np.zeros(R)
    for j in range(min(R, top + 1)):
        out[j] = float(np.float32(root_distribution.get(j, 0)) / np.float32(total))
    return out
