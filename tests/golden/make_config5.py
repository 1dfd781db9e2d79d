#!/usr/bin/env python3
"""Writes tests/golden/config5_sample.json: the oracle's gamma-model results
for a seeded 200-family sample of the optimizer-scale workload (seeded Yule
tree of 100 taxa, 200,000 families, size bound 500, K = 8 gamma categories,
200 seeded (lambda, alpha) trial points) at three finite trial points.

    python tests/golden/make_config5.py

The oracle builds every matrix in the reference's closed form
(probability.cpp:104-148): 1,584 matrices of 501 x 501 per trial point, about
6.7e10 exp/pow terms, minutes of CPU time -- too slow for a GPU test, hence
these vectors.  Each trial stores lambda, alpha, the K category frequencies
and multipliers (gamma.cpp), the sample's -lnL and every family's lnL and
category posteriors (gamma_core.cpp:97-109, :196-207).  Floats round-trip
exactly through JSON (repr)."""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle_lib import Oracle  # noqa: E402
from cafe5_amd import workloads as wl  # noqa: E402

F, K, B, T = 200000, 8, 500, 200
N_SAMPLE = 200


def finite_trials(w, longest):
    """Trial points where no category saturates the longest branch
    (matrix_cache.cpp:107-111); elsewhere the gamma model is +inf."""
    ok = []
    for t, (lam, a) in enumerate(zip(w.lambdas, w.alphas)):
        _, r = Oracle.get_gamma(K, float(a))
        lt = lam * float(r.max()) * longest
        if lt / (1 + lt) < 0.5:
            ok.append(t)
    return ok


def main():
    w = wl.config_optimizer(seed=5, F=F, K=K, bound=B, T=T)
    rng = np.random.Generator(np.random.PCG64(51))
    sample = np.sort(rng.choice(F, size=N_SAMPLE, replace=False))
    longest = max(float(x) for i, x in enumerate(w.tree.length) if i != w.tree.root)
    ok = finite_trials(w, longest)
    # small alpha gives categories with tiny rates, where families with large
    # count changes underflow to a zero likelihood and the evaluation is +inf
    # (gamma_core.cpp:151-152): the vectors use points with alpha >= 1
    moderate = [t for t in ok if w.alphas[t] >= 1.0]
    picks = [moderate[0], moderate[len(moderate) // 2], moderate[-1]]
    prior = Oracle.prior_uniform(B, B)
    out = {"F": F, "K": K, "B": B, "T": T, "seed": 5, "sample": sample.tolist(), "trials": []}
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    for t in picks:
        lam, alpha = float(w.lambdas[t]), float(w.alphas[t])
        f, r = Oracle.get_gamma(K, alpha)
        ref = Oracle.evaluate(w.tree, np.full(w.tree.n_nodes, lam), w.counts[sample], B, B, prior, n_cat=K,
                              multipliers=r, cat_probs=f, threads=threads)
        print(f"trial {t}: lambda={lam:.6g} alpha={alpha:.4g} -lnL(sample)={ref['total']!r}", flush=True)
        out["trials"].append({"t": t, "lambda": lam, "alpha": alpha, "freq": f.tolist(), "rate": r.tolist(),
                              "total": ref["total"], "lnl": ref["lnl"].tolist(), "post": ref["post"].tolist()})
    with open(os.path.join(HERE, "config5_sample.json"), "w") as fh:
        json.dump(out, fh)


if __name__ == "__main__":
    main()
