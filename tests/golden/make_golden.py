#!/usr/bin/env python3
"""Writes tests/golden/ref_vectors.json: input/output vectors of the compiled
CAFE5 reference (oracle/_ref/libcaferef.so, built by `make -C oracle ref` from
the CAFE5 sources) for every oracle function, so that the oracle stays pinned
where the reference tree is not available.

    python tests/golden/make_golden.py

Every case stores its inputs (tree arrays, counts, parameters) with the
reference's outputs; floats round-trip exactly through JSON (repr).
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle_lib import Oracle, Ref  # noqa: E402
from cafe5_amd import workloads as wl  # noqa: E402


def tree_json(t: wl.Tree):
    return {"parent": t.parent.tolist(), "leaf_col": t.leaf_col.tolist(), "length": t.length.tolist(),
            "names": list(t.names)}


def main():
    if not Ref.available():
        sys.exit("oracle/_ref/libcaferef.so is not built (make -C oracle ref)")
    out = {"source": "compiled CAFE5 reference (oracle/_ref) via oracle/ref_driver.cpp"}

    # matrix_cache::precalculate_matrices entries (matrix_cache.cpp:113-163)
    mats = []
    for lam, t, N in [(0.002, 12.345, 24), (0.0137, 3.3, 20), (0.0005, 61.7, 16), (0.3, 7.0, 8)]:
        mats.append({"lambda": lam, "t": t, "N": N, "P": Ref.matrix(lam, t, N).tolist()})
    out["matrices"] = mats

    # inference_prune root vectors (core.cpp:134-145), with and without an error model
    tree = wl.yule_tree(9, 11)
    B = R = 30
    err = wl.error_table(0.05, B)
    rng = np.random.default_rng(3)
    prunes = []
    for use_err in (False, True):
        for _ in range(3):
            counts = rng.integers(0, B - 1, tree.n_leaves)
            prunes.append({"counts": counts.tolist(), "lambda": 0.004, "B": B, "R": R,
                           "error_model": err.tolist() if use_err else None,
                           "root": Ref.prune(tree, counts, 0.004, 1.0, B, R, err if use_err else None).tolist()})
    out["prune"] = {"tree": tree_json(tree), "cases": prunes}

    # the same at bound 1,000, where the reference runs its heap path
    # (compute_node_probability_large_families, probability.cpp:239-310)
    tree = wl.parse_newick("((A:5,B:5):5,(C:5,D:5):5);")
    counts = np.array([960, 975, 990, 999])
    out["prune_heap"] = {"tree": tree_json(tree), "counts": counts.tolist(), "lambda": 0.002, "B": 1000, "R": 1000,
                         "root": Ref.prune(tree, counts, 0.002, 1.0, 1000, 1000).tolist()}

    # base_model::infer_family_likelihoods (base_model.cpp:53-100): uniform prior
    w = wl.config_base(seed=1, F=60)
    B = R = 40
    counts = np.minimum(w.counts, B)
    total, lnl = Ref.base_model(w.tree, counts, [0.002], B, R, prior_kind=0, prior_max=R)
    out["base_uniform"] = {"tree": tree_json(w.tree), "counts": counts.tolist(), "lambda": 0.002, "B": B, "R": R,
                           "prior": Ref.prior(0, R, 0.0, R).tolist(), "lnl": lnl.tolist(), "total": total}

    # multi-lambda + error model + Poisson prior
    w = wl.config_error_multi(seed=4, F=40, bound=40)
    B = R = 40
    err = wl.error_table(0.05, B)
    total, lnl = Ref.base_model(w.tree, w.counts, w.lambdas, B, R, prior_kind=1, prior_max=R, poisson_lambda=3.5,
                                err=err, node_class=w.node_class)
    out["base_multi_error_poisson"] = {
        "tree": tree_json(w.tree), "counts": w.counts.tolist(), "node_lambda": w.lambdas[w.node_class].tolist(),
        "B": B, "R": R, "prior": Ref.prior(1, R, 3.5, R).tolist(), "error_model": err.tolist(),
        "lnl": lnl.tolist(), "total": total}

    # gamma_model::infer_family_likelihoods (gamma_core.cpp:168-237)
    w = wl.config_gamma(seed=3, F=24, K=4, bound=30)
    B = R = 30
    f, r = Ref.get_gamma(4, 0.5)
    total, cat, fam, post = Ref.gamma_model(w.tree, w.counts, [0.003], B, R, f, r, prior_kind=0, prior_max=R)
    out["gamma"] = {"tree": tree_json(w.tree), "counts": w.counts.tolist(), "lambda": 0.003, "B": B, "R": R,
                    "alpha": 0.5, "cat_probs": f.tolist(), "multipliers": r.tolist(),
                    "prior": Ref.prior(0, R, 0.0, R).tolist(), "cat": cat.tolist(), "fam": fam.tolist(),
                    "post": post.tolist(), "total": total}

    # get_gamma (gamma.cpp:15-241)
    gam = []
    for K in (2, 4, 8):
        for alpha in (0.1, 0.5, 1.0, 2.0, 5.0):
            f, r = Ref.get_gamma(K, alpha)
            gam.append({"K": K, "alpha": alpha, "probs": f.tolist(), "multipliers": r.tolist()})
    out["discrete_gamma"] = gam

    # root_equilibrium_distribution (root_equilibrium_distribution.cpp:15-87)
    out["priors"] = {
        "uniform": [{"max": m, "R": m + 3, "p": Ref.prior(0, m, 0.0, m + 3).tolist()} for m in (8, 20)],
        "poisson": [{"lambda": lam, "max": 20, "R": 30, "p": Ref.prior(1, 20, lam, 30).tolist()}
                    for lam in (0.75, 3.0, 9.5)],
        "map": {"sizes": [2, 3, 4, 6], "counts": [11, 5, 7, 2], "R": 9,
                "p": Ref.prior(2, 0, 0.0, 9, [2, 3, 4, 6], [11, 5, 7, 2]).tolist()},
    }
    path = os.path.join(HERE, "ref_vectors.json")
    with open(path, "w") as fh:
        json.dump(out, fh, separators=(",", ":"))
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
