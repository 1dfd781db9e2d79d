"""base_model / gamma_model likelihood evaluation over the GPU engine.

Mirrors the reference's `infer_family_likelihoods` (src/base_model.cpp:53-100,
src/gamma_core.cpp:168-237): same arguments, same return value (-lnL, +inf
for invalid parameters or a saturated gamma family) and the same per-family
records.  Matrix keys are quantized exactly as matrix_cache_key does, so a
branch uses the matrix of (trunc(lambda 1e9)/1e9, trunc(t 1e3)/1e3).

Several parameter points (a lambda grid, or an optimizer's trial points) can be
evaluated in one call: the engine batches all of them in one launch.
"""
from __future__ import annotations

import math
from typing import Optional, Sequence

import numpy as np

from . import native
from .workloads import Tree


def is_saturated(t: float, lam: float) -> bool:
    """matrix_cache::is_saturated (matrix_cache.cpp:107-111)."""
    alpha = lam * t / (1 + lam * t)
    return (1 - 2 * alpha) < 0


class _ModelBase:
    def __init__(self, engine: native.Engine, tree: Tree, counts: np.ndarray, max_family_size,
                 max_root_family_size, error_model: Optional[np.ndarray] = None,
                 path: int = native.PATH_BATCHED32, build_algo: int = native.BUILD_RECURRENCE):
        self.engine = engine
        self.tree = tree
        self.counts = np.ascontiguousarray(counts, dtype=np.int32)
        self.F = self.counts.shape[0]
        self.B = np.broadcast_to(np.asarray(max_family_size, dtype=np.int32), (self.F,)).copy()
        self.R = np.broadcast_to(np.asarray(max_root_family_size, dtype=np.int32), (self.F,)).copy()
        self.path = path
        self.build_algo = build_algo
        self.error_model = error_model
        engine.set_tree(tree)
        engine.set_families(self.counts, bound_B=self.B, bound_R=self.R)
        engine.set_error_model(error_model)
        self.buckets = engine.buckets()
        self.N = int(max(self.B.max(), self.R.max())) + 1
        self.last = None

    def _prior_table(self, prior) -> np.ndarray:
        """prior: one vector (indexed by j, as root_equilibrium_distribution::
        compute), or a callable R -> vector giving each bucket's prior."""
        stride = max(r for _, r, _ in self.buckets)
        tab = np.zeros((len(self.buckets), stride))
        for i, (_, r, _) in enumerate(self.buckets):
            v = prior(r) if callable(prior) else np.asarray(prior, dtype=np.float64)
            n = min(r, len(v))
            tab[i, :n] = v[:n]
        return tab

    def _node_lambdas(self, lambdas: np.ndarray, node_class: Optional[np.ndarray]) -> np.ndarray:
        """(T, n_nodes) lambda of each branch at each trial point."""
        lambdas = np.atleast_2d(np.asarray(lambdas, dtype=np.float64))
        if node_class is None:
            return np.repeat(lambdas[:, :1], self.tree.n_nodes, axis=1)
        return lambdas[:, node_class]

    def _run(self, node_lams: np.ndarray, prior, model: int, cat_probs=None):
        keys = native.KeyTable()
        slots = keys.node_slots(self.tree, node_lams)
        lam_q, t_q = keys.arrays()
        eng = self.engine
        if eng.N != self.N or getattr(eng, "_capacity", 0) < len(lam_q):
            eng.matrix_reset(self.N, max(len(lam_q), 1))
            eng._capacity = max(len(lam_q), 1)
        eng.matrix_build(lam_q, t_q, 0, self.build_algo)
        self.build_ms = eng.timings()[0]
        res = eng.evaluate(slots, self._prior_table(prior), self.path, model, cat_probs)
        self.prune_ms = eng.timings()[1:]
        self.last = res
        return res


class BaseModel(_ModelBase):
    """base_model (src/base_model.h)."""

    def infer_family_likelihoods(self, prior, lam, node_class: Optional[np.ndarray] = None) -> float:
        """-lnL at one lambda (scalar, or per-class values with node_class).
        Per-family lnL (the stash's posterior_probability) in self.results."""
        lam = np.atleast_1d(np.asarray(lam, dtype=np.float64))
        valid = lam[0] > 0 if node_class is None else not np.any(lam < 0)   # lambda.h:58-60, lambda.cpp:58-61
        if not valid:
            self.results = None
            return math.inf
        res = self._run(self._node_lambdas(lam[None, :], node_class), prior, native.MODEL_BASE)
        self.results = res["lnl"][0]
        return float(res["total"][0])

    def infer_grid(self, prior, lambdas: Sequence[float]) -> np.ndarray:
        """-lnL at every lambda of a grid, evaluated in one batched launch
        (single-lambda model); per-family values in self.grid_lnl (T, F)."""
        lam = np.asarray(lambdas, dtype=np.float64)
        res = self._run(self._node_lambdas(lam[:, None], None), prior, native.MODEL_BASE)
        totals = res["total"].copy()
        totals[lam <= 0] = math.inf
        self.grid_lnl = res["lnl"]
        return totals


class GammaModel(_ModelBase):
    """gamma_model with explicit categories (src/gamma_core.h constructor 2)."""

    def __init__(self, engine, tree, counts, max_family_size, max_root_family_size, cat_probs, multipliers,
                 error_model=None, path=native.PATH_BATCHED32, build_algo=native.BUILD_RECURRENCE, alpha=0.0):
        super().__init__(engine, tree, counts, max_family_size, max_root_family_size, error_model, path,
                         build_algo)
        self.cat_probs = np.asarray(cat_probs, dtype=np.float64)
        self.multipliers = np.asarray(multipliers, dtype=np.float64)
        self.alpha = alpha

    def can_infer(self, lam: np.ndarray, node_class) -> bool:
        """gamma_core.cpp:123-141: valid lambda, alpha >= 0, and the largest
        (un-multiplied) lambda does not saturate the longest branch."""
        valid = lam[0] > 0 if node_class is None else not np.any(lam < 0)
        if not valid or self.alpha < 0:
            return False
        lengths = [float(x) for i, x in enumerate(self.tree.length) if i != self.tree.root and x > 0]
        return not is_saturated(max(lengths), float(np.max(lam)))

    def infer_family_likelihoods(self, prior, lam, node_class: Optional[np.ndarray] = None) -> float:
        lam = np.atleast_1d(np.asarray(lam, dtype=np.float64))
        if not self.can_infer(lam, node_class):
            self.results = None
            return math.inf
        base = self._node_lambdas(lam[None, :], node_class)[0]
        node_lams = np.stack([base * m for m in self.multipliers])
        res = self._run(node_lams, prior, native.MODEL_GAMMA, self.cat_probs)
        self.results = {"cat": res["cat"][0], "post": res["post"][0], "lnl": res["lnl"][0],
                        "fam": np.exp(res["lnl"][0])}
        return float(res["total"][0])
