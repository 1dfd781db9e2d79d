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

import ctypes
import math
import os
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


# --------------------------------------------------------------------------
# discrete gamma categories through the C++ mirror (gamma.cpp:15-52)
# --------------------------------------------------------------------------
_HOST = None


def _host_library():
    global _HOST
    if _HOST is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libcafe5_host.so")
        if not os.path.exists(path):
            raise native.CafeGpuError(f"{path} is not built (python -m cafe5_amd.build)")
        lib = ctypes.CDLL(path)
        lib.cafe5_get_gamma.restype = ctypes.c_int
        lib.cafe5_get_gamma.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_double)]
        _HOST = lib
    return _HOST


def get_gamma(K: int, alpha: float):
    """(frequencies, rate multipliers) of K median-discretised gamma
    categories of shape alpha, the PAML routines CAFE5 uses (gamma.cpp)."""
    freq = np.zeros(K)
    rate = np.zeros(K)
    dp = ctypes.POINTER(ctypes.c_double)
    if _host_library().cafe5_get_gamma(K, float(alpha), freq.ctypes.data_as(dp), rate.ctypes.data_as(dp)) != 0:
        raise ValueError(f"get_gamma failed for K={K} alpha={alpha}")
    return freq, rate


class ResidentGammaModel(_ModelBase):
    """gamma_model (gamma_core.cpp:168-237) for an optimizer's sequence of
    (lambda, alpha) trial points (gamma_lambda_optimizer,
    optimizer_scorer.cpp:147-180), with everything kept on the device.

    The first call captures one plan for K categories: key k * n_t + i is
    (lambda m_k, t_i) over the tree's distinct branch lengths t_i, and every
    branch's matrix slot is fixed.  A trial point only recomputes the K
    multipliers (get_gamma at alpha), quantizes the K n_t keys as
    matrix_cache_key does, sets them and replays the plan (matrix build,
    pruning, recompute, reductions)."""

    def __init__(self, engine, tree, counts, max_family_size, max_root_family_size, K: int, prior,
                 error_model=None, path=native.PATH_BATCHED32, build_algo=native.BUILD_RECURRENCE,
                 use_graph: bool = True):
        super().__init__(engine, tree, counts, max_family_size, max_root_family_size, error_model, path,
                         build_algo)
        self.K = K
        lengths = [float(tree.length[i]) for i in range(tree.n_nodes) if i != tree.root]
        self.t_keys = sorted({native.key_branch_length(t) for t in lengths})
        t_index = {t: i for i, t in enumerate(self.t_keys)}
        n_t = len(self.t_keys)
        slots = np.zeros((K, tree.n_nodes), dtype=np.int32)
        for k in range(K):
            for u in range(tree.n_nodes):
                if u != tree.root:
                    slots[k, u] = k * n_t + t_index[native.key_branch_length(float(tree.length[u]))]
        self.node_slots = slots
        self.longest = max(lengths)
        self.cat_probs = np.full(K, 1.0 / K)
        freq, rate = get_gamma(K, 1.0)
        lam_q, t_q = self._keys(0.001, rate)
        if engine.N != self.N or getattr(engine, "_capacity", 0) < len(lam_q):
            engine.matrix_reset(self.N, len(lam_q))
            engine._capacity = len(lam_q)
        self.plan = engine.plan(slots, self._prior_table(prior), lam_q, t_q, path=path, model=native.MODEL_GAMMA,
                                cat_probs=self.cat_probs, algo=build_algo, use_graph=use_graph)
        self.results = None

    def _keys(self, lam: float, rate: np.ndarray):
        n_t = len(self.t_keys)
        lam_q = np.empty(self.K * n_t)
        t_q = np.empty(self.K * n_t)
        for k in range(self.K):
            lam_q[k * n_t:(k + 1) * n_t] = native.key_lambda(lam * float(rate[k]))
            t_q[k * n_t:(k + 1) * n_t] = self.t_keys
        return lam_q, t_q

    def can_infer(self, lam: float, alpha: float) -> bool:
        """gamma_core.cpp:123-141: a valid lambda, alpha >= 0, and the
        un-multiplied lambda does not saturate the longest branch."""
        return lam > 0 and alpha >= 0 and not is_saturated(self.longest, lam)

    def infer_family_likelihoods(self, lam: float, alpha: float, records: bool = True) -> float:
        """-lnL at (lambda, alpha); +inf where the reference returns it.
        Per-family records in self.results (cat, post, lnl) unless
        records=False (an optimizer's score needs only the total)."""
        if not self.can_infer(lam, alpha):
            self.results = None
            return math.inf
        freq, rate = get_gamma(self.K, alpha)
        self.multipliers = rate
        self.plan.set_keys(*self._keys(lam, rate))
        self.plan.execute()
        res = self.plan.results(records)
        self.last = res
        self.results = {"cat": res["cat"][0], "post": res["post"][0], "lnl": res["lnl"][0]} if records else None
        return float(res["total"][0])

    def close(self):
        self.plan.close()

