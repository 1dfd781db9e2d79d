"""ctypes binding of libcafe_gpu.so (include/cafe_gpu.h).

The library is the product's compute path; this module only marshals numpy
arrays.  Loading fails loudly when the library has not been built -- there is
no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .workloads import Tree

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libcafe_gpu.so")

BUILD_CLOSED_FORM = 0
BUILD_RECURRENCE = 1
PATH_PLAIN64 = 0
PATH_BATCHED64 = 1
PATH_BATCHED32 = 2
PATH_BATCHED16 = 3
MODEL_BASE = 0
MODEL_GAMMA = 1
# the paths whose contractions run below FP64 (per-family accuracy bar, FP64
# recompute of uncertified families)
REDUCED_PATHS = (PATH_BATCHED32, PATH_BATCHED16)
PATH_NAMES = {PATH_PLAIN64: "plain64", PATH_BATCHED64: "batched64", PATH_BATCHED32: "batched32",
              PATH_BATCHED16: "batched16"}

_i = ctypes.c_int
_d = ctypes.c_double
_p = ctypes.c_void_p
_pi = ctypes.POINTER(ctypes.c_int)
_pd = ctypes.POINTER(ctypes.c_double)
_pll = ctypes.POINTER(ctypes.c_longlong)


class EvalDesc(ctypes.Structure):
    _fields_ = [("model", _i), ("path", _i), ("n_trials", _i), ("n_cat", _i), ("node_slot", _pi),
                ("cat_probs", _pd), ("prior", _pd), ("prior_stride", _i)]


class CafeGpuError(RuntimeError):
    pass


_LIB: Optional[ctypes.CDLL] = None


def load_library() -> ctypes.CDLL:
    """Load libcafe_gpu.so; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise CafeGpuError(f"{LIB_PATH} is missing: build it with `python -m cafe5_amd.build`")
    lib = ctypes.CDLL(LIB_PATH)
    sig = {
        "cafe_gpu_create": (_i, [_i, ctypes.POINTER(_p)]),
        "cafe_gpu_destroy": (_i, [_p]),
        "cafe_gpu_last_error": (ctypes.c_char_p, []),
        "cafe_gpu_matrix_reset": (_i, [_p, _i, _i]),
        "cafe_gpu_matrix_build": (_i, [_p, _i, _pd, _pd, _i, _i]),
        "cafe_gpu_matrix_get": (_i, [_p, _i, _pd]),
        "cafe_gpu_set_tree": (_i, [_p, _i, _pi, _pi]),
        "cafe_gpu_set_families": (_i, [_p, _i, _i, _pi, _pi, _pi, _i, _i]),
        "cafe_gpu_bucket_count": (_i, [_p, _pi]),
        "cafe_gpu_bucket_info": (_i, [_p, _i, _pi, _pi, _pi]),
        "cafe_gpu_set_error_model": (_i, [_p, _i, _i, _pd]),
        "cafe_gpu_eval": (_i, [_p, ctypes.POINTER(EvalDesc), _pd, _pd, _pd, _pd, _pll]),
        "cafe_gpu_root_vectors": (_i, [_p, ctypes.POINTER(EvalDesc), _i, _i, _i, _i, _pd]),
        "cafe_gpu_reconstruct": (_i, [_p, _pi, _pd, _i, _i, _i, _i, _pi]),
        "cafe_gpu_reconstruct_noalloc": (_i, [_p, _pi, _pd, _i, _i, _i, _i, _pi]),
        "cafe_gpu_last_timings": (_i, [_p, _pd]),
        "cafe_gpu_tuning_load": (_i, [_p, ctypes.c_char_p, _pi]),
        "cafe_gpu_tuning_set": (_i, [_p, _i, _i, _i, _i, _i, _i, _i, _i]),
        "cafe_gpu_tuning_clear": (_i, [_p]),
        "cafe_gpu_plan_create": (_i, [_p, ctypes.POINTER(EvalDesc), _i, _pd, _pd, _i, _i, ctypes.POINTER(_p)]),
        "cafe_gpu_plan_set_keys": (_i, [_p, _pd, _pd]),
        "cafe_gpu_plan_execute": (_i, [_p]),
        "cafe_gpu_plan_synchronize": (_i, [_p]),
        "cafe_gpu_plan_results": (_i, [_p, _pd, _pd, _pd, _pd, _pll]),
        "cafe_gpu_plan_profile": (_i, [_p, _pd, _pi]),
        "cafe_gpu_plan_recompute_counts": (_i, [_p, _pll, _pll]),
        "cafe_gpu_plan_class_times": (_i, [_p, _i, _pi, _pi, _pll, _pd, _pi]),
        "cafe_gpu_plan_destroy": (_i, [_p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


EXPORTED_SYMBOLS = ["cafe_gpu_create", "cafe_gpu_destroy", "cafe_gpu_last_error", "cafe_gpu_matrix_reset",
                    "cafe_gpu_matrix_build", "cafe_gpu_matrix_get", "cafe_gpu_set_tree",
                    "cafe_gpu_set_families", "cafe_gpu_bucket_count", "cafe_gpu_bucket_info",
                    "cafe_gpu_set_error_model", "cafe_gpu_eval", "cafe_gpu_root_vectors", "cafe_gpu_reconstruct",
                    "cafe_gpu_reconstruct_noalloc", "cafe_gpu_last_timings", "cafe_gpu_plan_create", "cafe_gpu_plan_set_keys",
                    "cafe_gpu_plan_execute", "cafe_gpu_plan_synchronize", "cafe_gpu_plan_results",
                    "cafe_gpu_plan_profile", "cafe_gpu_plan_recompute_counts", "cafe_gpu_plan_class_times",
                    "cafe_gpu_plan_destroy", "cafe_gpu_tuning_load", "cafe_gpu_tuning_set",
                    "cafe_gpu_tuning_clear"]


def _ip(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(_pi)


def _dp(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(_pd)


def key_lambda(lam: float) -> float:
    """matrix_cache_key lambda quantization (matrix_cache.h:50, :57-58)."""
    return float(int(lam * 1000000000)) / 1000000000.0


def key_branch_length(t: float) -> float:
    """matrix_cache_key branch-length quantization (matrix_cache.h:51, :60-61)."""
    return float(int(t * 1000)) / 1000.0


class Engine:
    """One libcafe_gpu context on one device."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = _p()
        self._check(self.lib.cafe_gpu_create(device, ctypes.byref(h)))
        self.ctx = h
        self.N = 0
        self.F = 0

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.cafe_gpu_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int):
        if rc != 0:
            raise CafeGpuError(f"libcafe_gpu error {rc}: {self.lib.cafe_gpu_last_error().decode()}")

    # ---- matrices ----
    def matrix_reset(self, N: int, capacity: int):
        self._check(self.lib.cafe_gpu_matrix_reset(self.ctx, N, capacity))
        self.N = N

    def matrix_build(self, lam_q: Sequence[float], t_q: Sequence[float], first_slot: int = 0,
                     algo: int = BUILD_RECURRENCE):
        lq = np.ascontiguousarray(lam_q, dtype=np.float64)
        tq = np.ascontiguousarray(t_q, dtype=np.float64)
        self._check(self.lib.cafe_gpu_matrix_build(self.ctx, len(lq), _dp(lq), _dp(tq), first_slot, algo))

    def matrix_get(self, slot: int) -> np.ndarray:
        out = np.zeros((self.N, self.N))
        self._check(self.lib.cafe_gpu_matrix_get(self.ctx, slot, _dp(out)))
        return out

    # ---- inputs ----
    def set_tree(self, tree: Tree):
        par = np.ascontiguousarray(tree.parent, dtype=np.int32)
        col = np.ascontiguousarray(tree.leaf_col, dtype=np.int32)
        self._check(self.lib.cafe_gpu_set_tree(self.ctx, tree.n_nodes, _ip(par), _ip(col)))
        self.tree = tree

    def set_families(self, counts: np.ndarray, B: int = 0, R: int = 0, bound_B=None, bound_R=None):
        c = np.ascontiguousarray(counts, dtype=np.int32)
        bb = None if bound_B is None else np.ascontiguousarray(bound_B, dtype=np.int32)
        br = None if bound_R is None else np.ascontiguousarray(bound_R, dtype=np.int32)
        self._check(self.lib.cafe_gpu_set_families(self.ctx, c.shape[0], c.shape[1], _ip(c), _ip(bb), _ip(br),
                                                   B, R))
        self.F = c.shape[0]

    def buckets(self) -> List[Tuple[int, int, int]]:
        n = ctypes.c_int()
        self._check(self.lib.cafe_gpu_bucket_count(self.ctx, ctypes.byref(n)))
        out = []
        for i in range(n.value):
            b, r, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
            self._check(self.lib.cafe_gpu_bucket_info(self.ctx, i, ctypes.byref(b), ctypes.byref(r),
                                                      ctypes.byref(c)))
            out.append((b.value, r.value, c.value))
        return out

    def set_error_model(self, rows: Optional[np.ndarray]):
        if rows is None:
            self._check(self.lib.cafe_gpu_set_error_model(self.ctx, 0, 0, None))
            return
        r = np.ascontiguousarray(rows, dtype=np.float64)
        self._check(self.lib.cafe_gpu_set_error_model(self.ctx, r.shape[1], r.shape[0], _dp(r)))

    # ---- evaluation ----
    def _desc(self, model, path, node_slot, prior, cat_probs):
        ns = np.ascontiguousarray(node_slot, dtype=np.int32)
        pr = np.ascontiguousarray(np.atleast_2d(prior), dtype=np.float64)
        cp = None if cat_probs is None else np.ascontiguousarray(cat_probs, dtype=np.float64)
        K = 1 if cp is None else len(cp)
        T = ns.shape[0] // K
        d = EvalDesc(model, path, T, K, _ip(ns), _dp(cp), _dp(pr), pr.shape[1])
        return d, (ns, pr, cp), T, K

    def evaluate(self, node_slot: np.ndarray, prior: np.ndarray, path: int = PATH_BATCHED32,
                 model: int = MODEL_BASE, cat_probs=None) -> Dict[str, np.ndarray]:
        """node_slot: (T*K, n_nodes) matrix slots; prior: (n_buckets, stride)."""
        d, keep, T, K = self._desc(model, path, node_slot, prior, cat_probs)
        lnl = np.zeros((T, self.F))
        cat = np.zeros((T, self.F, K)) if model == MODEL_GAMMA else None
        post = np.zeros((T, self.F, K)) if model == MODEL_GAMMA else None
        total = np.zeros(T)
        nfb = ctypes.c_longlong(0)
        self._check(self.lib.cafe_gpu_eval(self.ctx, ctypes.byref(d), _dp(lnl), _dp(cat), _dp(post), _dp(total),
                                           ctypes.byref(nfb)))
        return {"lnl": lnl, "cat": cat, "post": post, "total": total, "n_fallback": nfb.value}

    def root_vectors(self, node_slot: np.ndarray, prior: np.ndarray, path: int, trial: int = 0, cat: int = 0,
                     model: int = MODEL_BASE, cat_probs=None, R: Optional[int] = None) -> np.ndarray:
        d, keep, T, K = self._desc(model, path, node_slot, prior, cat_probs)
        R = R or int(np.atleast_2d(prior).shape[1])
        out = np.zeros(self.F * R)
        self._check(self.lib.cafe_gpu_root_vectors(self.ctx, ctypes.byref(d), trial, cat, 0, self.F, _dp(out)))
        return out.reshape(self.F, R)

    def reconstruct(self, node_slot: np.ndarray, prior: np.ndarray, B: int, R: int, f0: int = 0,
                    nf: Optional[int] = None, noalloc: bool = False) -> np.ndarray:
        """Pupko joint reconstruction of families f0 .. f0+nf-1 (all by
        default): states[nf, n_nodes] (gene_family_reconstructor.cpp:30-189);
        prior[j], j = 0..min(B, R).  noalloc: through
        cafe_gpu_reconstruct_noalloc (fixed device scratch, no allocation)."""
        n_nodes = self.tree.n_nodes
        nf = self.F - f0 if nf is None else nf
        ns = np.ascontiguousarray(np.asarray(node_slot).reshape(-1)[:n_nodes], dtype=np.int32)
        pr = np.ascontiguousarray(prior, dtype=np.float64)
        if pr.size < min(B, R) + 1:
            raise ValueError("prior needs min(B, R) + 1 entries")
        out = np.zeros((max(nf, 0), n_nodes), dtype=np.int32)
        fn = self.lib.cafe_gpu_reconstruct_noalloc if noalloc else self.lib.cafe_gpu_reconstruct
        self._check(fn(self.ctx, _ip(ns), _dp(pr), B, R, f0, nf, _ip(out)))
        return out

    # ---- launch shapes (see cafe_gpu_tuning_* in include/cafe_gpu.h) ----
    def tuning_load(self, path: str) -> int:
        """Replace the context's shape table with a cache file's entries."""
        n = ctypes.c_int(0)
        self._check(self.lib.cafe_gpu_tuning_load(self.ctx, path.encode(), ctypes.byref(n)))
        return n.value

    def tuning_set(self, path: int, tiles: int, leaves: int, kernel: int, ntw: int = 0, w: int = 0,
                   lds: bool = False, grid_mult: int = 0):
        self._check(self.lib.cafe_gpu_tuning_set(self.ctx, path, tiles, leaves, kernel, ntw, w, 1 if lds else 0,
                                                 grid_mult))

    def tuning_clear(self):
        self._check(self.lib.cafe_gpu_tuning_clear(self.ctx))

    def timings(self) -> np.ndarray:
        out = np.zeros(4)
        self._check(self.lib.cafe_gpu_last_timings(self.ctx, _dp(out)))
        return out

    def plan(self, node_slot: np.ndarray, prior: np.ndarray, lam_q, t_q, path: int = PATH_BATCHED32,
             model: int = MODEL_BASE, cat_probs=None, algo: int = BUILD_RECURRENCE, use_graph: bool = True):
        return Plan(self, node_slot, prior, lam_q, t_q, path, model, cat_probs, algo, use_graph)


class Plan:
    """A device-resident evaluation replayed as one HIP graph per execute()."""

    def __init__(self, eng: Engine, node_slot, prior, lam_q, t_q, path, model, cat_probs, algo, use_graph):
        self.eng = eng
        self.lib = eng.lib
        d, self._keep, self.T, self.K = eng._desc(model, path, node_slot, prior, cat_probs)
        self.model = model
        lq = np.ascontiguousarray(lam_q, dtype=np.float64)
        tq = np.ascontiguousarray(t_q, dtype=np.float64)
        self.n_keys = len(lq)
        h = _p()
        eng._check(self.lib.cafe_gpu_plan_create(eng.ctx, ctypes.byref(d), self.n_keys, _dp(lq), _dp(tq), algo,
                                                 1 if use_graph else 0, ctypes.byref(h)))
        self.h = h

    def set_keys(self, lam_q, t_q):
        lq = np.ascontiguousarray(lam_q, dtype=np.float64)
        tq = np.ascontiguousarray(t_q, dtype=np.float64)
        self.eng._check(self.lib.cafe_gpu_plan_set_keys(self.h, _dp(lq), _dp(tq)))

    def execute(self):
        self.eng._check(self.lib.cafe_gpu_plan_execute(self.h))

    def synchronize(self):
        self.eng._check(self.lib.cafe_gpu_plan_synchronize(self.h))

    def results(self, records: bool = True) -> Dict[str, np.ndarray]:
        """The last execution's outputs; records=False copies only the
        totals and the recompute count (what an optimizer's score needs)."""
        F = self.eng.F
        lnl = np.zeros((self.T, F)) if records else None
        cat = np.zeros((self.T, F, self.K)) if self.model == MODEL_GAMMA and records else None
        post = np.zeros((self.T, F, self.K)) if self.model == MODEL_GAMMA and records else None
        total = np.zeros(self.T)
        nfb = ctypes.c_longlong(0)
        self.eng._check(self.lib.cafe_gpu_plan_results(self.h, _dp(lnl), _dp(cat), _dp(post), _dp(total),
                                                       ctypes.byref(nfb)))
        return {"lnl": lnl, "cat": cat, "post": post, "total": total, "n_fallback": nfb.value}

    def profile(self):
        """One execution outside the graph with events between phases:
        ([build, main prune, recompute, finalize] ms, graph_ok)."""
        out = np.zeros(4)
        g = ctypes.c_int(0)
        self.eng._check(self.lib.cafe_gpu_plan_profile(self.h, _dp(out), ctypes.byref(g)))
        return out, bool(g.value)

    def class_times(self):
        """Per size class of the main pass, from the last profile() run:
        (16-row tiles per vector, tiles, kernel ms, launch) arrays; launch
        rows are (kernel: 0 matrix-core, 1 one family per lane, 2 PLAIN64,
        3 split-FP16; tiles per wave or rows per lane; waves per workgroup;
        partials in LDS; workgroups; families per tile: 16 or 32 for the
        matrix-core kernel, 64 per lane)."""
        n = ctypes.c_int(0)
        self.eng._check(self.lib.cafe_gpu_plan_class_times(self.h, 0, ctypes.byref(n), None, None, None, None))
        nt = np.zeros(n.value, dtype=np.int32)
        tiles = np.zeros(n.value, dtype=np.int64)
        ms = np.zeros(n.value)
        launch = np.zeros((n.value, 6), dtype=np.int32)
        self.eng._check(self.lib.cafe_gpu_plan_class_times(self.h, n.value, ctypes.byref(n), _ip(nt),
                                                           tiles.ctypes.data_as(_pll), _dp(ms), _ip(launch)))
        return nt, tiles, ms, launch

    def recompute_counts(self):
        """(families redone on BATCHED64, families redone on PLAIN64)."""
        a, b = ctypes.c_longlong(0), ctypes.c_longlong(0)
        self.eng._check(self.lib.cafe_gpu_plan_recompute_counts(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.cafe_gpu_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class KeyTable:
    """Matrix keys of a likelihood evaluation, as matrix_cache sees them.

    Every (lambda, branch length) pair is quantized like matrix_cache_key
    (truncation to 1e-9 / 1e-3) and de-duplicated; node_slot maps each branch
    of each parameter point to its matrix slot.
    """

    def __init__(self):
        self.keys: Dict[Tuple[float, float], int] = {}

    def slot(self, lam: float, t: float) -> int:
        k = (key_lambda(lam), key_branch_length(t))
        if k not in self.keys:
            self.keys[k] = len(self.keys)
        return self.keys[k]

    def node_slots(self, tree: Tree, node_lambdas: np.ndarray) -> np.ndarray:
        """node_lambdas: (P, n_nodes) lambda of every branch at every point."""
        P = node_lambdas.shape[0]
        out = np.zeros((P, tree.n_nodes), dtype=np.int32)
        for p in range(P):
            for i in range(tree.n_nodes):
                if i != tree.root:
                    out[p, i] = self.slot(float(node_lambdas[p, i]), float(tree.length[i]))
        return out

    def arrays(self) -> Tuple[np.ndarray, np.ndarray]:
        lam = np.zeros(len(self.keys))
        t = np.zeros(len(self.keys))
        for (lq, tq), s in self.keys.items():
            lam[s] = lq
            t[s] = tq
        return lam, t
