"""ctypes access to the checker libraries (tests only).

  oracle/_build/liboracle.so   FP64 C restatement of the reference (the oracle)
  oracle/_ref/libcaferef.so    the reference's own sources compiled by
                               oracle/Makefile (present only where the CAFE5
                               tree was available at build time)

Both are test infrastructure; the product never imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Optional, Sequence

import numpy as np

from cafe5_amd.workloads import Tree

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libcaferef.so")

_d = ctypes.c_double
_i = ctypes.c_int
_pd = ctypes.POINTER(ctypes.c_double)
_pi = ctypes.POINTER(ctypes.c_int)
_cp = ctypes.c_char_p


def _dp(a: Optional[np.ndarray]):
    if a is None:
        return None
    return a.ctypes.data_as(_pd)


def _ip(a: Optional[np.ndarray]):
    if a is None:
        return None
    return a.ctypes.data_as(_pi)


class OrcTree(ctypes.Structure):
    _fields_ = [("n_nodes", _i), ("parent", _pi), ("leaf_col", _pi)]


def _load_oracle() -> ctypes.CDLL:
    if not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                       stdout=subprocess.DEVNULL)
    lib = ctypes.CDLL(ORACLE_SO)
    sig = {
        "orc_init": (None, []),
        "orc_chooseln": (_d, [_d, _d]),
        "orc_birthdeath": (_d, [_i, _i, _d, _d]),
        "orc_transition": (_d, [_d, _d, _i, _i]),
        "orc_is_saturated": (_i, [_d, _d]),
        "orc_key_lambda": (_d, [_d]),
        "orc_key_branch_length": (_d, [_d]),
        "orc_build_matrix": (None, [_d, _d, _i, _pd]),
        "orc_matvec": (None, [_pd, _i, _pd, _i, _i, _i, _i, _pd]),
        "orc_leaf_vector": (_i, [_i, _pd, _i, _i, _pd]),
        "orc_prune": (_i, [ctypes.POINTER(OrcTree), _pi, _pd, _pi, _i, _pd, _i, _i, _i, _pd]),
        "orc_base_reduce": (_d, [_i, _i, _pd, _pd, _pd]),
        "orc_gamma_reduce": (_d, [_i, _i, _i, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pi]),
        "orc_evaluate": (_d, [ctypes.POINTER(OrcTree), _pd, _pd, _i, _i, _pi, _i, _i, _pd, _pd, _i,
                              _i, _pd, _pd, _pd, _pd, _pd, _pd, _i]),
        "orc_get_gamma": (None, [_i, _d, _pd, _pd]),
        "orc_prior_uniform": (None, [_i, _i, _pd]),
        "orc_prior_poisson": (None, [_d, _i, _i, _pd]),
        "orc_prior_map": (None, [_i, _pi, _pi, _i, _pd]),
        "orc_family_probability": (_i, [ctypes.POINTER(OrcTree), _pi, _pd, _pi, _i, _i, _i, _i, _pd]),
        "orc_pvalue": (_d, [_d, _pd, _i]),
        "orc_find_best_pvalue": (_d, [_pd, _i, _i, _pd, _i, _i]),
        "orc_reconstruct": (_i, [ctypes.POINTER(OrcTree), _pi, _pd, _pi, _i, _i, _i, _pd, _pi, _pd]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    lib.orc_init()
    return lib


def _load_ref() -> Optional[ctypes.CDLL]:
    if not os.path.exists(REF_SO):
        return None
    lib = ctypes.CDLL(REF_SO)
    sig = {
        "ref_init": (None, []),
        "ref_chooseln": (_d, [_d, _d]),
        "ref_birthdeath": (_d, [_i, _i, _d, _d]),
        "ref_transition": (_d, [_d, _d, _i, _i]),
        "ref_build_matrix": (_i, [_d, _d, _i, _pd]),
        "ref_prune": (_i, [_cp, _cp, _pi, _d, _d, _i, _i, _i, _i, _pd, _pd]),
        "ref_prior": (None, [_i, _i, _d, _i, _pi, _pi, _i, _pd]),
        "ref_base_model": (_d, [_cp, _cp, _i, _pi, _i, _pd, _cp, _i, _i, _i, _i, _d, _i, _i, _pd, _pd]),
        "ref_gamma_model": (_d, [_cp, _cp, _i, _pi, _i, _pd, _cp, _i, _i, _i, _pd, _pd, _i, _i, _d,
                                 _i, _i, _pd, _pd, _pd, _pd]),
        "ref_get_gamma": (None, [_i, _d, _pd, _pd]),
        "ref_family_probabilities": (_i, [_cp, _cp, _i, _pi, _pi, _d, _i, _i, _i, _pd]),
        "ref_pvalue": (_d, [_d, _pd, _i]),
        "ref_find_best_pvalue": (_d, [_i, _pd, _i, _pd, _i, _i]),
        "ref_create_families": (_i, [_cp, _cp, _d, _i, _i, _i, ctypes.c_uint, _i, _i, _pi]),
        "ref_reconstruct": (_i, [_cp, _cp, _cp, _pi, _d, _i, _i, _i, _i, _d, _i, _pi, _pi, _pi]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    lib.ref_init()
    return lib


ORC = _load_oracle()
REF = _load_ref()


class Oracle:
    """Numpy-level helpers over liboracle.so."""

    lib = ORC

    @staticmethod
    def tree_struct(tree: Tree):
        par = np.ascontiguousarray(tree.parent, dtype=np.int32)
        col = np.ascontiguousarray(tree.leaf_col, dtype=np.int32)
        st = OrcTree(tree.n_nodes, _ip(par), _ip(col))
        return st, (par, col)

    @staticmethod
    def matrix(lam_q: float, t_q: float, N: int) -> np.ndarray:
        out = np.zeros((N, N))
        ORC.orc_build_matrix(lam_q, t_q, N, _dp(out))
        return out

    @staticmethod
    def key(lam: float, t: float):
        return ORC.orc_key_lambda(lam), ORC.orc_key_branch_length(t)

    @staticmethod
    def node_matrices(tree: Tree, node_lambda: Sequence[float], N: int):
        """Pool + node_mat for the branches of tree (quantized keys, memoized)."""
        keys = {}
        mats = []
        node_mat = np.zeros(tree.n_nodes, dtype=np.int32)
        for i in range(tree.n_nodes):
            if i == tree.root:
                continue
            k = Oracle.key(float(node_lambda[i]), float(tree.length[i]))
            if k not in keys:
                keys[k] = len(mats)
                mats.append(Oracle.matrix(k[0], k[1], N))
            node_mat[i] = keys[k]
        pool = np.ascontiguousarray(np.stack(mats)) if mats else np.zeros((1, N, N))
        return pool, node_mat

    @staticmethod
    def prune(tree: Tree, counts_row: np.ndarray, node_lambda, B: int, R: int,
              err: Optional[np.ndarray] = None, N: Optional[int] = None) -> np.ndarray:
        N = N or max(R, B) + 1
        pool, node_mat = Oracle.node_matrices(tree, node_lambda, N)
        st, keep = Oracle.tree_struct(tree)
        cnt = np.ascontiguousarray(counts_row, dtype=np.int32)
        errc = None if err is None else np.ascontiguousarray(err, dtype=np.float64)
        out = np.zeros(R)
        rc = ORC.orc_prune(ctypes.byref(st), _ip(cnt), _dp(pool), _ip(node_mat), N, _dp(errc),
                           0 if err is None else errc.shape[1], B, R, _dp(out))
        if rc != 0:
            raise ValueError("orc_prune rejected its input")
        return out

    @staticmethod
    def evaluate(tree: Tree, node_lambda, counts: np.ndarray, B: int, R: int, prior: np.ndarray,
                 err: Optional[np.ndarray] = None, n_cat: int = 0, multipliers=None, cat_probs=None,
                 threads: int = 1):
        st, keep = Oracle.tree_struct(tree)
        F, n = counts.shape
        cnt = np.ascontiguousarray(counts, dtype=np.int32)
        nl = np.ascontiguousarray(node_lambda, dtype=np.float64)
        nt = np.ascontiguousarray(tree.length, dtype=np.float64)
        pr = np.ascontiguousarray(prior, dtype=np.float64)
        errc = None if err is None else np.ascontiguousarray(err, dtype=np.float64)
        K = max(n_cat, 1)
        lnl = np.zeros(F)
        fam = np.zeros(F)
        cat = np.zeros(F * K)
        post = np.zeros(F * K)
        mult = None if multipliers is None else np.ascontiguousarray(multipliers, dtype=np.float64)
        cp = None if cat_probs is None else np.ascontiguousarray(cat_probs, dtype=np.float64)
        total = ORC.orc_evaluate(ctypes.byref(st), _dp(nl), _dp(nt), F, n, _ip(cnt), B, R, _dp(pr),
                                 _dp(errc), 0 if err is None else errc.shape[1], n_cat, _dp(mult),
                                 _dp(cp), _dp(lnl), _dp(cat), _dp(fam), _dp(post), threads)
        return {"total": total, "lnl": lnl, "fam": fam, "cat": cat.reshape(F, K),
                "post": post.reshape(F, K)}

    @staticmethod
    def prior_uniform(max_size: int, R: int) -> np.ndarray:
        out = np.zeros(R)
        ORC.orc_prior_uniform(max_size, R, _dp(out))
        return out

    @staticmethod
    def prior_poisson(lam: float, num_values: int, R: int) -> np.ndarray:
        out = np.zeros(R)
        ORC.orc_prior_poisson(lam, num_values, R, _dp(out))
        return out

    @staticmethod
    def prior_map(sizes, counts, R: int) -> np.ndarray:
        s = np.ascontiguousarray(sizes, dtype=np.int32)
        c = np.ascontiguousarray(counts, dtype=np.int32)
        out = np.zeros(R)
        ORC.orc_prior_map(len(s), _ip(s), _ip(c), R, _dp(out))
        return out

    @staticmethod
    def get_gamma(K: int, alpha: float):
        f = np.zeros(K)
        r = np.zeros(K)
        ORC.orc_get_gamma(K, alpha, _dp(f), _dp(r))
        return f, r

    @staticmethod
    def family_probability(tree: Tree, counts_row, lam: float, B: int, max_size: int, r: int,
                           N: Optional[int] = None) -> float:
        """compute_family_probabilities for one family (single lambda)."""
        N = N or max(B, r) + 1
        pool, node_mat = Oracle.node_matrices(tree, [lam] * tree.n_nodes, N)
        st, keep = Oracle.tree_struct(tree)
        cnt = np.ascontiguousarray(counts_row, dtype=np.int32)
        out = np.zeros(1)
        rc = ORC.orc_family_probability(ctypes.byref(st), _ip(cnt), _dp(pool), _ip(node_mat), N, B,
                                        int(max_size), int(r), _dp(out))
        if rc != 0:
            raise ValueError("orc_family_probability rejected its input")
        return float(out[0])

    @staticmethod
    def reconstruct(tree: Tree, counts_row, lam, B: int, R: int, prior, N: Optional[int] = None,
                    with_L: bool = False, matrices=None):
        """Pupko joint reconstruction of one family: states per node (and the
        nodes' L vectors).  lam: one value or one per node.  matrices: the
        (pool, node_mat) of node_matrices at these lambdas and N, to share
        one build between families."""
        N = N or max(B, R) + 1
        node_lambda = np.full(tree.n_nodes, lam) if np.ndim(lam) == 0 else np.asarray(lam)
        pool, node_mat = matrices if matrices is not None else Oracle.node_matrices(tree, node_lambda, N)
        st, keep = Oracle.tree_struct(tree)
        cnt = np.ascontiguousarray(counts_row, dtype=np.int32)
        pr = np.ascontiguousarray(prior, dtype=np.float64)
        states = np.zeros(tree.n_nodes, dtype=np.int32)
        L = np.zeros((tree.n_nodes, B + 1))
        rc = ORC.orc_reconstruct(ctypes.byref(st), _ip(cnt), _dp(pool), _ip(node_mat), N, B, R, _dp(pr),
                                 _ip(states), _dp(L))
        if rc != 0:
            raise ValueError(f"orc_reconstruct failed ({rc})")
        return (states, L) if with_L else states

    @staticmethod
    def pvalue(v: float, conddist) -> float:
        cd = np.ascontiguousarray(conddist, dtype=np.float64)
        return ORC.orc_pvalue(float(v), _dp(cd), len(cd))

    @staticmethod
    def find_best_pvalue(root_probs, max_size: int, conddist: np.ndarray) -> float:
        rp = np.ascontiguousarray(root_probs, dtype=np.float64)
        cd = np.ascontiguousarray(conddist, dtype=np.float64)
        return ORC.orc_find_best_pvalue(_dp(rp), len(rp), int(max_size), _dp(cd), cd.shape[0], cd.shape[1])


class Ref:
    """Numpy-level helpers over the compiled reference (libcaferef.so)."""

    lib = REF

    @staticmethod
    def available() -> bool:
        return REF is not None

    @staticmethod
    def matrix(lam: float, t: float, N: int) -> np.ndarray:
        out = np.zeros((N, N))
        if REF.ref_build_matrix(lam, t, N, _dp(out)) != 0:
            raise RuntimeError("ref_build_matrix failed")
        return out

    @staticmethod
    def _err_args(err: Optional[np.ndarray]):
        if err is None:
            return 0, 0, None
        e = np.ascontiguousarray(err, dtype=np.float64)
        return e.shape[1], e.shape[0], e

    @staticmethod
    def prune(tree: Tree, counts_row, lam: float, multiplier: float, B: int, R: int,
              err: Optional[np.ndarray] = None) -> np.ndarray:
        nd, nr, e = Ref._err_args(err)
        cnt = np.ascontiguousarray(counts_row, dtype=np.int32)
        out = np.zeros(R)
        rc = REF.ref_prune(tree.newick().encode(), " ".join(tree.leaf_names()).encode(), _ip(cnt),
                           lam, multiplier, B, R, nd, nr, _dp(e), _dp(out))
        if rc != 0:
            raise RuntimeError("ref_prune failed")
        return out

    @staticmethod
    def prior(kind: int, max_size: int, poisson_lambda: float, R: int, sizes=None, counts=None):
        out = np.zeros(R)
        s = None if sizes is None else np.ascontiguousarray(sizes, dtype=np.int32)
        c = None if counts is None else np.ascontiguousarray(counts, dtype=np.int32)
        REF.ref_prior(kind, max_size, poisson_lambda, 0 if s is None else len(s), _ip(s), _ip(c), R,
                      _dp(out))
        return out

    @staticmethod
    def _lambda_args(tree: Tree, lambdas, node_class):
        lam = np.ascontiguousarray(np.atleast_1d(lambdas), dtype=np.float64)
        if node_class is None:
            return lam, b""
        labels = [int(c) + 1 for c in node_class]
        return lam, tree.newick(labels=labels).encode()

    @staticmethod
    def base_model(tree: Tree, counts: np.ndarray, lambdas, B: int, R: int, prior_kind: int = 0,
                   prior_max: int = 100, poisson_lambda: float = 0.0, err=None, node_class=None):
        F = counts.shape[0]
        lam, lt = Ref._lambda_args(tree, lambdas, node_class)
        nd, nr, e = Ref._err_args(err)
        cnt = np.ascontiguousarray(counts, dtype=np.int32)
        out = np.zeros(F)
        total = REF.ref_base_model(tree.newick().encode(), " ".join(tree.leaf_names()).encode(), F,
                                   _ip(cnt), len(lam), _dp(lam), lt, B, R, prior_kind, prior_max,
                                   poisson_lambda, nd, nr, _dp(e), _dp(out))
        return total, out

    @staticmethod
    def gamma_model(tree: Tree, counts: np.ndarray, lambdas, B: int, R: int, cat_probs, multipliers,
                    prior_kind: int = 0, prior_max: int = 100, poisson_lambda: float = 0.0, err=None,
                    node_class=None):
        F = counts.shape[0]
        K = len(cat_probs)
        lam, lt = Ref._lambda_args(tree, lambdas, node_class)
        nd, nr, e = Ref._err_args(err)
        cnt = np.ascontiguousarray(counts, dtype=np.int32)
        cp = np.ascontiguousarray(cat_probs, dtype=np.float64)
        mu = np.ascontiguousarray(multipliers, dtype=np.float64)
        cat = np.zeros(F * K)
        post = np.zeros(F * K)
        fam = np.zeros(F)
        total = REF.ref_gamma_model(tree.newick().encode(), " ".join(tree.leaf_names()).encode(), F,
                                    _ip(cnt), len(lam), _dp(lam), lt, B, R, K, _dp(cp), _dp(mu),
                                    prior_kind, prior_max, poisson_lambda, nd, nr, _dp(e), _dp(cat),
                                    _dp(fam), _dp(post))
        return total, cat.reshape(F, K), fam, post.reshape(F, K)

    @staticmethod
    def get_gamma(K: int, alpha: float):
        f = np.zeros(K)
        r = np.zeros(K)
        REF.ref_get_gamma(K, alpha, _dp(f), _dp(r))
        return f, r

    @staticmethod
    def family_probabilities(tree: Tree, counts: np.ndarray, inner, lam: float, B: int, R: int,
                             root_size: int) -> np.ndarray:
        """compute_family_probabilities over families whose interior nodes
        all hold inner[f] in the clademap."""
        F = counts.shape[0]
        cnt = np.ascontiguousarray(counts, dtype=np.int32)
        inn = np.ascontiguousarray(inner, dtype=np.int32)
        out = np.zeros(F)
        rc = REF.ref_family_probabilities(tree.newick().encode(), " ".join(tree.leaf_names()).encode(), F,
                                          _ip(cnt), _ip(inn), lam, B, R, root_size, _dp(out))
        if rc != 0:
            raise RuntimeError("ref_family_probabilities failed")
        return out

    @staticmethod
    def pvalue(v: float, conddist) -> float:
        cd = np.ascontiguousarray(conddist, dtype=np.float64)
        return REF.ref_pvalue(float(v), _dp(cd), len(cd))

    @staticmethod
    def find_best_pvalue(root_probs, max_size: int, conddist: np.ndarray) -> float:
        rp = np.ascontiguousarray(root_probs, dtype=np.float64)
        cd = np.ascontiguousarray(conddist, dtype=np.float64)
        return REF.ref_find_best_pvalue(int(max_size), _dp(rp), len(rp), _dp(cd), cd.shape[0], cd.shape[1])

    @staticmethod
    def create_families(tree: Tree, lam: float, B: int, R: int, N: int, seed: int, root_size: int,
                        n_draws: int) -> np.ndarray:
        """create_family n_draws times after seeding the reference's engine:
        leaf sizes [n_draws, n_leaves] in leaf_names order."""
        out = np.zeros((n_draws, tree.n_leaves), dtype=np.int32)
        rc = REF.ref_create_families(tree.newick().encode(), " ".join(tree.leaf_names()).encode(), lam, B, R,
                                     N, seed, root_size, n_draws, _ip(out))
        if rc != 0:
            raise RuntimeError("ref_create_families failed")
        return out

    @staticmethod
    def reconstruct(tree: Tree, counts_row, lam: float, B: int, R: int, prior_kind: int = 0, prior_max: int = 0,
                    poisson_lambda: float = 0.0, map_sizes=None, map_counts=None) -> np.ndarray:
        """reconstruct_gene_family of one family: the state of every node, in
        the tree's node order."""
        cnt = np.ascontiguousarray(counts_row, dtype=np.int32)
        ms = None if map_sizes is None else np.ascontiguousarray(map_sizes, dtype=np.int32)
        mc = None if map_counts is None else np.ascontiguousarray(map_counts, dtype=np.int32)
        out = np.zeros(tree.n_nodes, dtype=np.int32)
        rc = REF.ref_reconstruct(tree.newick().encode(), " ".join(tree.leaf_names()).encode(),
                                 " ".join(tree.names).encode(), _ip(cnt), lam, B, R, prior_kind, prior_max,
                                 poisson_lambda, 0 if ms is None else len(ms), _ip(ms), _ip(mc), _ip(out))
        if rc != 0:
            raise RuntimeError("ref_reconstruct failed")
        return out
