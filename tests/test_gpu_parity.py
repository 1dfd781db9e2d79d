"""GPU parity of libcafe_gpu.so against the FP64 oracle (through the C ABI).

Tolerances:
  * matrices: <= 2e-12 relative to the oracle for entries above 1e-280.  The
    closed form multiplies log(alpha) by s+c-2j before exp, so one ulp of
    difference between the device and host libm's log becomes ~(s+c)|log a|
    ulps (~1e-13 at N = 201); the recurrence is a different, cancellation-free
    summation of the same positive terms;
  * PLAIN64 / BATCHED64 per-family lnL: <= 1e-9 relative;
  * BATCHED32: per-family |dlnL| <= 5e-6 + 1e-7 |lnL| and <= 1e-5 relative
    (tests/bars.py), the total within 1e-6 relative, with families the FP32
    flush bound cannot certify recomputed in FP64.
"""
import math

import numpy as np
import pytest

from cafe5_amd import native, workloads as wl
from cafe5_amd.model import BaseModel, GammaModel
from bars import reduced_ok
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

PATHS = [native.PATH_PLAIN64, native.PATH_BATCHED64, native.PATH_BATCHED32, native.PATH_BATCHED16]


@pytest.fixture(scope="module")
def eng():
    e = native.Engine(0)
    yield e
    e.close()


def _rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    both_inf = np.isinf(a) & np.isinf(b) & (np.sign(a) == np.sign(b))
    d = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
    d[both_inf] = 0.0
    return d


def assert_lnl(got, want, path, what=""):
    got = np.asarray(got)
    want = np.asarray(want)
    assert np.array_equal(np.isfinite(got), np.isfinite(want)), (what, np.flatnonzero(np.isfinite(got) != np.isfinite(want))[:5])
    fin = np.isfinite(want)
    if path in native.REDUCED_PATHS:
        ok = reduced_ok(got[fin], want[fin])
        assert np.all(ok), (what, np.abs(got[fin] - want[fin]).max())
    else:
        r = _rel(got[fin], want[fin])
        assert np.all(r <= 1e-9), (what, r.max())


# --------------------------------------------------------------------------
# matrix build
# --------------------------------------------------------------------------

@pytest.mark.parametrize("algo", [native.BUILD_CLOSED_FORM, native.BUILD_RECURRENCE])
def test_matrix_build_vs_oracle(eng, algo):
    keys = [(0.05, 5.0), (0.002, 37.4), (0.0137, 11.0), (0.001, 1.0), (0.0099, 70.0), (1e-4, 3.0),
            (0.05, 25.0), (0.05, 20.0), (0.01, 0.0004)]
    for N in (5, 21, 101, 201):
        lq = [native.key_lambda(l) for l, _ in keys]
        tq = [native.key_branch_length(t) for _, t in keys]
        eng.matrix_reset(N, len(keys))
        eng.matrix_build(lq, tq, 0, algo)
        for i, (l, t) in enumerate(zip(lq, tq)):
            got = eng.matrix_get(i)
            want = Oracle.matrix(l, t, N)
            tol = 2e-12
            mask = want > 1e-280
            assert np.array_equal(got == 0, want == 0) or algo == native.BUILD_RECURRENCE, (N, i)
            assert np.all(_rel(got[mask], want[mask]) <= tol), (N, i, _rel(got[mask], want[mask]).max())
            assert np.all(np.abs(got[~mask]) <= 1e-270)


def test_matrix_5x5_known_values(eng):
    # test.cpp:595-617
    eng.matrix_reset(5, 1)
    eng.matrix_build([0.05], [5.0], 0, native.BUILD_RECURRENCE)
    m = eng.matrix_get(0)
    assert abs(m[1, 1] - 0.64) < 1e-15 and abs(m[4, 4] - 0.305562) < 1e-6 and abs(m[3, 2] - 0.26112) < 1e-15


def test_matrix_slot_miss_is_an_error(eng):
    eng.matrix_reset(11, 2)
    with pytest.raises(native.CafeGpuError):
        eng.matrix_get(1)


# --------------------------------------------------------------------------
# pruning (root vectors) and known-answer likelihoods
# --------------------------------------------------------------------------

def _setup_prune(eng, tree, counts, lam, B, R, algo=native.BUILD_RECURRENCE):
    model = BaseModel(eng, tree, counts, B, R, path=native.PATH_PLAIN64, build_algo=algo)
    keys = native.KeyTable()
    slots = keys.node_slots(tree, np.full((1, tree.n_nodes), lam))
    lq, tq = keys.arrays()
    eng.matrix_reset(max(B, R) + 1, len(lq))
    eng.matrix_build(lq, tq, 0, algo)
    return model, slots


@pytest.mark.parametrize("path", PATHS)
def test_root_vector_known_values(eng, path):
    # test.cpp:1662-1683: (A:1,B:3):7, A=3, B=6, lambda 0.045, B=R=20
    tree = wl.parse_newick("(A:1,B:3):7")
    counts = np.array([[3, 6]], dtype=np.int32)
    model, slots = _setup_prune(eng, tree, counts, 0.045, 20, 20, native.BUILD_CLOSED_FORM)
    got = eng.root_vectors(slots, np.full((1, 20), 0.05), path)[0]
    want = Oracle.prune(tree, counts[0], np.full(tree.n_nodes, 0.045), 20, 20)
    tol = 1e-5 if path in native.REDUCED_PATHS else 1e-12
    assert np.all(_rel(got, want) <= tol), _rel(got, want).max()


@pytest.mark.parametrize("path", PATHS)
def test_root_vectors_seeded_tree(eng, path):
    w = wl.config_base(seed=1, F=64)
    B = R = 40
    counts = np.minimum(w.counts, B)
    model, slots = _setup_prune(eng, w.tree, counts, 0.002, B, R)
    got = eng.root_vectors(slots, np.full((1, R), 1.0 / R), path)
    for f in range(counts.shape[0]):
        want = Oracle.prune(w.tree, counts[f], np.full(w.tree.n_nodes, 0.002), B, R)
        # split FP16 keeps about FP32 accuracy for entries within ~2^-50 of the
        # vector's maximum (what the likelihood uses); FP32 down to 1e-30
        big = want > want.max() * (1e-8 if path == native.PATH_BATCHED16 else 1e-30)
        tol = 2e-5 if path in native.REDUCED_PATHS else 1e-11
        assert np.all(_rel(got[f][big], want[big]) <= tol), (f, _rel(got[f][big], want[big]).max())


@pytest.mark.parametrize("path", PATHS)
def test_base_model_known_value(eng, path):
    # test.cpp:461-488
    tree = wl.parse_newick("(A:1,B:1);")
    counts = np.array([[1, 2], [2, 1], [3, 6], [6, 3]], dtype=np.int32)
    m = BaseModel(eng, tree, counts, 56, 8, path=path)
    total = m.infer_family_likelihoods(Oracle.prior_uniform(100, 8), 0.01)
    assert abs(total - 46.56632) < 1e-5
    res = Oracle.evaluate(tree, np.full(tree.n_nodes, 0.01), counts, 56, 8, Oracle.prior_uniform(100, 8))
    assert_lnl(m.results, res["lnl"], path)


@pytest.mark.parametrize("path", PATHS)
def test_gamma_known_values(eng, path):
    # test.cpp:1269-1294
    tree = wl.parse_newick("(A:1,B:3):7")
    prior = Oracle.prior_map([1, 2, 3, 4, 5], [2, 2, 2, 2, 1], 8)
    m = GammaModel(eng, tree, np.array([[3, 6]], dtype=np.int32), 10, 8, [0.01, 0.05], [0.1, 0.5], path=path)
    m.infer_family_likelihoods(prior, 0.005)
    cat = m.results["cat"][0]
    assert abs(math.log(cat[0]) - -23.04433) < 1e-5
    assert abs(math.log(cat[1]) - -16.68005) < 1e-5


@pytest.mark.parametrize("path", PATHS)
def test_gamma_saturation_returns_inf(eng, path):
    # test.cpp:1296-1311 (saturated t=3 branch at lambda 0.9 x 0.5)
    tree = wl.parse_newick("(A:1,B:3):7")
    m = GammaModel(eng, tree, np.array([[3, 6]], dtype=np.int32), 10, 8, [1.0, 1.0], [0.1, 0.5], path=path)
    # can_infer uses the longest branch (7) and the un-multiplied lambda: 0.9*7 > 1
    assert m.infer_family_likelihoods(Oracle.prior_uniform(100, 8), 0.9) == math.inf
    # a lambda that passes can_infer but saturates a category on the t=3 branch
    m2 = GammaModel(eng, tree, np.array([[3, 6]], dtype=np.int32), 10, 8, [0.5, 0.5], [0.1, 9.0], path=path)
    assert m2.infer_family_likelihoods(Oracle.prior_uniform(100, 8), 0.05) == math.inf


# --------------------------------------------------------------------------
# seeded workloads against the oracle
# --------------------------------------------------------------------------

@pytest.mark.parametrize("path", PATHS)
def test_base_model_seeded(eng, path):
    w = wl.config_base(seed=1, F=400)
    B = R = 100
    prior = Oracle.prior_uniform(R, R)
    m = BaseModel(eng, w.tree, w.counts, B, R, path=path)
    total = m.infer_family_likelihoods(prior, 0.002)
    ref = Oracle.evaluate(w.tree, np.full(w.tree.n_nodes, 0.002), w.counts, B, R, prior, threads=8)
    assert_lnl(m.results, ref["lnl"], path)
    assert abs(total - ref["total"]) <= 1e-6 * abs(ref["total"]) if path in native.REDUCED_PATHS else \
        abs(total - ref["total"]) <= 1e-9 * abs(ref["total"])


@pytest.mark.parametrize("path", PATHS)
def test_uneven_buckets_lambda_grid(eng, path):
    w = wl.config_uneven(seed=2, F=500, T=5)
    m = BaseModel(eng, w.tree, w.counts, w.bounds, w.bounds, path=path)
    totals = m.infer_grid(lambda r: Oracle.prior_uniform(r, r), w.lambdas)
    for ti, lam in enumerate(w.lambdas):
        for (B, R, _) in m.buckets:
            sel = np.flatnonzero(w.bounds == B)
            ref = Oracle.evaluate(w.tree, np.full(w.tree.n_nodes, lam), w.counts[sel], B, R,
                                  Oracle.prior_uniform(R, R), threads=8)
            assert_lnl(m.grid_lnl[ti, sel], ref["lnl"], path, f"lambda {lam} bucket {B}")
        want_total = -np.sum([Oracle.evaluate(w.tree, np.full(w.tree.n_nodes, lam),
                                              w.counts[w.bounds == B], B, B, Oracle.prior_uniform(B, B),
                                              threads=8)["lnl"].sum() for (B, _, _) in m.buckets])
        tol = 1e-6 if path in native.REDUCED_PATHS else 1e-9
        assert abs(totals[ti] - want_total) <= tol * abs(want_total), (lam, totals[ti], want_total)


@pytest.mark.parametrize("path", PATHS)
def test_gamma_model_seeded(eng, path):
    w = wl.config_gamma(seed=3, F=300, K=4, bound=60)
    B = R = 60
    f, r = Oracle.get_gamma(4, 0.5)
    prior = Oracle.prior_uniform(R, R)
    m = GammaModel(eng, w.tree, w.counts, B, R, f, r, path=path, alpha=0.5)
    total = m.infer_family_likelihoods(prior, 0.003)
    ref = Oracle.evaluate(w.tree, np.full(w.tree.n_nodes, 0.003), w.counts, B, R, prior, n_cat=4,
                          multipliers=r, cat_probs=f, threads=8)
    assert_lnl(m.results["lnl"], ref["lnl"], path)
    tol = 1e-5 if path in native.REDUCED_PATHS else 1e-9
    assert np.all(np.abs(m.results["post"] - ref["post"]) <= tol)
    assert abs(total - ref["total"]) <= (1e-6 if path in native.REDUCED_PATHS else 1e-9) * abs(ref["total"])


@pytest.mark.parametrize("path", PATHS)
def test_error_model_multi_lambda(eng, path):
    w = wl.config_error_multi(seed=4, F=200, bound=60)
    B = R = 60
    err = wl.error_table(0.05, B)
    prior = Oracle.prior_poisson(3.0, R, R)
    m = BaseModel(eng, w.tree, w.counts, B, R, error_model=err, path=path)
    total = m.infer_family_likelihoods(prior, w.lambdas, node_class=w.node_class)
    ref = Oracle.evaluate(w.tree, w.lambdas[w.node_class], w.counts, B, R, prior, err=err, threads=8)
    assert_lnl(m.results, ref["lnl"], path)


@pytest.mark.parametrize("path", PATHS)
def test_tiny_likelihood_semantics(eng, path):
    """A family far from the model (huge jumps) has a likelihood near or below
    the FP64 range; every path must reproduce the reference's value or -inf."""
    tree = wl.parse_newick("((A:10,B:10):30,(C:20,D:20):20);")
    counts = np.array([[0, 95, 0, 95], [1, 1, 2, 1], [90, 0, 0, 90], [0, 0, 99, 0]], dtype=np.int32)
    B = R = 100
    prior = Oracle.prior_uniform(R, R)
    for lam in (1e-4, 2e-3):
        m = BaseModel(eng, tree, counts, B, R, path=path)
        m.infer_family_likelihoods(prior, lam)
        ref = Oracle.evaluate(tree, np.full(tree.n_nodes, lam), counts, B, R, prior)
        assert np.array_equal(np.isfinite(m.results), np.isfinite(ref["lnl"])), (lam, m.results, ref["lnl"])
        assert_lnl(m.results, ref["lnl"], path, f"lambda {lam}")


# --------------------------------------------------------------------------
# large bounds: vectors of 32 (B = 500) and 38 (B = 600) 16-row tiles, the
# widest kernel shapes (8 and 16 waves per workgroup)
# --------------------------------------------------------------------------

_LARGE = {}


def _large_case(B, K):
    """10-taxon tree, 40 families with root sizes spread up to 0.85 B, lambda
    0.002; oracle results cached per (B, K) across the three paths."""
    if (B, K) not in _LARGE:
        tree = wl.yule_tree(10, 11)
        roots = np.linspace(1, int(0.85 * B), 40).astype(np.int64)
        counts = wl.simulate_families(tree, roots, np.full(tree.n_nodes, 0.002), 13, cap=B - 1)
        prior = Oracle.prior_uniform(B, B)
        if K:
            f, r = Oracle.get_gamma(K, 1.0)
            ref = Oracle.evaluate(tree, np.full(tree.n_nodes, 0.002), counts, B, B, prior, n_cat=K,
                                  multipliers=r, cat_probs=f, threads=16)
        else:
            f = r = None
            ref = Oracle.evaluate(tree, np.full(tree.n_nodes, 0.002), counts, B, B, prior, threads=16)
        _LARGE[(B, K)] = (tree, counts, prior, f, r, ref)
    return _LARGE[(B, K)]


@pytest.mark.parametrize("path", PATHS)
def test_large_bound_gamma_500(eng, path):
    tree, counts, prior, f, r, ref = _large_case(500, 2)
    assert counts.max() > 400   # rows of the last tiles carry data
    m = GammaModel(eng, tree, counts, 500, 500, f, r, path=path, alpha=1.0)
    total = m.infer_family_likelihoods(prior, 0.002)
    assert_lnl(m.results["lnl"], ref["lnl"], path)
    tol = 1e-6 if path in native.REDUCED_PATHS else 1e-9
    assert abs(total - ref["total"]) <= tol * abs(ref["total"])


@pytest.mark.parametrize("path", PATHS)
def test_large_bound_base_600(eng, path):
    tree, counts, prior, _, _, ref = _large_case(600, 0)
    assert counts.max() > 512   # rows beyond the first 32 tiles carry data
    m = BaseModel(eng, tree, counts, 600, 600, path=path)
    total = m.infer_family_likelihoods(prior, 0.002)
    assert_lnl(m.results, ref["lnl"], path)
    tol = 1e-6 if path in native.REDUCED_PATHS else 1e-9
    assert abs(total - ref["total"]) <= tol * abs(ref["total"])
