"""Error paths of the C ABI (include/cafe_gpu.h): inputs the engine must
reject before any kernel runs, with the message cafe_gpu_last_error carries,
and an engine that still evaluates correctly afterwards.  The reference's
counterparts throw std::runtime_error (io.cpp:134-217 for family files,
error_model.cpp:31-50 for error rows, matrix_cache.cpp:88-104 for a missing
matrix)."""
import numpy as np
import pytest

from cafe5_amd import native, workloads as wl
from cafe5_amd.native import CafeGpuError, _dp, _ip
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

TREE = "((A:2,B:3):4,(C:1,D:5):2);"


@pytest.fixture(scope="module")
def eng():
    e = native.Engine(0)
    yield e
    e.close()


def _raises(fn, text):
    with pytest.raises(CafeGpuError) as ei:
        fn()
    assert text in str(ei.value), str(ei.value)


def test_family_table_errors(eng):
    tree = wl.parse_newick(TREE)
    eng.set_tree(tree)
    _raises(lambda: eng.set_families(np.zeros((0, 4), dtype=np.int32), 10, 10), "at least one family")
    _raises(lambda: eng.set_families(np.array([[1, 2, 11, 0]]), 10, 10), "count 11 outside 0..10")
    _raises(lambda: eng.set_families(np.array([[1, -1, 3, 0]]), 10, 10), "count -1 outside 0..10")
    _raises(lambda: eng.set_families(np.array([[1, 2, 3, 0]]), bound_B=np.array([0]), bound_R=np.array([5])),
            "non-positive bound")


def test_tree_errors(eng):
    def set_tree(parent, leaf_col):
        p = np.ascontiguousarray(parent, dtype=np.int32)
        c = np.ascontiguousarray(leaf_col, dtype=np.int32)
        eng._check(eng.lib.cafe_gpu_set_tree(eng.ctx, len(p), _ip(p), _ip(c)))

    _raises(lambda: set_tree([2, 2, -1, -1], [0, 1, -1, -1]), "more than one root")
    _raises(lambda: set_tree([2, 0, -1], [0, 1, -1]), "leaf 0 has children")
    _raises(lambda: set_tree([2, 2, -1], [0, 0, -1]), "used twice")
    _raises(lambda: set_tree([-1], [-1]), "tree needs >= 2 nodes")


def test_matrix_and_error_model_errors(eng):
    _raises(lambda: eng.matrix_reset(1, 4), "out of range")
    _raises(lambda: eng.matrix_reset(11, 0), "capacity must be positive")
    _raises(lambda: eng.set_error_model(np.full((11, 2), 0.5)), "odd number of deviations")
    eng.matrix_reset(11, 8)
    _raises(lambda: eng.matrix_build([0.01], [-1.0], 0, native.BUILD_RECURRENCE), "invalid lambda or branch length")
    _raises(lambda: eng.matrix_build([0.01] * 9, [1.0] * 9, 0, native.BUILD_RECURRENCE), "exceed capacity")


def test_evaluation_errors_then_a_valid_evaluation(eng):
    tree = wl.parse_newick(TREE)
    eng.set_tree(tree)
    counts = np.array([[1, 2, 3, 0], [4, 4, 5, 6], [0, 0, 1, 0]], dtype=np.int32)
    B = R = 10
    eng.set_families(counts, B, R)
    keys = native.KeyTable()
    slots = keys.node_slots(tree, np.full((1, tree.n_nodes), 0.02))
    lq, tq = keys.arrays()
    prior = Oracle.prior_uniform(R, R)[None, :]
    # matrices smaller than the bucket needs
    eng.matrix_reset(8, len(lq))
    eng.matrix_build(lq, tq, 0, native.BUILD_RECURRENCE)
    _raises(lambda: eng.evaluate(slots, prior, path=native.PATH_PLAIN64), "needs matrices of size 11")
    eng.matrix_reset(B + 1, len(lq))
    # a slot that was never built
    _raises(lambda: eng.evaluate(slots, prior, path=native.PATH_PLAIN64), "not built")
    eng.matrix_build(lq, tq, 0, native.BUILD_RECURRENCE)
    # a prior row shorter than R, an error model that stops below the bound
    _raises(lambda: eng.evaluate(slots, prior[:, :5], path=native.PATH_PLAIN64), "prior_stride 5 < max R 10")
    eng.set_error_model(wl.error_table(0.05, 5))
    _raises(lambda: eng.evaluate(slots, prior, path=native.PATH_PLAIN64), "does not cover bucket")
    eng.set_error_model(None)
    # the engine still evaluates correctly
    for path in (native.PATH_PLAIN64, native.PATH_BATCHED32):
        got = eng.evaluate(slots, prior, path=path)["lnl"][0]
        want = Oracle.evaluate(tree, np.full(tree.n_nodes, 0.02), counts, B, R, prior[0])["lnl"]
        tol = 1e-9 if path == native.PATH_PLAIN64 else 1e-6
        assert np.all(np.abs(got - want) <= tol * np.abs(want)), (path, got, want)


def _small_setup(eng):
    tree = wl.parse_newick(TREE)
    eng.set_tree(tree)
    counts = np.array([[1, 2, 3, 0], [4, 4, 5, 6], [0, 0, 1, 0]], dtype=np.int32)
    B = R = 10
    eng.set_families(counts, B, R)
    keys = native.KeyTable()
    slots = keys.node_slots(tree, np.full((1, tree.n_nodes), 0.02))
    lq, tq = keys.arrays()
    prior = Oracle.prior_uniform(R, R)[None, :]
    return tree, counts, B, R, slots, lq, tq, prior


def test_reconstruct_rejects_mismatched_leaves_and_unbuilt_slots(eng):
    tree, counts, B, R, slots, lq, tq, prior = _small_setup(eng)
    eng.matrix_reset(B + 1, len(lq))
    rprior = Oracle.prior_uniform(R, R + 1)
    # gene_family_reconstructor.cpp:84-85 / matrix_cache.cpp:88-104: no matrix, no reconstruction
    _raises(lambda: eng.reconstruct(slots[0], rprior, B, R), "not built")
    eng.matrix_build(lq, tq, 0, native.BUILD_RECURRENCE)
    # a family table with fewer leaf columns than the tree has leaves
    eng.set_families(counts[:, :3], B, R)
    _raises(lambda: eng.reconstruct(slots[0], rprior, B, R), "families have 3 leaves, tree has 4")
    eng.set_families(counts, B, R)
    got = eng.reconstruct(slots[0], rprior, B, R)
    for f in range(counts.shape[0]):
        want = Oracle.reconstruct(tree, counts[f], np.full(tree.n_nodes, 0.02), B, R, rprior)
        assert np.array_equal(got[f], want), (f, got[f], want)


def test_plan_is_stale_after_inputs_change(eng):
    tree, counts, B, R, slots, lq, tq, prior = _small_setup(eng)
    eng.set_error_model(wl.error_table(0.05, B))
    eng.matrix_reset(B + 1, len(lq))
    p = eng.plan(slots, prior, lq, tq, path=native.PATH_BATCHED32)
    p.execute()
    first = p.results()["total"][0]
    # same-shape error rows are updated in place: the plan stays valid
    eng.set_error_model(wl.error_table(0.06, B))
    p.execute()
    assert p.results()["total"][0] != first
    # a new family table, tree or error-model shape makes the plan stale
    for change in (lambda: eng.set_families(counts, B, R), lambda: eng.set_tree(tree),
                   lambda: eng.set_error_model(None)):
        change()
        _raises(p.execute, "changed after the plan was made")
        _raises(p.results, "changed after the plan was made")
    p.close()
    eng.set_error_model(None)


def test_plan_built_slots_serve_the_fp64_paths(eng):
    """A plan's build writes the FP64 matrices and only the FP32 copies its
    own kernels read, so an eager FP32 evaluation of its slots asks for a
    full build first; the FP64 paths use them as they are."""
    tree, counts, B, R, slots, lq, tq, prior = _small_setup(eng)
    eng.matrix_reset(B + 1, len(lq))
    p = eng.plan(slots, prior, lq, tq, path=native.PATH_BATCHED32)
    p.execute()
    p.synchronize()
    planned = p.results()["lnl"][0].copy()
    p.close()
    for path in (native.PATH_BATCHED32, native.PATH_BATCHED16):
        _raises(lambda: eng.evaluate(slots, prior, path=path), "only the FP32 copies a plan read")
    want = Oracle.evaluate(tree, np.full(tree.n_nodes, 0.02), counts, B, R, prior[0])["lnl"]
    for path in (native.PATH_PLAIN64, native.PATH_BATCHED64):
        got = eng.evaluate(slots, prior, path=path)["lnl"][0]
        assert np.allclose(got, want, rtol=1e-9, atol=0), (path, got, want)
    assert np.allclose(planned, want, rtol=1e-6, atol=0)
    eng.matrix_build(lq, tq, 0, native.BUILD_RECURRENCE)
    got = eng.evaluate(slots, prior, path=native.PATH_BATCHED32)["lnl"][0]
    assert np.allclose(got, want, rtol=1e-6, atol=0)


def test_failed_plan_leaves_slots_unbuilt(eng):
    tree, counts, B, R, slots, lq, tq, prior = _small_setup(eng)
    eng.matrix_reset(B + 1, len(lq))
    with pytest.raises(CafeGpuError):
        eng.plan(slots, prior[:, :5], lq, tq, path=native.PATH_PLAIN64)   # prior shorter than R
    _raises(lambda: eng.matrix_get(0), "not built")
    p = eng.plan(slots, prior, lq, tq, path=native.PATH_PLAIN64)
    _raises(lambda: eng.matrix_get(0), "not built")   # built by the first execution
    p.execute()
    p.synchronize()
    want = Oracle.matrix(lq[0], tq[0], B + 1)
    assert np.allclose(eng.matrix_get(0), want, rtol=1e-12, atol=1e-300)
    p.close()
