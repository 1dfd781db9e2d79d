"""cafe_gpu_reconstruct_noalloc -- the reconstruction on the library's fixed
32 MiB device scratch, with no device allocation -- against
cafe_gpu_reconstruct on the same inputs.  States are integers and must be
equal: on two calls in a row, the second with fewer families and smaller
vectors (the fixed scratch keeps the first call's contents), and against the
oracle on a few families.  A family too large for the fixed scratch is
rejected before anything runs."""
import numpy as np
import pytest

from cafe5_amd import native, priors, workloads as wl
from cafe5_amd.native import CafeGpuError
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = native.Engine(0)
    yield e
    e.close()


def _slots(eng, tree, lam, N):
    keys = native.KeyTable()
    slots = keys.node_slots(tree, np.full((1, tree.n_nodes), lam))
    lq, tq = keys.arrays()
    eng.matrix_reset(N, len(lq))
    eng.matrix_build(lq, tq, 0, native.BUILD_RECURRENCE)
    return slots


def test_noalloc_matches_reconstruct_on_two_calls(eng):
    w = wl.config_uneven(seed=2, F=3000, T=2)
    lam = 0.003
    # call 1: 3,000 families at B = R = 60 (67 nodes, about 49 KB of rows per
    # family: several launches through the fixed scratch)
    B, R = 60, 60
    counts = np.minimum(w.counts, B).astype(np.int32)
    eng.set_tree(w.tree)
    eng.set_families(counts, B, R)
    slots = _slots(eng, w.tree, lam, B + 1)
    prior = priors.uniform(R, R + 1)
    want = eng.reconstruct(slots, prior, B, R)
    got = eng.reconstruct(slots, prior, B, R, noalloc=True)
    print(f"noalloc call 1: {got.shape[0]} families, B={B}", flush=True)
    assert np.array_equal(got, want), np.argwhere(got != want)[:5]
    # call 2: fewer families, smaller vectors, a family range inside the table
    B2, R2 = 30, 24
    counts2 = np.minimum(counts[:500], B2).astype(np.int32)
    eng.set_families(counts2, B2, R2)
    prior2 = priors.uniform(R2, min(B2, R2) + 1)
    want2 = eng.reconstruct(slots, prior2, B2, R2, f0=7, nf=300)
    got2 = eng.reconstruct(slots, prior2, B2, R2, f0=7, nf=300, noalloc=True)
    print(f"noalloc call 2: {got2.shape[0]} families, B={B2}", flush=True)
    assert np.array_equal(got2, want2), np.argwhere(got2 != want2)[:5]
    for f in (0, 151, 299):
        ref = Oracle.reconstruct(w.tree, counts2[7 + f], lam, B2, R2, prior2, N=B + 1)
        assert np.array_equal(got2[f], ref), f


def test_noalloc_rejects_a_family_larger_than_the_fixed_scratch(eng):
    # a caterpillar of 751 leaves (1,501 nodes) at B = 2000: one family's rows
    # need 2001 x (1501 x 12 + 8) bytes, more than the 32 MiB fixed scratch
    n_leaves = 751
    n = 2 * n_leaves - 1
    parent = np.full(n, -1, dtype=np.int32)
    leaf_col = np.full(n, -1, dtype=np.int32)
    # leaves 0..750, interior nodes 751..1499 chained, root = n - 1
    for i in range(n_leaves):
        leaf_col[i] = i
    interior = list(range(n_leaves, n))
    parent[0] = interior[0]
    parent[1] = interior[0]
    for k in range(1, len(interior)):
        parent[interior[k - 1]] = interior[k]
        parent[k + 1] = interior[k]
    tree = wl.Tree(parent, leaf_col, np.ones(n), [f"n{i}" for i in range(n)])
    B = R = 2000
    eng.set_tree(tree)
    eng.set_families(np.zeros((1, n_leaves), dtype=np.int32), B, R)
    eng.matrix_reset(B + 1, 1)
    prior = priors.uniform(R, R + 1)
    with pytest.raises(CafeGpuError) as ei:
        eng.reconstruct(np.zeros(n, dtype=np.int32), prior, B, R, noalloc=True)
    assert "fixed scratch holds" in str(ei.value), str(ei.value)
