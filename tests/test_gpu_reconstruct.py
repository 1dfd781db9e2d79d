"""Pupko joint reconstruction on the GPU (cafe_gpu_reconstruct) against the
oracle's restatement (pinned to the reference in test_oracle_reconstruct):
the reconstructed states are integers and must be equal.  The products and
comparisons are the reference's in its order on FP64 matrices; the closed-form
build agrees with the oracle's matrices to 2e-12, so a state could only differ
where two candidates tie to that precision, which the seeded cases avoid."""
import numpy as np
import pytest

from cafe5_amd import native, priors, workloads as wl
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = native.Engine(0)
    yield e
    e.close()


def _setup(eng, tree, counts, lam, B, R, algo):
    eng.set_tree(tree)
    eng.set_families(counts, B, R)
    keys = native.KeyTable()
    slots = keys.node_slots(tree, np.full((1, tree.n_nodes), lam))
    lq, tq = keys.arrays()
    eng.matrix_reset(max(B, R) + 1, len(lq))
    eng.matrix_build(lq, tq, 0, algo)
    return slots


def test_reference_fixture(eng):
    # test.cpp:1090-1118: the root (AB) reconstructs to 4
    tree = wl.parse_newick("(A:1,B:3):7;")
    counts = np.array([[{"A": 3, "B": 6}[n] for n in tree.leaf_names()]], dtype=np.int32)
    B, R = 10, 8
    slots = _setup(eng, tree, counts, 0.005, B, R, native.BUILD_CLOSED_FORM)
    prior = Oracle.prior_map(list(range(9)), [1, 2, 3, 4, 5, 4, 3, 2, 1], 9)
    states = eng.reconstruct(slots, prior, B, R)
    assert states[0][tree.names.index("AB")] == 4
    assert np.array_equal(states[0], Oracle.reconstruct(tree, counts[0], 0.005, B, R, prior, N=max(B, R) + 1))


@pytest.mark.parametrize("algo", [native.BUILD_CLOSED_FORM, native.BUILD_RECURRENCE])
@pytest.mark.parametrize("B,R", [(30, 30), (70, 40), (130, 130)])
def test_seeded_families(eng, algo, B, R):
    w = wl.config_base(seed=3, F=96)
    counts = np.minimum(w.counts, B).astype(np.int32)
    lam = 0.004
    slots = _setup(eng, w.tree, counts, lam, B, R, algo)
    prior = np.zeros(min(B, R) + 1)
    prior[:] = priors.poisson(2.5, R, min(B, R) + 1) if R == 40 else priors.uniform(R, min(B, R) + 1)
    got = eng.reconstruct(slots, prior, B, R)
    for f in range(counts.shape[0]):
        want = Oracle.reconstruct(w.tree, counts[f], lam, B, R, prior)
        assert np.array_equal(got[f], want), (f, np.flatnonzero(got[f] != want))


def test_large_bound_one_family_per_wave(eng):
    # B >= 256 runs k_reconstruct (one family per workgroup, matrices read from
    # memory); smaller bounds run k_reconstruct_shared.  Three families: the
    # oracle builds its 301 x 301 matrices for every family (~6 s each)
    algo = native.BUILD_RECURRENCE
    w = wl.config_base(seed=5, F=3)
    B = R = 300
    counts = np.minimum(w.counts, B).astype(np.int32)
    lam = 0.004
    slots = _setup(eng, w.tree, counts, lam, B, R, algo)
    prior = priors.uniform(R, R + 1)
    got = eng.reconstruct(slots, prior, B, R)
    for f in range(counts.shape[0]):
        want = Oracle.reconstruct(w.tree, counts[f], lam, B, R, prior)
        assert np.array_equal(got[f], want), (f, np.flatnonzero(got[f] != want))


def test_bound_past_1024_rows(eng):
    # bounds of 1,100: k_reconstruct on vectors of 1,101 rows, the sizes the
    # likelihood path runs on its plain kernel; one branch length, so the
    # oracle builds its one 1,101 x 1,101 matrix once for every family
    tree = wl.parse_newick("((A:5,B:5):5,(C:5,D:5):5);")
    B = R = 1100
    lam = 0.002
    counts = wl.simulate_families(tree, np.array([1, 350, 700, 900]), np.full(tree.n_nodes, lam), 9, cap=B - 1)
    assert counts.max() > 850
    slots = _setup(eng, tree, counts, lam, B, R, native.BUILD_RECURRENCE)
    prior = priors.uniform(R, R + 1)
    got = eng.reconstruct(slots, prior, B, R)
    mats = Oracle.node_matrices(tree, np.full(tree.n_nodes, lam), B + 1)
    for f in range(counts.shape[0]):
        want = Oracle.reconstruct(tree, counts[f], lam, B, R, prior, matrices=mats)
        assert np.array_equal(got[f], want), (f, np.flatnonzero(got[f] != want))


def test_many_families_in_chunks(eng):
    # more families than one launch holds (8192): every chunk's rows land in place
    w = wl.config_base(seed=8, F=9000)
    B = R = 24
    counts = np.minimum(w.counts, B).astype(np.int32)
    slots = _setup(eng, w.tree, counts, 0.003, B, R, native.BUILD_CLOSED_FORM)
    prior = priors.uniform(R, R + 1)
    got = eng.reconstruct(slots, prior, B, R)
    rng = np.random.default_rng(0)
    for f in np.concatenate([rng.choice(counts.shape[0], 40, replace=False), [0, 8191, 8192, 8999]]):
        want = Oracle.reconstruct(w.tree, counts[f], 0.003, B, R, prior)
        assert np.array_equal(got[f], want), f
