"""Pins of the oracle's p-value functions (compute_family_probabilities,
pvalue, find_best_pvalue; src/probability.cpp:448-541): the reference's own
test literals, exact rational brute force on small trees, and, where
oracle/_ref is built, the compiled reference itself."""
from fractions import Fraction

import numpy as np
import pytest

from cafe5_amd import workloads as wl
from oracle_lib import Oracle, Ref
from test_oracle_pins import exact_matrix


def _steps_of_001(n):
    # test.cpp:1229-1231: n += 0.01, ten times
    out, v = [], 0.0
    for _ in range(n):
        v += 0.01
        out.append(v)
    return out


def test_pvalue_reference_values():
    cd = _steps_of_001(10)
    assert Oracle.pvalue(0.05, cd) == 0.5           # test.cpp:1232
    assert Oracle.pvalue(0.0001, cd) == 0.0         # :1233
    assert Oracle.pvalue(0.099, cd) == 0.9          # :1234
    v = Oracle.pvalue(0.35, [.1, .2, .3, .4, .5, .6, .7, .8, .9])
    assert abs(v - 3.0 / 9.0) < 1e-15               # :1237-1243
    # nothing above v: the last index
    assert Oracle.pvalue(5.0, [1.0, 2.0, 3.0]) == 2.0 / 3.0
    # ties: upper_bound passes every equal value
    assert Oracle.pvalue(2.0, [1.0, 2.0, 2.0, 3.0]) == 3.0 / 4.0


def test_find_best_pvalue_reference_values():
    # test.cpp:2276-2300, family of the Inference fixture (A = 1, B = 2)
    cd = np.array([[.1, .2, .3, .4]] * 3)
    root = np.zeros(3)
    root[0] = .15
    assert Oracle.find_best_pvalue(root, 2, cd) == 0.25
    root[2] = .35                                     # j = 2 is past rint(2 * 1.25) = 2
    assert Oracle.find_best_pvalue(root, 2, cd) == 0.25
    root[1] = .35
    assert Oracle.find_best_pvalue(root, 2, cd) == 0.75


def _alpha(lam, t):
    lq, tq = Oracle.key(lam, t)
    return Fraction(lq * tq / (1 + lq * tq))


def test_family_probability_cherry_exact():
    tree = wl.parse_newick("(A:2,B:5);")
    lam, B = 0.03, 14
    counts = np.array([4, 7])
    a, b = [i for i in range(tree.n_nodes) if tree.leaf_col[i] >= 0]
    ma = exact_matrix(_alpha(lam, 2.0), B + 2)
    mb = exact_matrix(_alpha(lam, 5.0), B + 2)
    for r in (1, 3, 6, 9):
        got = Oracle.family_probability(tree, counts, lam, B, max_size=7, r=r)
        e = ma[r][4] * mb[r][7]
        assert abs(got - float(e)) <= 1e-13 * float(e), (r, got, float(e))


def test_family_probability_truncation():
    # interior vectors stop at m = min(B, max + max(50, max / 5)); a root of
    # 60 under a near-critical process (lambda t = 0.99) leaves mass on
    # interior sizes past m = 112 (2.6e-12 of the total), so summing over
    # sizes 0..m and 0..B differ by far more than rounding
    tree = wl.parse_newick("((A:9,B:9):9,C:9);")
    lam, B, r = 0.11, 200, 60
    counts = np.array([60, 62, 58])
    max_size = 62
    m = min(B, max_size + max(50, max_size // 5))
    assert m == 112
    lq, tq = Oracle.key(lam, 9.0)
    M = Oracle.matrix(lq, tq, B + 1)

    def by_sum(limit):
        inner = sum(M[r][x] * M[x][60] * M[x][62] for x in range(limit + 1))
        return inner * M[r][58]

    got = Oracle.family_probability(tree, counts, lam, B, max_size, r)
    assert abs(got - by_sum(m)) <= 1e-14 * by_sum(m)
    assert abs(by_sum(B) - by_sum(m)) > 1e-13 * by_sum(m)


def test_family_probability_reference_fixture():
    # test.cpp:1245-1266: the Reconstruction fixture's family with every
    # interior size 5, root size 5, B = 20, R = 15, matrices of size 100;
    # the test checks Approx(1e-10) at doctest's default epsilon
    tree = wl.parse_newick("((A:1,B:3):7,(C:11,D:17):23);")
    names = tree.leaf_names()
    fam = {"A": 11, "B": 2, "C": 5, "D": 6}
    counts = np.array([fam[n] for n in names])
    got = Oracle.family_probability(tree, counts, 0.03, 20, max_size=11, r=5, N=100)
    eps = 1.1920928955078125e-05 * (1 + max(abs(got), 1e-10))
    assert abs(got - 1e-10) < eps
    assert got > 0


needs_ref = pytest.mark.skipif(not Ref.available(), reason="oracle/_ref not built (no CAFE5 tree)")


@needs_ref
def test_ref_family_probabilities_bitwise():
    rng = np.random.default_rng(11)
    tree = wl.yule_tree(7, 5)
    for lam, B, R in ((0.004, 30, 25), (0.02, 60, 40)):
        F = 12
        counts = rng.integers(0, 12, (F, tree.n_leaves)).astype(np.int32)
        inner = rng.integers(1, 15, F).astype(np.int32)
        for r in (1, 5, 17):
            ref = Ref.family_probabilities(tree, counts, inner, lam, B, R, r)
            got = np.array([Oracle.family_probability(tree, counts[f], lam, B,
                                                      max(int(counts[f].max()), int(inner[f])), r,
                                                      N=max(R, B) + 1) for f in range(F)])
            assert np.array_equal(ref, got), (lam, r)


@needs_ref
def test_ref_pvalues():
    rng = np.random.default_rng(5)
    for _ in range(40):
        n = int(rng.integers(1, 30))
        cd = np.sort(rng.choice(np.linspace(0, 1, 11), n))
        v = float(rng.choice(np.concatenate([cd, rng.uniform(-0.1, 1.1, 3)])))
        assert Ref.pvalue(v, cd) == Oracle.pvalue(v, cd)
    for _ in range(30):
        n_root = int(rng.integers(3, 20))
        max_size = int(rng.integers(0, int(n_root / 1.25) + 1))
        cd = np.sort(rng.uniform(0, 1, (n_root, int(rng.integers(1, 12)))), axis=1)
        root = rng.uniform(0, 1, n_root)
        assert Ref.find_best_pvalue(root, max_size, cd) == Oracle.find_best_pvalue(root, max_size, cd)


@needs_ref
def test_ref_create_family_reference_values():
    # test.cpp:636-654: engine seeded 10, lambda 0.2, matrices of size 15,
    # B = 10, R = 8, root size 6
    tree = wl.parse_newick("((A:1,B:1):1,(C:1,D:1):1);")
    got = Ref.create_families(tree, 0.2, 10, 8, 15, 10, 6, 1)[0]
    sizes = dict(zip(tree.leaf_names(), got))
    assert (sizes["A"], sizes["B"], sizes["C"], sizes["D"]) == (5, 5, 7, 4)
