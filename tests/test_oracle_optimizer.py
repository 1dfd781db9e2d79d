"""Pins of the oracle's Nelder-Mead restatement (oracle/nelder_mead.py,
src/optimizer.cpp:61-330): the reference's step tests (test.cpp:2802-3030)
and convergence to known minima."""
import math
import os
import sys

import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import nelder_mead as nmr  # noqa: E402


def product(x):
    # test.cpp:2844-2862: the product of the first two values
    return x[0] * x[1]


def fresh():
    return nmr.NelderMead(product, 2)


def test_sort():
    nm = fresh()
    c = list(nm.candidates)
    c[0].score, c[1].score, c[2].score = 3, 5, 1
    nm.sort()
    assert nm.candidates == [c[2], c[0], c[1]]          # test.cpp:2802-2813


def test_check_values_and_scores():
    nm = fresh()
    for c, v in zip(nm.candidates, [(1, 3), (2, 4), (3, 5)]):
        c.values = list(v)
    nm.tolx = 3
    assert nm.check_values()                              # test.cpp:2815-2827
    nm.tolx = .5
    assert not nm.check_values()
    for c, s in zip(nm.candidates, (1.0, 3.0, 5.0)):
        c.score = s
    nm.tolf = 5
    assert nm.check_scores()                              # :2829-2839
    nm.tolf = 1
    assert not nm.check_scores()


def test_min_init():
    nm = fresh()
    nm.min_init([5, 3])                                   # :2864-2882
    assert nm.candidates[0].score == 15
    assert nm.candidates[1].score == 15.75
    assert abs(nm.candidates[2].score - 15.75) < 1e-12


def test_mean_reflection_expansion_contractions():
    nm = fresh()
    nm.candidates[0].values = [300, 12]
    nm.candidates[1].values = [200, 44]
    nm.mean()
    assert nm.x_mean == [250, 28]                         # :2884-2898
    assert nm.reflection() == 28000 and nm.x_r == [500, 56]      # :2900-2918 (third point (0, 0))
    assert nm.expansion() == 63000 and nm.x_tmp == [750, 84]     # :2920-2938
    assert nm.contract_outside() == 15750 and nm.x_tmp == [375, 42]   # :2940-2958
    nm.candidates[2].values = [26, 12]
    assert nm.contract_inside() == 13032 and nm.x_tmp == [362, 36]    # :2961-2981


def test_shrink_and_set_last():
    nm = fresh()
    for c, v in zip(nm.candidates, [(300, 200), (42, 64), (26, 12)]):
        c.values = list(v)
    nm.shrink()                                           # :2983-3008
    assert [c.values for c in nm.candidates] == [[300, 200], [163, 106], [171, 132]]
    nm = fresh()
    for c, v, s in zip(nm.candidates, [(300, 200), (42, 64), (26, 12)], (2, 4, 6)):
        c.values, c.score = list(v), s
    nm.set_last([99, 14], 3)                              # :3010-3033
    assert [c.score for c in nm.candidates] == [2, 3, 4]


@pytest.mark.parametrize("x0", [[0.5, 0.5], [-1.2, 1.0], [3.0, -2.0]])
def test_converges_on_a_quadratic(x0):
    f = lambda x: (x[0] - 1) ** 2 + 2 * (x[1] - 2) ** 2 + 3
    vals, score, iters = nmr.optimize(f, x0)
    assert abs(vals[0] - 1) < 2e-3 and abs(vals[1] - 2) < 2e-3
    assert abs(score - 3) < 1e-5 and iters < 300


def test_converges_in_one_dimension_and_respects_the_cap():
    f = lambda x: 1e6 * (x[0] - 0.002) ** 2 if x[0] > 0 else math.inf
    vals, score, iters = nmr.optimize(f, [0.004])
    assert abs(vals[0] - 0.002) < 2e-5
    vals, score, iters = nmr.optimize(lambda x: (x[0] - 1) ** 2 + (x[1] + 1) ** 2, [40.0, 40.0], iterations=5)
    assert iters == 5


def test_zero_start_uses_the_absolute_step():
    seen = []
    nmr.optimize(lambda x: (seen.append(list(x)), (x[0] - 1) ** 2)[1], [0.0], iterations=0)
    assert seen == [[0.0], [0.00025]]                     # optimizer.cpp:170-181
