"""The per-lane kernel's program on trees of other shapes.

The lane program keeps the vector each op reads from the op just before it in
a register array and stores only the first of two interior siblings in slots;
a tree whose program needs more arrays than the kernel's instances hold (five
for vectors of 21 rows, four for 33) runs on the matrix cores.  Each tree is
checked against the FP64 oracle on both reduced paths at bounds 20 and 33
rows (B = 20 and 32, the per-lane size classes), and the launched kernel is
the one the program's needs call for: a polytomy (one stored slot), a
caterpillar (none), balanced trees of 16 (three), 32 (four) and 64 leaves
(five).
"""
import numpy as np
import pytest

from cafe5_amd import native, priors, workloads as wl
from cafe5_amd.model import BaseModel
from bars import reduced_ok
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu


def _balanced(depth: int, bl: float = 4.0) -> str:
    names = iter(range(1 << 20))

    def rec(k):
        if k == 0:
            return f"T{next(names)}:{bl}"
        return f"({rec(k - 1)},{rec(k - 1)}):{bl}"

    return f"({rec(depth - 1)},{rec(depth - 1)});"


TREES = {
    "polytomy": "((A:5,B:5,C:5):10,(D:3,E:3,F:3,G:3):12,H:15);",
    "caterpillar": "((((((A:2,B:2):3,C:5):4,D:9):2,E:11):5,F:16):3,G:19);",
    "balanced16": _balanced(4),
    "balanced32": _balanced(5),
    "balanced64": _balanced(6),
}
# stored slots of the lane program, plus the hot array
ARRAYS = {"polytomy": 2, "caterpillar": 2, "balanced16": 4, "balanced32": 5, "balanced64": 6}
MAX_ARRAYS = {20: 5, 32: 4}   # the instances for 21 and 33 rows
LAM = 0.01


@pytest.fixture(scope="module")
def eng():
    e = native.Engine(0)
    yield e
    e.close()


def _families(tree, B, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    roots = 1 + rng.poisson(3.0, 300)
    return wl.simulate_families(tree, roots, np.full(tree.n_nodes, LAM), seed + 1, cap=B)


@pytest.mark.parametrize("B", [20, 32])
@pytest.mark.parametrize("name", list(TREES))
def test_lane_program_matches_oracle(eng, name, B):
    tree = wl.parse_newick(TREES[name])
    counts = _families(tree, B, 11 + B)
    prior = Oracle.prior_uniform(B, B)
    ref = Oracle.evaluate(tree, np.full(tree.n_nodes, LAM), counts, B, B, prior, threads=8)
    fin = np.isfinite(ref["lnl"])
    for path in (native.PATH_BATCHED32, native.PATH_BATCHED16):
        m = BaseModel(eng, tree, counts, B, B, path=path)
        total = m.infer_family_likelihoods(prior, LAM)
        got = np.asarray(m.results)
        assert np.array_equal(np.isfinite(got), fin), (name, B, path)
        assert np.all(reduced_ok(got[fin], ref["lnl"][fin])), (name, B, path, np.abs(got[fin] - ref["lnl"][fin]).max())
        assert abs(total - ref["total"]) <= 1e-6 * abs(ref["total"]), (name, B, path, total, ref["total"])

    # the launched kernel: per lane when the program fits an instance
    keys = native.KeyTable()
    slots = keys.node_slots(tree, np.full((1, tree.n_nodes), LAM))
    lq, tq = keys.arrays()
    eng.matrix_reset(B + 1, len(lq))
    p = eng.plan(slots, priors.uniform(B, B)[None, :], lq, tq, path=native.PATH_BATCHED32)
    p.execute()
    p.synchronize()
    res = p.results()
    p.profile()
    nts, _, _, launch = p.class_times()
    p.close()
    assert np.all(reduced_ok(res["lnl"][0][fin], ref["lnl"][fin]))
    want = 1 if ARRAYS[name] <= MAX_ARRAYS[B] else 0
    print(f"{name} B={B}: {ARRAYS[name]} arrays, classes {nts.tolist()} kernels {[int(l[0]) for l in launch]}",
          flush=True)
    assert [int(l[0]) for l in launch] == [want], (name, B, launch)
