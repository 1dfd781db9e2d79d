"""The tile-shape table (cafe_gpu_tuning_*): a cache file's choice replaces a
size class's built-in launch shape -- seen in the launch the plan records --
and the results stay those of the built-in shapes (the per-family bar of the
reduced paths); entries naming shapes that do not exist are ignored, and a
malformed file is rejected."""
import os

import numpy as np
import pytest

from cafe5_amd import native, priors, workloads as wl
from cafe5_amd.native import CafeGpuError
from bars import reduced_ok

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    w = wl.config_uneven(seed=2, F=3000, T=4)
    eng = native.Engine(0)
    eng.set_tree(w.tree)
    eng.set_families(w.counts, bound_B=w.bounds, bound_R=w.bounds)
    buckets = eng.buckets()
    keys = native.KeyTable()
    slots = keys.node_slots(w.tree, np.repeat(w.lambdas[:, None], w.tree.n_nodes, axis=1))
    lq, tq = keys.arrays()
    eng.matrix_reset(int(w.bounds.max()) + 1, len(lq))
    stride = max(r for _, r, _ in buckets)
    prior = np.zeros((len(buckets), stride))
    for i, (_, r, _) in enumerate(buckets):
        prior[i, :r] = priors.uniform(r, r)
    yield w, eng, slots, prior, lq, tq
    eng.close()


def _run(eng, slots, prior, lq, tq, path):
    p = eng.plan(slots, prior, lq, tq, path=path)
    p.execute()
    p.synchronize()
    res = p.results()
    p.profile()
    nts, _, _, launch = p.class_times()
    p.close()
    return res, {int(n): tuple(int(x) for x in l) for n, l in zip(nts, launch)}


@pytest.mark.parametrize("path", [native.PATH_BATCHED32, native.PATH_BATCHED16])
def test_cached_shapes_are_applied(setup, tmp_path, path):
    w, eng, slots, prior, lq, tq = setup
    eng.tuning_clear()
    base, shapes = _run(eng, slots, prior, lq, tq, path)
    assert shapes[2][0] == 1                      # built in: the 2-tile class runs one family per lane
    big = max(shapes)
    assert shapes[big][0] in (0, 3)               # the largest class on the matrix cores
    leaves = w.tree.n_leaves
    cache = tmp_path / "tune.txt"
    cache.write_text(
        "# test cache\n"
        f"{path} 2 {leaves} 0 2 1 1 4   0.5\n"            # the 2-tile class on the matrix cores, LDS, 4x grid
        f"{path} {big} {leaves} 0 4 4 0 1\n"             # the largest class with global partials, 1x grid
        f"{path} 3 {leaves} 0 7 7 1 4\n")                # no such instance: ignored
    assert eng.tuning_load(str(cache)) == 3
    tuned, tshapes = _run(eng, slots, prior, lq, tq, path)
    print(f"{native.PATH_NAMES[path]}: built-in {shapes}\n   tuned {tshapes}", flush=True)
    assert tshapes[2][:4] in ((0, 2, 1, 1), (3, 2, 1, 1))       # matrix cores, 2 tiles x 1 wave, LDS
    assert tshapes[2][5] == 16                                   # 16-family tiles
    assert tshapes[big][1:4] == (4, 4, 0)                        # 4 tiles x 4 waves, global partials
    assert tshapes[big][4] <= shapes[big][4]                     # a 1x grid
    assert tshapes[3] == shapes[3]                               # the unknown shape left the class alone
    fin = np.isfinite(base["lnl"])
    assert np.all(reduced_ok(tuned["lnl"][fin], base["lnl"][fin]))
    assert np.allclose(tuned["total"], base["total"], rtol=1e-6, atol=0)
    eng.tuning_clear()


def test_malformed_cache_is_rejected(setup, tmp_path):
    _, eng, *_ = setup
    bad = tmp_path / "bad.txt"
    for text in ("2 13 34 0 4\n", "2 13 34 0 4 4 1 16  0.8 plan=\n", "2 13 34 0 4 4 1 16 plan=2,x\n",
                 "2 13 34 0 4 4 1 16 plan=2,7\n", "2 13 34 0 4 4 1 16 plan=13,65\n"):
        bad.write_text(text)
        with pytest.raises(CafeGpuError) as ei:
            eng.tuning_load(str(bad))
        assert "malformed tuning entry" in str(ei.value), text
    with pytest.raises(CafeGpuError):
        eng.tuning_load(str(tmp_path / "missing.txt"))
    eng.tuning_clear()


# 32-family tiles (kernel 2): every instantiated shape on every class it
# covers.  Each family's column of an MFMA depends only on the matrix
# operand and that family's own column, in the same k order as with
# 16-family tiles, so the per-family results (after the same certification
# and FP64 recompute) are bit-identical to the built-in shapes', including
# the ragged last tile of every (bucket, point).
@pytest.mark.parametrize("shape", [(2, 4, 1), (2, 4, 0), (1, 8, 1), (1, 8, 0), (2, 8, 1), (2, 8, 0), (1, 16, 0),
                                   (2, 16, 0)])
def test_32_family_tiles_match_16_family_tiles(setup, shape):
    w, eng, slots, prior, lq, tq = setup
    path = native.PATH_BATCHED32
    eng.tuning_clear()
    base, shapes = _run(eng, slots, prior, lq, tq, path)
    ntw, wv, lds = shape
    leaves = w.tree.n_leaves
    tried = [nt for nt, l in shapes.items() if l[0] == 0 and nt <= 16 and ntw * wv >= nt]
    assert tried
    for nt in tried:
        eng.tuning_set(path, nt, leaves, 2, ntw, wv, bool(lds), 4 if lds else 2)
    tuned, tshapes = _run(eng, slots, prior, lq, tq, path)
    eng.tuning_clear()
    wide = [nt for nt in tried if tshapes[nt][5] == 32]
    print(f"shape {shape}: 32-family tiles on classes {wide} of {tried}; {tshapes}", flush=True)
    for nt in wide:
        assert tshapes[nt][:4] == (0, ntw, wv, lds)
    if not lds:
        assert wide == tried                 # global partials always fit
    assert wide
    assert np.array_equal(np.isfinite(tuned["lnl"]), np.isfinite(base["lnl"]))
    fin = np.isfinite(base["lnl"])
    assert np.array_equal(tuned["lnl"][fin], base["lnl"][fin]), np.abs(tuned["lnl"][fin] - base["lnl"][fin]).max()
    assert np.array_equal(tuned["total"], base["total"])
    assert tuned["n_fallback"] == base["n_fallback"]


def test_32_family_tiles_on_501_row_vectors():
    """The optimizer-scale shape (100 taxa, B = 500: 32-tile vectors, K = 8
    gamma categories) on 3,000 families: 32-family tiles of sixteen waves
    (the built-in choice at this size) against 16-family tiles of eight
    waves of four, bit for bit."""
    from cafe5_amd.model import ResidentGammaModel
    w = wl.config_optimizer(seed=5, F=3000, K=8, bound=500, T=200)
    B, K = 500, 8
    eng = native.Engine(0)
    out = []
    for entry in ((0, 4, 8, 0, 2), None):
        eng.tuning_clear()
        if entry:
            eng.tuning_set(native.PATH_BATCHED32, 32, w.tree.n_leaves, entry[0], *entry[1:3], bool(entry[3]),
                           entry[4])
        m = ResidentGammaModel(eng, w.tree, w.counts, B, B, K, priors.uniform(B, B), path=native.PATH_BATCHED32)
        totals, lnls = [], []
        for t in (3, 7):
            totals.append(m.infer_family_likelihoods(float(w.lambdas[t]), max(1.0, float(w.alphas[t]))))
            lnls.append(m.results["lnl"].copy())
        m.plan.profile()
        nts, _, _, launch = m.plan.class_times()
        m.close()
        out.append((totals, lnls, {int(n): tuple(int(x) for x in l) for n, l in zip(nts, launch)}))
    eng.tuning_clear()
    eng.close()
    (t16, l16, s16), (t32, l32, s32) = out
    print(f"16-family {s16}\n32-family {s32}\ntotals {t16} {t32}", flush=True)
    assert s16[32][:4] == (0, 4, 8, 0) and s16[32][5] == 16
    assert s32[32][:4] == (0, 2, 16, 0) and s32[32][5] == 32
    for a, b in zip(l16, l32):
        assert np.array_equal(np.isfinite(a), np.isfinite(b))
        fin = np.isfinite(a)
        assert np.array_equal(a[fin], b[fin])
    assert t16 == t32


# Band-limited k-loops: an FP32 interior step multiplies only the k-tiles
# inside the union of its wave's row tiles' bands (the blocks it skips are
# exactly 0 in the FP32 pool).  One output tile per wave skips by each tile's
# own band, the built-in shapes by unions of three or four; results that do
# not depend on the skipped set are bit-identical.
@pytest.mark.parametrize("shape", [(1, 16, 0), (1, 4, 1)])
def test_band_limited_k_loops_are_exact(setup, shape):
    w, eng, slots, prior, lq, tq = setup
    path = native.PATH_BATCHED32
    eng.tuning_clear()
    base, shapes = _run(eng, slots, prior, lq, tq, path)
    ntw, wv, lds = shape
    leaves = w.tree.n_leaves
    tried = [nt for nt, l in shapes.items() if l[0] == 0 and ntw * wv >= nt]
    assert tried
    for nt in tried:
        eng.tuning_set(path, nt, leaves, 0, ntw, wv, bool(lds), 2)
    tuned, tshapes = _run(eng, slots, prior, lq, tq, path)
    eng.tuning_clear()
    one = [nt for nt in tried if tshapes[nt][:4] == (0, ntw, wv, lds)]
    print(f"shape {shape}: one tile per wave on classes {one} of {tried}; {tshapes}", flush=True)
    assert one
    assert np.array_equal(np.isfinite(tuned["lnl"]), np.isfinite(base["lnl"]))
    fin = np.isfinite(base["lnl"])
    assert np.array_equal(tuned["lnl"][fin], base["lnl"][fin]), np.abs(tuned["lnl"][fin] - base["lnl"][fin]).max()
    assert np.array_equal(tuned["total"], base["total"])
    assert tuned["n_fallback"] == base["n_fallback"]


# Entries measured in a plan (plan=, the tiles of every size class) apply to
# plans with exactly those classes and are preferred there to entries for any
# plan; a plan of other classes ignores them.
def test_plan_keyed_entries(setup, tmp_path):
    w, eng, slots, prior, lq, tq = setup
    path = native.PATH_BATCHED32
    eng.tuning_clear()
    base, shapes = _run(eng, slots, prior, lq, tq, path)
    big = max(shapes)
    mine = ",".join(str(n) for n in sorted(shapes))
    leaves = w.tree.n_leaves
    cache = tmp_path / "plan.txt"
    cache.write_text(
        f"{path} {big} {leaves} 0 4 4 1 4\n"                         # any plan: LDS partials
        f"{path} {big} {leaves} 0 4 4 0 1  0.7  plan={mine}\n"       # this plan's classes: global partials
        f"{path} 2 {leaves} 0 2 1 1 4  0.5  plan=2,{big}\n")          # other classes: ignored
    assert eng.tuning_load(str(cache)) == 3
    tuned, tshapes = _run(eng, slots, prior, lq, tq, path)
    eng.tuning_clear()
    print(f"classes {mine}: built-in {shapes}\n   tuned {tshapes}", flush=True)
    assert tshapes[big][1:4] == (4, 4, 0)       # the plan-keyed entry, not the one for any plan
    assert tshapes[2] == shapes[2]              # a plan of other classes: built-in
    assert np.array_equal(tuned["total"], base["total"])   # where partials live changes no arithmetic
