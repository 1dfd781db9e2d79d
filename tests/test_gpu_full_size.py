"""The headline workload at its full size (seeded Yule tree of 34 taxa, 20,000
families in 8 size buckets up to B = 200, a 64-point lambda grid, 3,648 matrix
keys) through the captured plan the benchmark replays, checked by properties
that hold at any size:

  * BATCHED64 against PLAIN64 on every (point, family): 1e-9 relative;
  * BATCHED32 and BATCHED16 against PLAIN64 on every (point, family): the
    per-family bar of the reduced paths (tests/bars.py), and every point's
    total within 1e-6 relative;
  * PLAIN64 against the oracle on a seeded sample of (point, family) pairs
    drawn from every bucket: 1e-9 relative;
  * replays of one plan give bit-identical per-family results (the graph
    clears its counters and recompute lists itself), equal to the eager
    evaluation of the same inputs.
"""
import numpy as np
import pytest

from cafe5_amd import native, priors, workloads as wl
from bars import reduced_ok
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def full():
    w = wl.config_uneven(seed=2, F=20000, T=64)
    eng = native.Engine(0)
    eng.set_tree(w.tree)
    eng.set_families(w.counts, bound_B=w.bounds, bound_R=w.bounds)
    buckets = eng.buckets()
    keys = native.KeyTable()
    slots = keys.node_slots(w.tree, np.repeat(w.lambdas[:, None], w.tree.n_nodes, axis=1))
    lam_q, t_q = keys.arrays()
    eng.matrix_reset(int(w.bounds.max()) + 1, len(lam_q))
    stride = max(r for _, r, _ in buckets)
    prior = np.zeros((len(buckets), stride))
    for i, (_, r, _) in enumerate(buckets):
        prior[i, :r] = priors.uniform(r, r)
    out = {}
    for path in (native.PATH_PLAIN64, native.PATH_BATCHED64, native.PATH_BATCHED32, native.PATH_BATCHED16):
        plan = eng.plan(slots, prior, lam_q, t_q, path=path)
        plan.execute()
        plan.synchronize()
        out[path] = plan.results()
        if path == native.PATH_BATCHED32:
            plan.execute()
            plan.synchronize()
            out["replay"] = plan.results()
        plan.close()
    # eager evaluation of the same inputs (matrices built by the same algorithm)
    eng.matrix_build(lam_q, t_q, 0, native.BUILD_RECURRENCE)
    out["eager"] = eng.evaluate(slots, prior, path=native.PATH_BATCHED32)
    yield w, buckets, prior, out
    eng.close()


def test_batched64_matches_plain64(full):
    _, _, _, out = full
    a, b = out[native.PATH_BATCHED64]["lnl"], out[native.PATH_PLAIN64]["lnl"]
    assert a.shape == (64, 20000)
    assert np.array_equal(np.isfinite(a), np.isfinite(b))
    fin = np.isfinite(b)
    assert np.max(np.abs(a[fin] - b[fin]) / np.abs(b[fin])) <= 1e-9


@pytest.mark.parametrize("path", [native.PATH_BATCHED32, native.PATH_BATCHED16])
def test_reduced_paths_match_plain64(full, path):
    _, _, _, out = full
    a, b = out[path]["lnl"], out[native.PATH_PLAIN64]["lnl"]
    assert np.array_equal(np.isfinite(a), np.isfinite(b))
    fin = np.isfinite(b)
    err = np.abs(a[fin] - b[fin])
    ta, tb = out[path]["total"], out[native.PATH_PLAIN64]["total"]
    print(f"{native.PATH_NAMES[path]}: max|dlnL|={err.max():.3e} max total rel="
          f"{np.max(np.abs(ta - tb) / np.abs(tb)):.3e} recomputed={out[path]['n_fallback']}", flush=True)
    assert np.all(reduced_ok(a[fin], b[fin])), err.max()
    assert np.all(np.abs(ta - tb) <= 1e-6 * np.abs(tb)), np.max(np.abs(ta - tb) / np.abs(tb))
    # the families the bounds could not certify were recomputed in FP64
    assert out[path]["n_fallback"] > 0


def test_plain64_matches_oracle_on_a_sample(full):
    w, buckets, prior, out = full
    lnl = out[native.PATH_PLAIN64]["lnl"]
    rng = np.random.default_rng(2024)
    for bi, (B, R, _) in enumerate(buckets):
        fams = np.flatnonzero(w.bounds == B)
        for f in rng.choice(fams, size=min(4, len(fams)), replace=False):
            t = int(rng.integers(0, len(w.lambdas)))
            ref = Oracle.evaluate(w.tree, np.full(w.tree.n_nodes, w.lambdas[t]), w.counts[f:f + 1], B, R,
                                  prior[bi, :R], threads=8)["lnl"][0]
            assert abs(lnl[t, f] - ref) <= 1e-9 * abs(ref), (B, f, t, lnl[t, f], ref)


def test_replays_and_eager_agree(full):
    _, _, _, out = full
    first, again, eager = out[native.PATH_BATCHED32], out["replay"], out["eager"]
    assert np.array_equal(first["lnl"], again["lnl"])
    assert np.array_equal(first["total"], again["total"])
    assert first["n_fallback"] == again["n_fallback"]
    assert np.array_equal(first["lnl"], eager["lnl"])
    assert np.allclose(first["total"], eager["total"], rtol=1e-12, atol=0)
