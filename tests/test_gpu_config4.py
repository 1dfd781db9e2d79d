"""The error-model workload with local lambdas at its full size: a seeded Yule
tree of 60 taxa with 3 branch-lambda classes (0.001, 0.004, 0.01), 100,000
families at size bound 150, the symmetric +-1 leaf error model with
epsilon = 0.05 (probability.cpp:187-198 leaves, lambda.cpp:32-40 classes).

  * the split-FP16 path (BATCHED16) and the FP32 path against the oracle on a
    seeded sample of 600 families: the per-family bar of the reduced paths
    (tests/bars.py) and the sample's total to 1e-6 relative; the largest
    per-family error is printed;
  * BATCHED64 against the oracle on the same sample: 1e-9 relative;
  * all 100,000 families: BATCHED16 and BATCHED32 against BATCHED64 with the
    per-family bar, the totals to 1e-6 relative; the recompute lists hold
    the families the bounds did not certify (a small fraction).
"""
import numpy as np
import pytest

from cafe5_amd import native, priors, workloads as wl
from cafe5_amd.model import BaseModel
from bars import reduced_ok, TOTAL
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

B = R = 150
EPS = 0.05


@pytest.fixture(scope="module")
def case():
    w = wl.config_error_multi(seed=4, F=100000, bound=B)
    err = wl.error_table(EPS, B)
    prior = priors.uniform(R, R)
    rng = np.random.Generator(np.random.PCG64(41))
    sample = np.sort(rng.choice(w.counts.shape[0], size=600, replace=False))
    ref = Oracle.evaluate(w.tree, w.lambdas[w.node_class], w.counts[sample], B, R, prior, err=err, threads=16)
    eng = native.Engine(0)
    out = {}
    for path in (native.PATH_BATCHED64, native.PATH_BATCHED32, native.PATH_BATCHED16):
        m = BaseModel(eng, w.tree, w.counts, B, R, error_model=err, path=path)
        total = m.infer_family_likelihoods(prior, w.lambdas, node_class=w.node_class)
        out[path] = (total, m.results.copy(), m.last["n_fallback"])
        print(f"config4 {native.PATH_NAMES[path]}: -lnL={total:.6f} recomputed={m.last['n_fallback']}", flush=True)
    eng.close()
    return w, sample, ref, out


@pytest.mark.parametrize("path", [native.PATH_BATCHED64, native.PATH_BATCHED32, native.PATH_BATCHED16])
def test_sample_against_oracle(case, path):
    w, sample, ref, out = case
    total, lnl, _ = out[path]
    got = lnl[sample]
    want = ref["lnl"]
    assert np.all(np.isfinite(want)) and np.all(np.isfinite(got))
    d = np.abs(got - want)
    print(f"config4 {native.PATH_NAMES[path]} vs oracle on {len(sample)} families: max|dlnL|={d.max():.3e} "
          f"max rel={np.max(d / np.abs(want)):.3e} sample total rel="
          f"{abs(got.sum() - want.sum()) / abs(want.sum()):.3e}", flush=True)
    if path == native.PATH_BATCHED64:
        assert np.all(d <= 1e-9 * np.abs(want)), np.max(d / np.abs(want))
    else:
        assert np.all(reduced_ok(got, want)), d.max()
        assert abs(got.sum() - want.sum()) <= TOTAL * abs(want.sum())


@pytest.mark.parametrize("path", [native.PATH_BATCHED32, native.PATH_BATCHED16])
def test_full_size_against_batched64(case, path):
    w, sample, ref, out = case
    total64, lnl64, _ = out[native.PATH_BATCHED64]
    total, lnl, recomputed = out[path]
    assert np.array_equal(np.isfinite(lnl), np.isfinite(lnl64))
    fin = np.isfinite(lnl64)
    d = np.abs(lnl[fin] - lnl64[fin])
    print(f"config4 {native.PATH_NAMES[path]} vs batched64 on {fin.sum()} families: max|dlnL|={d.max():.3e} "
          f"total rel={abs(total - total64) / abs(total64):.3e} recomputed={recomputed}", flush=True)
    assert np.all(reduced_ok(lnl[fin], lnl64[fin])), d.max()
    assert abs(total - total64) <= TOTAL * abs(total64)
    # the bounds certify nearly every family: the FP64 recompute is a small tail
    assert recomputed <= 0.05 * w.counts.shape[0], recomputed
