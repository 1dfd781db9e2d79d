"""Two configurations at their full size against the oracle.

The plain base model (base_model.cpp:53-100): a seeded Yule tree of 12 taxa,
all 2,000 families simulated from a Poisson root (mean 5) at lambda = 0.002,
size bound 100.  Every family on every path against the oracle: 1e-9
relative on the FP64 paths, the reduced-path bar on BATCHED32 / BATCHED16,
and the total to 1e-9 / 1e-6.

Gamma rate variation (gamma_core.cpp:97-237): a seeded Yule tree of 34 taxa,
50,000 families, K = 4 categories at alpha = 0.5, lambda = 0.003, size
bound 100.  A seeded 500-family sample against the oracle on every path --
the family likelihoods and the category posteriors (p_k applied twice,
gamma_core.cpp:102) -- and all 50,000 families of BATCHED32 / BATCHED16
against BATCHED64, posteriors included.
"""
import numpy as np
import pytest

from cafe5_amd import native, priors, workloads as wl
from cafe5_amd.model import BaseModel, GammaModel, get_gamma
from bars import reduced_ok, TOTAL
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

PATHS = [native.PATH_PLAIN64, native.PATH_BATCHED64, native.PATH_BATCHED32, native.PATH_BATCHED16]


@pytest.fixture(scope="module")
def eng():
    e = native.Engine(0)
    yield e
    e.close()


@pytest.fixture(scope="module")
def base_case():
    w = wl.config_base(seed=1, F=2000)
    B = R = 100
    prior = priors.uniform(R, R)
    ref = Oracle.evaluate(w.tree, np.full(w.tree.n_nodes, 0.002), w.counts, B, R, prior, threads=16)
    return w, B, R, prior, ref


@pytest.mark.parametrize("path", PATHS)
def test_base_model_all_2000_families(eng, base_case, path):
    w, B, R, prior, ref = base_case
    m = BaseModel(eng, w.tree, w.counts, B, R, path=path)
    total = m.infer_family_likelihoods(prior, 0.002)
    got, want = m.results, ref["lnl"]
    assert np.all(np.isfinite(got)) and np.all(np.isfinite(want))
    d = np.abs(got - want)
    print(f"base 2000 {native.PATH_NAMES[path]}: max|dlnL|={d.max():.3e} total rel="
          f"{abs(total - ref['total']) / abs(ref['total']):.3e}", flush=True)
    if path in native.REDUCED_PATHS:
        assert np.all(reduced_ok(got, want)), d.max()
        assert abs(total - ref["total"]) <= TOTAL * abs(ref["total"])
    else:
        assert np.all(d <= 1e-9 * np.abs(want)), np.max(d / np.abs(want))
        assert abs(total - ref["total"]) <= 1e-9 * abs(ref["total"])


@pytest.fixture(scope="module")
def gamma_case(eng):
    w = wl.config_gamma(seed=3, F=50000, K=4, alpha=0.5, lam=0.003)
    B = R = int(w.bounds[0])
    K = 4
    freq, rate = get_gamma(K, 0.5)
    prior = priors.uniform(R, R)
    rng = np.random.Generator(np.random.PCG64(33))
    sample = np.sort(rng.choice(w.counts.shape[0], size=500, replace=False))
    ref = Oracle.evaluate(w.tree, np.full(w.tree.n_nodes, 0.003), w.counts[sample], B, R, prior, n_cat=K,
                          multipliers=rate, cat_probs=freq, threads=16)
    out = {}
    for path in (native.PATH_BATCHED64, native.PATH_BATCHED32, native.PATH_BATCHED16, native.PATH_PLAIN64):
        m = GammaModel(eng, w.tree, w.counts, B, R, freq, rate, path=path, alpha=0.5)
        total = m.infer_family_likelihoods(prior, 0.003)
        out[path] = (total, m.results["lnl"].copy(), m.results["post"].copy(), m.last["n_fallback"])
        print(f"gamma 50000 {native.PATH_NAMES[path]}: -lnL={total:.6f} recomputed={m.last['n_fallback']}",
              flush=True)
    return w, sample, ref, out


@pytest.mark.parametrize("path", PATHS)
def test_gamma_sample_against_oracle(gamma_case, path):
    w, sample, ref, out = gamma_case
    total, lnl, post, _ = out[path]
    got, want = lnl[sample], ref["lnl"]
    assert np.array_equal(np.isfinite(got), np.isfinite(want))
    d = np.abs(got - want)
    dp = np.abs(post[sample] - ref["post"])
    print(f"gamma sample {native.PATH_NAMES[path]}: max|dlnL|={d.max():.3e} max|dpost|={dp.max():.3e}",
          flush=True)
    if path in native.REDUCED_PATHS:
        assert np.all(reduced_ok(got, want)), d.max()
        assert dp.max() <= 1e-5
    else:
        assert np.all(d <= 1e-9 * np.abs(want)), np.max(d / np.abs(want))
        assert np.allclose(post[sample], ref["post"], rtol=1e-9, atol=1e-13)


@pytest.mark.parametrize("path", [native.PATH_BATCHED32, native.PATH_BATCHED16, native.PATH_PLAIN64])
def test_gamma_full_size_against_batched64(gamma_case, path):
    w, sample, ref, out = gamma_case
    t64, l64, p64, _ = out[native.PATH_BATCHED64]
    total, lnl, post, rc = out[path]
    assert np.isfinite(t64) == np.isfinite(total)
    if not np.isfinite(t64):   # a family with a zero-sum category: +inf (gamma_core.cpp:151-152)
        return
    assert np.array_equal(np.isfinite(lnl), np.isfinite(l64))
    fin = np.isfinite(l64)
    d = np.abs(lnl[fin] - l64[fin])
    print(f"gamma 50000 {native.PATH_NAMES[path]} vs batched64: max|dlnL|={d.max():.3e} total rel="
          f"{abs(total - t64) / abs(t64):.3e} max|dpost|={np.abs(post - p64).max():.3e} recomputed={rc}",
          flush=True)
    if path in native.REDUCED_PATHS:
        assert np.all(reduced_ok(lnl[fin], l64[fin])), d.max()
        assert abs(total - t64) <= TOTAL * abs(t64)
        assert np.abs(post - p64).max() <= 1e-5
    else:
        assert np.all(d <= 1e-9 * np.abs(l64[fin]))
        assert abs(total - t64) <= 1e-9 * abs(t64)
