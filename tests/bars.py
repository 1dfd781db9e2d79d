"""Per-family accuracy bars of the reduced-precision paths (FP32 MFMA and
split-FP16 MFMA, with certified FP64 recompute) against an FP64 result:

    |dlnL| <= ABS + REL_LNL |lnL|   and   |dlnL| / |lnL| <= REL

The total of every parameter point must agree to TOTAL relative (the target
the engine is built to).  The largest per-family error measured on the
1.28 M family-likelihoods of the uneven-size workload is 1.5e-6, so the bar
leaves about 3x headroom and catches any larger regression."""
import numpy as np

ABS = 5e-6
REL_LNL = 1e-7
REL = 1e-5
TOTAL = 1e-6


def reduced_ok(got, want) -> np.ndarray:
    """Elementwise: got within the per-family bar of want (finite entries)."""
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    err = np.abs(got - want)
    return (err <= ABS + REL_LNL * np.abs(want)) & (err <= REL * np.abs(want))
