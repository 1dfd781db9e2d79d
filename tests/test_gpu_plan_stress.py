"""Plan bookkeeping under load: tests/cpp/plan_stress drives the C ABI with a
plan of every path alive at once (graph capture and replay on the context's
side streams, interleaved executions, profiles, class times, recompute
counts, new keys, destruction in a different order each round), eager
evaluations with a tuning entry, and both reconstruction entry points; it
checks that the paths' totals agree and the two reconstructions are equal.

It runs twice: over libcafe_gpu.so, and over a copy whose host code is built
with AddressSanitizer (tests/cpp/_asan), so that a use after free, an
overflow or a double free in the planner's stream, event, graph and buffer
bookkeeping ends the run with a report."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVERS = {"plain": os.path.join(HERE, "cpp", "plan_stress"),
           "asan": os.path.join(HERE, "cpp", "_asan", "plan_stress")}


@pytest.mark.parametrize("kind", ["plain", "asan"])
def test_plan_stress(kind):
    driver = DRIVERS[kind]
    if not os.path.exists(driver):
        pytest.fail(f"{driver} is not built (python -m cafe5_amd.build --asan)")
    env = dict(os.environ)
    if kind == "asan":
        # host memory only; the GPU runtime holds its allocations to the end
        env["ASAN_OPTIONS"] = "detect_leaks=0:protect_shadow_gap=0"
    print(f"plan_stress {kind}", flush=True)
    res = subprocess.run([driver, "3"], capture_output=True, text=True, timeout=240, env=env)
    print(res.stdout[-3000:], res.stderr[-3000:])
    assert res.returncode == 0
    assert "plan_stress: ok" in res.stdout
    assert "AddressSanitizer" not in res.stderr
