"""CPU-side checks of the engine library: it is built, it exports every entry
point include/cafe_gpu.h declares, and the host-side key quantization matches
the oracle.  No compute call is made here (there is no GPU)."""
import ctypes
import os
import re

import numpy as np
import pytest

from cafe5_amd import native
from oracle_lib import ORC

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _ensure_built():
    if not os.path.exists(native.LIB_PATH):
        from cafe5_amd import build
        build.build_native()


def header_functions():
    text = open(os.path.join(ROOT, "include", "cafe_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cafe_gpu_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    _ensure_built()
    lib = ctypes.CDLL(native.LIB_PATH)
    names = header_functions()
    assert len(names) >= 14
    for name in names:
        assert hasattr(lib, name), name
    assert sorted(native.EXPORTED_SYMBOLS) == names


def test_binding_signatures_load():
    _ensure_built()
    lib = native.load_library()
    assert lib.cafe_gpu_last_error() is not None


def test_engine_without_gpu_fails_loudly():
    _ensure_built()
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(native.CafeGpuError):
        native.Engine(0)


def test_key_quantization_matches_oracle():
    rng = np.random.default_rng(5)
    lams = np.concatenate([np.exp(rng.uniform(np.log(1e-6), np.log(0.5), 2000)),
                           [1.57e-05, 0.0025, 0.3, 3.0, 0.1 * 3]])
    ts = np.concatenate([rng.uniform(0.0, 120.0, 2000), [1.001, 0.29, 68.7105, 2.5, 0.0004]])
    for lam in lams:
        assert native.key_lambda(float(lam)) == ORC.orc_key_lambda(float(lam))
    for t in ts:
        assert native.key_branch_length(float(t)) == ORC.orc_key_branch_length(float(t))


def test_key_table_dedups_quantized_keys():
    from cafe5_amd import workloads as wl
    tree = wl.parse_newick("((A:1.0004,B:1.0001):2,C:3.0009);")
    kt = native.KeyTable()
    slots = kt.node_slots(tree, np.full((2, tree.n_nodes), 0.01))
    # A and B quantize to t = 1.000 and share a matrix; both points share keys
    leaves = {tree.names[i]: i for i in range(tree.n_nodes) if tree.leaf_col[i] >= 0}
    assert slots[0, leaves["A"]] == slots[0, leaves["B"]]
    assert (slots[0] == slots[1]).all()
    assert len(kt.keys) == 3


def test_newick_without_outer_parentheses():
    # clade.cpp:334-341: a top-level comma makes a new root over the clade read
    # so far and the clades after it
    from cafe5_amd import workloads as wl

    def same(a, b):
        return (all(np.array_equal(getattr(a, k), getattr(b, k)) for k in ("parent", "leaf_col", "length"))
                and list(a.names) == list(b.names))

    assert same(wl.parse_newick("(A:1,B:1):2,C:3;"), wl.parse_newick("((A:1,B:1):2,C:3);"))
    assert same(wl.parse_newick("(A:1,B:1):2,C:3,D:4;"), wl.parse_newick("((A:1,B:1):2,C:3,D:4);"))
    assert same(wl.parse_newick("A:1,B:2"), wl.parse_newick("(A:1,B:2);"))
    t = wl.parse_newick("(A:1,B:1):2,C:3;")
    assert t.n_nodes == 5 and t.leaf_names() == ["A", "B", "C"] and t.names[t.root] == "ABC"
    with pytest.raises(ValueError):
        wl.parse_newick("(A:1,B:1))")   # text past the tree is an error, not dropped


def test_cpp_interface_built_and_fails_loudly_without_gpu():
    import subprocess
    import torch
    driver = os.path.join(ROOT, "tests", "cpp", "adapter_test")
    if not os.path.exists(driver) or not os.path.exists(os.path.join(ROOT, "cafe5_amd", "libcafe5_host.so")):
        from cafe5_amd import build
        build.build_native()
        build.build_host()
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present")
    res = subprocess.run([driver, "kats"], capture_output=True, text=True, timeout=60)
    assert res.returncode == 3 and "no HIP device" in res.stderr


def test_reference_file_readers_through_cpp_interface():
    """The reference's reader and writer tests (test.cpp:367-426,
    :1552-1608) and its error-model epsilon tests (:1483-1495, :1527-1549)
    through the C++ interface, and Newick text without the outer parentheses
    (clade.cpp:334-341); no device is used."""
    import subprocess
    driver = os.path.join(ROOT, "tests", "cpp", "adapter_test")
    res = subprocess.run([driver, "io"], capture_output=True, text=True, timeout=60)
    lines = res.stdout.splitlines()
    assert res.returncode == 0 and not [l for l in lines if l.startswith("FAIL")], res.stdout + res.stderr
    assert sum(l.startswith("PASS") for l in lines) == 11


def test_tuner_candidates_and_cache_format(tmp_path):
    from cafe5_amd import tune
    c13 = tune.candidates(native.PATH_BATCHED32, 13)
    assert all(k in (0, 2) and ntw * w >= 13 for k, ntw, w, _, _ in c13)
    # 16-family tiles: throughput and latency shapes; 32-family tiles
    assert {(ntw, w) for k, ntw, w, _, _ in c13 if k == 0} == {(4, 4), (2, 8), (1, 16)}
    assert {(ntw, w) for k, ntw, w, _, _ in c13 if k == 2} == {(2, 8), (1, 16)}
    assert {(ntw, w) for k, ntw, w, _, _ in tune.candidates(native.PATH_BATCHED32, 32) if k == 2} == {(2, 16)}
    assert not any(k == 2 for k, *_ in tune.candidates(native.PATH_BATCHED32, 33))   # up to 32 tiles
    for path in (native.PATH_BATCHED64, native.PATH_BATCHED32, native.PATH_BATCHED16):
        assert tune.candidates(path, 69) == []   # past 64 tiles: the plain kernel, no shape to tune
    assert tune.plan_key_of([69, 2, 13, 3]) == "2,3,13"   # the engine's plan key holds classes of 1..64 tiles
    assert not any(k == 2 for k, *_ in tune.candidates(native.PATH_BATCHED64, 13))   # FP32 only
    c2 = tune.candidates(native.PATH_BATCHED32, 2)
    assert any(k == 1 for k, *_ in c2) and any(k == 0 for k, *_ in c2)
    assert not any(k == 1 for k, *_ in tune.candidates(native.PATH_BATCHED64, 2))   # FP64: matrix cores only
    assert all(w <= 8 for _, _, w, _, _ in tune.candidates(native.PATH_BATCHED16, 32))
    out = tmp_path / "cache.txt"
    tune.write_cache(str(out), [(2, 13, 34, 0, 4, 4, 1, 16, 0.83, ""), (2, 7, 34, 2, 1, 8, 1, 4, 3.6, "2,7,13")],
                     "test")
    lines = [l for l in out.read_text().splitlines() if l and not l.startswith("#")]
    assert lines == ["2 13 34 0 4 4 1 16  0.8300", "2 7 34 2 1 8 1 4  3.6000  plan=2,7,13"]


def test_committed_tuning_cache_parses():
    path = os.path.join(ROOT, "cafe5_amd", "tune_cache.txt")
    if not os.path.exists(path):
        pytest.skip("no tile-shape cache committed")
    for line in open(path):
        f = line.split("#", 1)[0].split()
        if f:
            assert 8 <= len(f) <= 10, line
            path_, tiles, leaves, kernel, ntw, w, lds, gm = (int(x) for x in f[:8])
            assert 0 <= path_ <= 3 and tiles >= 1 and leaves >= 1 and kernel in (0, 1, 2) and lds in (0, 1)
            plans = [t[5:] for t in f[8:] if t.startswith("plan=")]
            for pl in plans:   # the plan's classes include the entry's own
                assert tiles in [int(x) for x in pl.split(",")], line
