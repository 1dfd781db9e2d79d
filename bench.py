#!/usr/bin/env python3
"""Family-likelihood throughput of the MI355X engine on the uneven-size
lambda-grid workload.

Workload (one "step" = one pass of the hot path over the whole batch):
  seeded Yule tree with 34 taxa (height 100); 20,000 families with
  heavy-tailed root sizes (85% <= 10, tail to 150) simulated by birth-death;
  per-family size bounds 20..200 in 8 buckets; a 64-point lambda grid
  1e-4 .. 1e-2.  A step builds every (lambda x branch) transition matrix
  (3,648 matrices of 201 x 201), prunes all 20,000 x 64 family likelihoods,
  reduces them and returns the 64 totals -- replayed as one HIP graph with all
  inputs resident on the GPU.

The headline path is BATCHED32 (FP32 MFMA contractions, per-family rescaling,
FP64 recompute of families the flush bound cannot certify).  The plain FP64
drop-in (per-family matrix-vector products) and the batched FP64 MFMA path are
timed on the same inputs, the FP32 results are compared with FP64 ones, and the
compiled reference (oracle/_ref, or the oracle port) is timed on a bounded
sample on the host cores.

    python bench.py [--steps K] [--warmup W]      # one JSON line on stdout
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from cafe5_amd import native, priors, workloads as wl  # noqa: E402

METRIC = "family likelihoods/sec, batched FP32-MFMA vs plain FP64 on GPU at |ΔlnL|≤1e-6"
PEAK_FP32_MFMA = 157.3e12     # MI355X dense FP32 matrix, FLOP/s
PEAK_HBM = 8.0e12             # MI355X HBM3E, B/s
PEAK_FP32_VALU_UNPACKED = 78.6e12   # FP32 vector FMAs without packed FP32 (half the 157.3 figure)
PEAK_FP64 = 78.6e12           # MI355X FP64 (vector and matrix), FLOP/s
KERNEL_KIND = {0: "k_prune_mfma", 1: "k_prune_lane", 2: "k_prune_plain"}


def log(msg: str):
    print(msg, file=sys.stderr, flush=True)


def load_tuning(eng: native.Engine, path: str):
    """Load the tile-shape cache into the engine; what was loaded, for the
    output line."""
    if not path or path == "none" or not os.path.exists(path):
        return {"file": None, "entries": 0, "note": "built-in shapes"}
    n = eng.tuning_load(path)
    return {"file": os.path.relpath(path, ROOT), "entries": n}


def tree_flop_factors(tree: wl.Tree):
    """(n1, n2): interior non-root nodes whose parent is interior / the root.
    A family likelihood at bound B (L = B+1 entries, R root sizes) costs
    2 L (n1 L + n2 R) flops of contractions; leaf branches are column gathers."""
    n1 = n2 = 0
    for u in range(tree.n_nodes):
        if tree.leaf_col[u] >= 0 or u == tree.root:
            continue
        if tree.parent[u] == tree.root:
            n2 += 1
        else:
            n1 += 1
    return n1, n2


def setup(eng: native.Engine, w: wl.Workload):
    eng.set_tree(w.tree)
    eng.set_families(w.counts, bound_B=w.bounds, bound_R=w.bounds)
    buckets = eng.buckets()
    keys = native.KeyTable()
    node_lams = np.repeat(w.lambdas[:, None], w.tree.n_nodes, axis=1)
    slots = keys.node_slots(w.tree, node_lams)
    lam_q, t_q = keys.arrays()
    N = int(w.bounds.max()) + 1
    eng.matrix_reset(N, len(lam_q))
    stride = max(r for _, r, _ in buckets)
    prior = np.zeros((len(buckets), stride))
    for i, (_, r, _) in enumerate(buckets):
        prior[i, :r] = priors.uniform(r, r)
    return buckets, slots, lam_q, t_q, prior, N


def time_plan(plan: native.Plan, steps: int, warmup: int, profile_steps: int):
    """ms per step over `steps` graph replays (each waits for its totals, as
    the optimizer does), then the phase times of `profile_steps` executions
    with HIP events between the phases (same kernels, same stream)."""
    for _ in range(warmup):
        plan.execute()
        plan.synchronize()
    plan.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.execute()
        plan.synchronize()
    t1 = time.perf_counter()
    phase = np.zeros(4)
    graph_ok = False
    cls = None
    for _ in range(profile_steps):
        ms, graph_ok = plan.profile()
        phase += ms
        nt, tiles, cms, launch = plan.class_times()
        cls = (nt, tiles, cms if cls is None else cls[2] + cms, launch)
    n = max(1, profile_steps)
    if cls is not None:
        cls = (cls[0], cls[1], cls[2] / n, cls[3])
    return (t1 - t0) * 1e3 / steps, phase / n, graph_ok, cls


def read_traffic(nt: int, kernel: str):
    """HBM bytes per launch of the FP32 main-pass kernel of the nt-tile size
    class, from the committed rocprofv3 counter summary (FETCH_SIZE and
    WRITE_SIZE of that kernel, each scaled by what the counter reads for a
    device copy of known size); None when the summary does not cover it or
    was taken on another kernel (name prefix `kernel`)."""
    path = os.path.join(ROOT, "profiles", "prune_fp32_uneven34_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        t = json.load(fh)
    k = t.get("kernels", {}).get(str(nt))
    if not k or not k["kernel"].startswith(kernel):
        return None
    return {"bytes": k["fetch_kib"] * 1024 * t["fetch_scale"] + k["write_kib"] * 1024 * t["write_scale"],
            "source": f"profiles/prune_fp32_uneven34_traffic.json ({t['command']})"}


def kernels_only(w: wl.Workload, steps: int, paths=(native.PATH_BATCHED32,)):
    """Each path's plan executed `steps` times with its size classes launched
    one after another (plan.profile), then two device copies of 256 MiB (the
    known-size reference for counter scaling): the program that rocprofv3
    counter passes run."""
    eng = native.Engine(0)
    buckets, slots, lam_q, t_q, prior, N = setup(eng, w)
    for path in paths:
        p = eng.plan(slots, prior, lam_q, t_q, path=path, use_graph=False)
        for _ in range(steps):
            p.profile()   # the size classes one after another: each launch alone on the device
        p.close()
        log(f"kernels-only: {steps} profiled {native.PATH_NAMES[path]} executions")
    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = ctypes.c_size_t(256 << 20)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), nbytes) != 0 or hip.hipMalloc(ctypes.byref(b), nbytes) != 0:
        raise RuntimeError("hipMalloc failed")
    for _ in range(2):
        if hip.hipMemcpy(b, a, nbytes, 3) != 0:   # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    eng.close()
    log("kernels-only: 2 x 256 MiB device copies")


# ----------------------------------------------------------------------------
# Pupko joint reconstruction (gene_family_reconstructor.cpp:30-189) of every
# family of the workload at one lambda, matrices of the largest bound
# ----------------------------------------------------------------------------
RECON_LAMBDA = 0.003


def time_reconstruct(eng: native.Engine, w: wl.Workload, reps: int = 3):
    B = R = int(w.bounds.max())
    eng.set_families(w.counts, B, R)
    keys = native.KeyTable()
    slots = keys.node_slots(w.tree, np.full((1, w.tree.n_nodes), RECON_LAMBDA))
    lq, tq = keys.arrays()
    eng.matrix_reset(B + 1, len(lq))
    eng.matrix_build(lq, tq, 0, native.BUILD_RECURRENCE)
    prior = priors.uniform(R, min(B, R) + 1)
    states = eng.reconstruct(slots, prior, B, R)   # first call sizes the device buffers
    t0 = time.perf_counter()
    for _ in range(reps):
        states = eng.reconstruct(slots, prior, B, R)
    ms = (time.perf_counter() - t0) * 1e3 / reps
    F = w.counts.shape[0]
    return {"families": F, "B": B, "R": R, "lambda": RECON_LAMBDA, "ms_per_call": ms,
            "families_per_s": F / (ms * 1e-3),
            "timed": "cafe_gpu_reconstruct over all families, counts and states as host arrays (PCIe included), "
                     "matrices built beforehand"}, states


def _cpu_reconstruct(so: str, w: wl.Workload, threads: int, gpu_states: np.ndarray, n_sample: int = 512):
    lib = ctypes.CDLL(so)
    lib.ref_init.restype = None
    lib.ref_init()
    lib.ref_set_threads.restype = ctypes.c_int
    lib.ref_set_threads.argtypes = [ctypes.c_int]
    used = lib.ref_set_threads(threads)
    lib.ref_reconstruct_many.restype = ctypes.c_int
    lib.ref_reconstruct_many.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_int), ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_int)]
    B = R = int(w.bounds.max())
    idx = _sample(w, n_sample, seed=13)
    ip = ctypes.POINTER(ctypes.c_int)
    t, outs = [], []
    for n in (min(64, len(idx)), len(idx)):
        c = np.ascontiguousarray(w.counts[idx[:n]], dtype=np.int32)
        out = np.zeros((n, w.tree.n_nodes), dtype=np.int32)
        t0 = time.perf_counter()
        rc = lib.ref_reconstruct_many(w.tree.newick().encode(), " ".join(w.tree.leaf_names()).encode(),
                                      " ".join(w.tree.names).encode(), n, c.ctypes.data_as(ip), RECON_LAMBDA, B, R,
                                      out.ctypes.data_as(ip))
        t.append(time.perf_counter() - t0)
        if rc != 0:
            raise RuntimeError("ref_reconstruct_many failed")
        outs.append(out)
    F = w.counts.shape[0]
    est = _extrapolate(t[0], t[1], min(64, len(idx)), len(idx), F)
    differ = int(np.sum(np.any(outs[1] != gpu_states[idx], axis=1)))
    return {"families_per_s": F / est, "cores": int(used), "kind": "reference",
            "sample": (f"reference pupko_reconstructor::reconstruct_gene_family (oracle/_ref, OpenMP over families, "
                       f"one matrix cache) on seeded {min(64, len(idx))}- and {len(idx)}-family samples, matrix build "
                       f"and per-family cost extrapolated to all {F} families ({sum(t):.1f} s of CPU wall time)"),
            "families_whose_states_differ_from_gpu": differ}


# ----------------------------------------------------------------------------
# CPU baseline (checker libraries under oracle/, loaded only here)
# ----------------------------------------------------------------------------

def cpu_baseline(w: wl.Workload, n_lambda: int, n_sample: int, threads: int):
    ref_so = os.path.join(ROOT, "oracle", "_ref", "libcaferef.so")
    if os.path.exists(ref_so):
        return _cpu_reference(ref_so, w, n_lambda, n_sample, threads)
    return _cpu_port(w, n_lambda, n_sample, threads)


def _sample(w: wl.Workload, n_sample: int, seed: int = 11):
    rng = np.random.Generator(np.random.PCG64(seed))
    F = w.counts.shape[0]
    idx = np.sort(rng.choice(F, size=min(n_sample, F), replace=False))
    return idx


def _extrapolate(times_tiny, times_sample, n_tiny, n_sample, n_full):
    """Matrix build + per-family cost, amortized over the full family count."""
    if n_sample <= n_tiny:
        return times_sample * (n_full / max(n_sample, 1))
    per_fam = max(0.0, (times_sample - times_tiny) / (n_sample - n_tiny))
    return times_tiny + per_fam * (n_full - n_tiny)


def _cpu_reference(so, w, n_lambda, n_sample, threads):
    """The compiled reference (base_model::infer_family_likelihoods with its
    OpenMP loops) on the host: per size bucket a session on 8 families and one
    on the bucket's share of a seeded n_sample-family subsample, each
    evaluated at n_lambda grid lambdas with one warm-up and the median of 3;
    the per-evaluation matrix build and the per-family cost are separated by
    the two session sizes and extrapolated to every family.  The sessions'
    construction (the duplicate-family scan of base_model.cpp:27-51) is timed
    apart.  Run on `threads` threads and on one thread (an eighth of the
    sample, two lambdas)."""
    lib = ctypes.CDLL(so)
    lib.ref_init.restype = None
    lib.ref_init()
    lib.ref_set_threads.restype = ctypes.c_int
    lib.ref_set_threads.argtypes = [ctypes.c_int]
    lib.ref_base_session_create.restype = ctypes.c_void_p
    lib.ref_base_session_create.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int,
                                            ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int, ctypes.c_double]
    lib.ref_base_session_eval.restype = ctypes.c_double
    lib.ref_base_session_eval.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p]
    lib.ref_base_session_free.argtypes = [ctypes.c_void_p]
    newick = w.tree.newick().encode()
    names = " ".join(w.tree.leaf_names()).encode()
    F = w.counts.shape[0]

    def run(nt, n_fam, n_lam):
        used = lib.ref_set_threads(nt)
        idx = _sample(w, n_fam)
        lams = w.lambdas[np.linspace(0, len(w.lambdas) - 1, n_lam).round().astype(int)]
        est_total = wall = scan = 0.0
        for B in sorted(set(w.bounds.tolist())):
            members = np.flatnonzero(w.bounds == B)
            full = len(members)
            sel = idx[w.bounds[idx] == B]
            if len(sel) < 24:   # enough families to separate the build from the per-family cost
                sel = members[:min(24, full)]
            n_tiny = min(8, max(1, len(sel) // 3))
            sessions = []
            for part in (sel[:n_tiny], sel):
                c = np.ascontiguousarray(w.counts[part], dtype=np.int32)
                t0 = time.perf_counter()
                h = lib.ref_base_session_create(newick, names, len(part),
                                                c.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), B, B, 0, B, 0.0)
                scan += time.perf_counter() - t0
                sessions.append((h, len(part)))
            for lam in lams:
                med = []
                for h, _ in sessions:
                    ts = []
                    for _ in range(4):
                        t0 = time.perf_counter()
                        lib.ref_base_session_eval(h, float(lam), None)
                        ts.append(time.perf_counter() - t0)
                    wall += sum(ts)
                    med.append(float(np.median(ts[1:])))
                est_total += _extrapolate(med[0], med[1], sessions[0][1], sessions[1][1], full)
            for h, _ in sessions:
                lib.ref_base_session_free(h)
        return F * len(lams) / est_total, int(used), wall, scan, len(idx), len(lams)

    value, used, wall, scan, n_s, n_l = run(threads, n_sample, n_lambda)
    v1, _, wall1, _, n_s1, n_l1 = run(1, max(64, n_sample // 8), 2)
    return {"value": value, "unit": "family likelihoods/s", "cores": used, "kind": "reference",
            "one_thread_value": v1, "cpu_model": _cpu_model(),
            "duplicate_scan_s": scan,
            "sample": (f"reference base_model::infer_family_likelihoods (oracle/_ref, -O3 -march=x86-64-v3, OpenMP "
                       f"over families) at {n_l} of the 64 grid lambdas on a seeded {n_s}-family subsample plus "
                       f"8-family sessions per bucket, one warm-up and the median of 3 per evaluation; matrix "
                       f"build and per-family cost extrapolated to all {F} families ({wall:.1f} s of CPU wall "
                       f"time); one thread: {n_l1} lambdas on {n_s1} families ({wall1:.1f} s); the sessions' "
                       f"duplicate scan (base_model.cpp:27-51) is reported apart")}


def _cpu_port(w, n_lambda, n_sample, threads):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle  # checker only
    idx = _sample(w, n_sample)
    lams = w.lambdas[np.linspace(0, len(w.lambdas) - 1, n_lambda).round().astype(int)]
    est_total = wall = 0.0
    for B in sorted(set(w.bounds.tolist())):
        full = int((w.bounds == B).sum())
        sel = idx[w.bounds[idx] == B]
        if len(sel) == 0:
            sel = np.flatnonzero(w.bounds == B)[:1]
        n_tiny = min(8, len(sel))
        prior = priors.uniform(B, B)
        for lam in lams:
            t = []
            for part in (sel[:n_tiny], sel):
                t0 = time.perf_counter()
                Oracle.evaluate(w.tree, np.full(w.tree.n_nodes, lam), w.counts[part], B, B, prior, threads=threads)
                t.append(time.perf_counter() - t0)
            wall += sum(t)
            est_total += _extrapolate(t[0], t[1], n_tiny, len(sel), full)
    F = w.counts.shape[0]
    return {"value": F * len(lams) / est_total, "unit": "family likelihoods/s", "cores": threads, "kind": "port",
            "sample": f"FP64 oracle port at {len(lams)} lambdas, {len(idx)}-family subsample, extrapolated "
                      f"({wall:.1f} s)"}


# ----------------------------------------------------------------------------

# ----------------------------------------------------------------------------
# Error model + local lambdas (60 taxa, 3 lambda classes, eps 0.05, B = 150,
# 100,000 families): split-FP16 MFMA against FP32, FP64 MFMA and PLAIN64
# ----------------------------------------------------------------------------
PEAK_FP16_MFMA = 2.5e15     # MI355X dense FP16 matrix, FLOP/s
KERNEL_KIND[3] = "k_prune_f16"


def _path_runs(eng, slots, prior, lam_q, t_q, paths, steps, warmup, fp64_steps, model=0, cat_probs=None):
    out = {}
    for path in paths:
        p = eng.plan(slots, prior, lam_q, t_q, path=path, model=model, cat_probs=cat_probs, use_graph=True)
        reduced = path in native.REDUCED_PATHS
        n = steps if reduced else fp64_steps
        ms, ph, graph_ok, cls = time_plan(p, n, warmup if reduced else 1, n if reduced else 1)
        r = p.results()
        rc = p.recompute_counts()
        out[path] = {"ms": ms, "phases": ph, "graph": graph_ok, "classes": cls, "res": r, "recomputed": rc}
        log(f"{native.PATH_NAMES[path]}: {ms:.3f} ms/step phases={np.round(ph, 3).tolist()} recomputed={rc}")
        p.close()
    return out


def _accuracy(runs, test_path, ref_path):
    a = runs[test_path]["res"]["lnl"]
    b = runs[ref_path]["res"]["lnl"]
    fin = np.isfinite(b)
    d = np.abs(a[fin] - b[fin])
    tr = np.abs(runs[test_path]["res"]["total"] - runs[ref_path]["res"]["total"]) / np.abs(runs[ref_path]["res"]["total"])
    return {"reference_for_comparison": f"{native.PATH_NAMES[ref_path]} GPU path (matches the FP64 oracle to 1e-9 "
                                        f"in tests)",
            "max_abs_dlnL_per_family": float(d.max()) if d.size else 0.0,
            "max_rel_dlnL_per_family": float((d / np.abs(b[fin])).max()) if d.size else 0.0,
            "max_rel_dlnL_total": float(np.max(tr)),
            "nonfinite_families_match": bool(np.array_equal(np.isfinite(a), fin)),
            "families_recomputed_fp64": int(runs[test_path]["recomputed"][0]),
            "families_recomputed_plain64": int(runs[test_path]["recomputed"][1])}


def _class_roofline(run, flops, peak, unit_desc):
    nt_cls, tiles_cls, ms_cls, launch_cls = run["classes"]
    dom = int(np.argmax(ms_cls))
    ms = float(ms_cls[dom])
    dl = launch_cls[dom]
    achieved = flops / (ms * 1e-3)
    return {"bound": "mfma", "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
            "frac": achieved / peak, "traffic": None,
            "kernel": f"{KERNEL_KIND[int(dl[0])]}<{int(dl[1])},{int(dl[2])},{'lds' if dl[3] else 'global'}> main-pass "
                      f"launch of the {int(nt_cls[dom])}-tile size class ({int(tiles_cls[dom])} tiles of "
                      f"{int(dl[5])} families, {int(dl[4])} workgroups)",
            "algorithmic_flops_per_launch": flops, "avg_ms": ms, "flops_per_family_likelihood": unit_desc,
            "executed": "the FP32 matrix-core kernel multiplies only the k-tiles inside each step's nonzero band "
                        "(matrix blocks that are exactly 0 in FP32 are skipped), so it executes fewer matrix-core "
                        "flops than the dense algorithmic count above; the counters under profiles/ give the "
                        "executed rate"}


def _ref_cpu_base(w, B, R, err, node_lambda_classes, n_sample, threads, gamma=None):
    """The compiled reference's base_model::infer_family_likelihoods (error
    model, lambda tree) -- or, with gamma = (category probabilities,
    multipliers), gamma_model's -- on seeded subsamples of 32 and n_sample
    families, one warm-up and the median of 3 each: the per-evaluation matrix
    build and the per-family cost, separated by the two sizes and
    extrapolated to every family, on `threads` OpenMP threads and on one
    thread (a quarter of the sample).  Each base-model call also runs the
    constructor's duplicate scan (base_model.cpp:27-51), O(n^2) in the
    subsample and included."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Ref  # checker only
    if not Ref.available():
        return None
    Ref.lib.ref_set_threads.restype = ctypes.c_int
    Ref.lib.ref_set_threads.argtypes = [ctypes.c_int]
    idx = _sample(w, n_sample, seed=17)
    F = w.counts.shape[0]
    res = {}
    for nt, n in ((threads, len(idx)), (1, max(64, len(idx) // 4))):
        used = Ref.lib.ref_set_threads(nt)
        med = []
        for m in (32, n):
            c = np.ascontiguousarray(w.counts[idx[:m]])
            times = []
            for rep in range(4):
                t0 = time.perf_counter()
                if gamma is None:
                    Ref.base_model(w.tree, c, w.lambdas, B, R, err=err, node_class=node_lambda_classes)
                else:
                    Ref.gamma_model(w.tree, c, w.lambdas, B, R, gamma[0], gamma[1], err=err,
                                    node_class=node_lambda_classes)
                times.append(time.perf_counter() - t0)
            med.append(float(np.median(times[1:])))
        res[nt] = (used, n, _extrapolate(med[0], med[1], 32, n, F), med)
    used, n, est, med = res[threads]
    _, n1, est1, med1 = res[1]
    return {"value": F / est, "unit": "family likelihoods/s", "cores": int(used), "kind": "reference",
            "one_thread_value": F / est1, "cpu_model": _cpu_model(),
            "sample": (f"reference {'gamma_model' if gamma is not None else 'base_model'}::infer_family_likelihoods "
                       f"(oracle/_ref, -O3 -march=x86-64-v3, OpenMP) "
                       f"{'with the error model and the lambda tree ' if err is not None else ''}"
                       f"on seeded 32- and {n}-family subsamples "
                       f"({med[0]:.2f} s and {med[1]:.2f} s; one thread: 32 and {n1} families, {med1[0]:.2f} s and "
                       f"{med1[1]:.2f} s), one warm-up and the median of 3; matrix build and per-family cost "
                       f"extrapolated to all {F} families")}


def _cpu_model() -> str:
    try:
        res = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10)
        for line in res.stdout.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _host_threads() -> int:
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    return env or aff


def main_config4(args):
    F = 100000 if args.families == 20000 else args.families
    B = R = 150
    w = wl.config_error_multi(seed=4, F=F, bound=B)
    err = wl.error_table(0.05, B)
    log(f"workload: {w.name} F={F} taxa={w.tree.n_leaves} B={B} classes={w.lambdas.tolist()} eps=0.05")
    eng = native.Engine(0)
    tuning = load_tuning(eng, args.tuning)
    eng.set_tree(w.tree)
    eng.set_families(w.counts, B, R)
    eng.set_error_model(err)
    keys = native.KeyTable()
    slots = keys.node_slots(w.tree, w.lambdas[w.node_class][None, :])
    lam_q, t_q = keys.arrays()
    eng.matrix_reset(B + 1, len(lam_q))
    prior = priors.uniform(R, R)[None, :]
    paths = (native.PATH_BATCHED16, native.PATH_BATCHED32, native.PATH_BATCHED64, native.PATH_PLAIN64)
    runs = _path_runs(eng, slots, prior, lam_q, t_q, paths, args.steps, args.warmup, args.fp64_steps)
    n1, n2 = tree_flop_factors(w.tree)
    L = B + 1
    flops = F * 2.0 * L * (n1 * L + n2 * R)
    r16 = runs[native.PATH_BATCHED16]
    out = {
        "metric": "family likelihoods/sec, batched split-FP16 MFMA (error model + lambda classes) vs plain FP64 "
                  "on GPU at |ΔlnL_total|≤1e-6",
        "value": F / (r16["ms"] * 1e-3), "unit": "family likelihoods/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r16["ms"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f16x2", "data": "synthetic",
        "config": {"workload": f"errmodel-60taxa: seeded Yule tree (60 taxa, height 100), 3 branch-lambda classes "
                               f"(0.001, 0.004, 0.01), {F} birth-death families, size bound 150, leaf error model "
                               f"eps=0.05, {len(lam_q)} matrices of {B + 1}x{B + 1} per step",
                   "path": "batched16 (split-FP16 MFMA + per-entry range bound + FP64 recompute)",
                   "graph": bool(r16["graph"]), "families": F, "trials": 1},
        "paths": {native.PATH_NAMES[p]: {"value": F / (runs[p]["ms"] * 1e-3), "ms_per_step": runs[p]["ms"],
                                         "phases_ms": np.round(runs[p]["phases"], 4).tolist()} for p in paths},
        "speedup_vs_plain64": runs[native.PATH_PLAIN64]["ms"] / r16["ms"],
        "speedup_vs_batched32": runs[native.PATH_BATCHED32]["ms"] / r16["ms"],
        "accuracy": _accuracy(runs, native.PATH_BATCHED16, native.PATH_BATCHED64),
        "accuracy_batched32": _accuracy(runs, native.PATH_BATCHED32, native.PATH_BATCHED64),
        "roofline": _class_roofline(r16, flops, PEAK_FP16_MFMA,
                                    "2 L (n1 L + n2 R), n1=%d n2=%d, L=%d (algorithmic; the split operands issue "
                                    "3 value MFMAs and 1 bound MFMA per product on 16-padded sizes)" % (n1, n2, L)),
        "roofline_batched32": _class_roofline(runs[native.PATH_BATCHED32], flops, PEAK_FP32_MFMA,
                                              "2 L (n1 L + n2 R)"),
        "tuning": tuning,
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = _ref_cpu_base(w, B, R, err, w.node_class, min(2000, F), _host_threads())
        log(f"cpu baseline: {out['cpu_baseline']}")
    eng.close()
    print(json.dumps(out))


def _ref_cpu_prune(w, B, R, lam, n_sample, n_one, threads, seed=19):
    """The compiled reference's two per-evaluation costs on the workload's
    tree at bounds (B, R) and one lambda, timed apart in a pruning session:
    its matrix_cache build (precalculate_matrices: one key per branch length)
    and inference_prune over a seeded n_sample-family subsample in an OpenMP
    loop over families (one warm-up and the median of 3) on `threads`
    threads, and over n_one families on one thread.  Returns (build seconds,
    seconds per family, seconds per family on one thread, threads used)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Ref  # checker only
    if not Ref.available():
        return None
    lib = Ref.lib
    lib.ref_set_threads.restype = ctypes.c_int
    lib.ref_set_threads.argtypes = [ctypes.c_int]
    lib.ref_prune_session_create.restype = ctypes.c_void_p
    lib.ref_prune_session_create.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int,
                                             ctypes.POINTER(ctypes.c_int), ctypes.c_double, ctypes.c_double,
                                             ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    lib.ref_prune_session_run.restype = ctypes.c_double
    lib.ref_prune_session_run.argtypes = [ctypes.c_void_p]
    lib.ref_prune_session_free.argtypes = [ctypes.c_void_p]
    idx = _sample(w, n_sample, seed=seed)
    newick = w.tree.newick().encode()
    names = " ".join(w.tree.leaf_names()).encode()
    out = []
    for nt, n in ((threads, n_sample), (1, n_one)):
        used = lib.ref_set_threads(nt)
        c = np.ascontiguousarray(w.counts[idx[:n]], dtype=np.int32)
        build = ctypes.c_double(0.0)
        h = lib.ref_prune_session_create(newick, names, n, c.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                         float(lam), 1.0, B, R, ctypes.byref(build))
        if not h:
            return None
        ts = [lib.ref_prune_session_run(h) for _ in range(4)]
        lib.ref_prune_session_free(h)
        out.append((build.value, float(np.median(ts[1:])) / n, int(used)))
        if nt == 1:
            break
    (build, per_fam, used), (_, per_fam1, _) = out
    return build, per_fam, per_fam1, used


def main_config5(args):
    """Optimizer-scale gamma run: 100 taxa, 200,000 families, B = 500, K = 8;
    the first --evals of the 200 seeded (lambda, alpha) trial points scored
    one after another through one captured plan per path (the scoring loop
    of gamma_lambda_optimizer: keys set, graph replayed, total read)."""
    from cafe5_amd.model import ResidentGammaModel
    F = 200000 if args.families == 20000 else args.families
    K, B, T = 8, 500, 200
    w = wl.config_optimizer(seed=5, F=F, K=K, bound=B, T=T)
    n_eval = max(1, min(args.evals, T))
    log(f"workload: {w.name} F={F} taxa={w.tree.n_leaves} B={B} K={K}; scoring {n_eval} of {T} trial points")
    eng = native.Engine(0)
    tuning = load_tuning(eng, args.tuning)
    prior = priors.uniform(B, B)
    n1, n2 = tree_flop_factors(w.tree)
    L = B + 1
    flops_fam = K * 2.0 * L * (n1 * L + n2 * B)
    runs = {}
    for path in (native.PATH_BATCHED16, native.PATH_BATCHED32, native.PATH_BATCHED64):
        m = ResidentGammaModel(eng, w.tree, w.counts, B, B, K, prior, path=path)
        m.infer_family_likelihoods(float(w.lambdas[0]), float(w.alphas[0]), records=False)   # warm-up
        totals, recomputed = [], []
        t0 = time.perf_counter()
        for t in range(n_eval):
            totals.append(m.infer_family_likelihoods(float(w.lambdas[t]), float(w.alphas[t]), records=False))
            recomputed.append(int(m.last["n_fallback"]))
        ms = (time.perf_counter() - t0) * 1e3 / n_eval
        # one profiled execution at trial 0: phases and per-class kernel times
        m.plan.set_keys(*m._keys(float(w.lambdas[0]), get_gamma_rates(K, float(w.alphas[0]))))
        ph, graph_ok = m.plan.profile()
        cls = m.plan.class_times()
        m.plan.execute()
        lnl0 = m.plan.results()["lnl"][0]
        runs[path] = {"ms": ms, "totals": np.array(totals), "recomputed": recomputed, "phases": ph,
                      "graph": graph_ok, "classes": cls, "lnl0": lnl0}
        log(f"{native.PATH_NAMES[path]}: {ms:.1f} ms per evaluation, phases={np.round(ph, 2).tolist()} "
            f"recomputed={recomputed[:4]}...")
        m.close()
    head = native.PATH_BATCHED32
    r = runs[head]
    b64 = runs[native.PATH_BATCHED64]
    fin = np.isfinite(b64["totals"])
    acc = {}
    for path in (native.PATH_BATCHED32, native.PATH_BATCHED16):
        tr = np.abs(runs[path]["totals"][fin] - b64["totals"][fin]) / np.abs(b64["totals"][fin])
        a, b = runs[path]["lnl0"], b64["lnl0"]
        ok = np.isfinite(b)
        acc[native.PATH_NAMES[path]] = {
            "reference_for_comparison": "batched64 GPU path (matches the FP64 oracle to 1e-9 in tests)",
            "max_rel_dlnL_total_over_evaluations": float(tr.max()) if tr.size else None,
            "max_abs_dlnL_per_family_trial0": float(np.max(np.abs(a[ok] - b[ok]))),
            "nonfinite_totals_match": bool(np.array_equal(np.isfinite(runs[path]["totals"]), fin)),
            "families_recomputed_fp64_mean": float(np.mean(runs[path]["recomputed"]))}
    out = {
        "metric": "family likelihoods/sec, batched FP32-MFMA gamma model (K=8) over an optimizer's trial points "
                  "vs FP64 on GPU at |ΔlnL_total|≤1e-6",
        "value": F / (r["ms"] * 1e-3), "unit": "family likelihoods/s", "n_gpus": 1, "steps": n_eval,
        "warmup": 1, "ms_per_step": r["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"optimizer-100taxa: seeded Yule tree (100 taxa, height 100), {F} birth-death "
                               f"families, size bound 500, K=8 gamma categories, {n_eval} of 200 seeded (lambda, "
                               f"alpha) trial points scored one after another through one captured plan "
                               f"(1584 matrices of 501x501 rebuilt per evaluation)",
                   "path": "batched32", "graph": bool(r["graph"]), "families": F, "evaluations": n_eval},
        "ms_per_200_evaluations": r["ms"] * 200,
        "paths": {native.PATH_NAMES[p]: {"value": F / (runs[p]["ms"] * 1e-3), "ms_per_evaluation": runs[p]["ms"],
                                         "ms_per_200_evaluations": runs[p]["ms"] * 200,
                                         "phases_ms": np.round(runs[p]["phases"], 3).tolist()} for p in runs},
        "finite_evaluations": int(fin.sum()),
        "accuracy": acc,
        "roofline": _class_roofline(r, F * flops_fam, PEAK_FP32_MFMA,
                                    "K 2 L (n1 L + n2 R), K=%d n1=%d n2=%d L=%d" % (K, n1, n2, L)),
        "roofline_batched16": _class_roofline(runs[native.PATH_BATCHED16], F * flops_fam, PEAK_FP16_MFMA,
                                              "K 2 L (n1 L + n2 R) against the dense FP16 peak"),
        "roofline_batched64": _class_roofline(b64, F * flops_fam, 78.6e12, "K 2 L (n1 L + n2 R) against FP64"),
        "tuning": tuning,
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        threads = _host_threads()
        est = _ref_cpu_prune(w, B, B, float(w.lambdas[0]), 64, 4, threads)
        if est:
            build, per_fam, per_fam1, used = est
            eval_s = K * build + K * per_fam * F
            eval1 = K * build * used + K * per_fam1 * F   # the build taken to scale with its threads
            out["cpu_baseline"] = {
                "value": F / eval_s, "unit": "family likelihoods/s", "cores": used, "kind": "reference",
                "one_thread_value": F / eval1, "cpu_model": _cpu_model(), "seconds_per_evaluation": eval_s,
                "sample": (f"reference matrix_cache::precalculate_matrices and inference_prune (oracle/_ref, -O3 "
                           f"-march=x86-64-v3, OpenMP over families as in gamma_model) on the same tree at B=R=500 "
                           f"and one lambda: the build of its 198 matrices of 501x501 ({build:.2f} s on {used} "
                           f"threads) and the pruning of a seeded 64-family subsample (one warm-up, median of 3: "
                           f"{per_fam * 1e3:.2f} ms per family; one thread on 4 families: {per_fam1 * 1e3:.1f} ms), "
                           f"scaled by K=8 (8x the keys, 8 prunes per family; gamma_model has no duplicate scan) to "
                           f"all {F} families; the one-thread figure scales the build by {used}")}
            log(f"cpu baseline: {out['cpu_baseline']}")
    eng.close()
    print(json.dumps(out))


def main_config13(args):
    """One parameter point through a captured plan per path.  Config 1: the
    plain base model (seeded Yule tree of 12 taxa, 2,000 families, lambda
    0.002, B = R = 100; base_model.cpp:53-100).  Config 3: gamma rate
    variation (34 taxa, 50,000 families, K = 4 categories at alpha = 0.5,
    lambda 0.003, B = R = 100; gamma_core.cpp:97-237), with the category
    posteriors compared across paths."""
    from cafe5_amd.model import get_gamma
    gamma = args.config == 3
    if gamma:
        F = 50000 if args.families == 20000 else args.families
        K, alpha = 4, 0.5
        w = wl.config_gamma(seed=3, F=F, K=K, alpha=alpha, lam=0.003)
        freq, rate = get_gamma(K, alpha)
    else:
        F = 2000 if args.families == 20000 else args.families
        K, alpha = 1, None
        w = wl.config_base(seed=1, F=F)
        freq, rate = np.ones(1), np.ones(1)
    B = R = int(w.bounds[0])
    lam = float(w.lambdas[0])
    log(f"workload: {w.name} F={F} taxa={w.tree.n_leaves} B={B} lambda={lam} K={K}")
    eng = native.Engine(0)
    tuning = load_tuning(eng, args.tuning)
    eng.set_tree(w.tree)
    eng.set_families(w.counts, B, R)
    keys = native.KeyTable()
    slots = keys.node_slots(w.tree, np.stack([np.full(w.tree.n_nodes, lam * r) for r in rate]))
    lam_q, t_q = keys.arrays()
    eng.matrix_reset(B + 1, len(lam_q))
    prior = priors.uniform(R, R)[None, :]
    paths = (native.PATH_BATCHED32, native.PATH_BATCHED16, native.PATH_BATCHED64, native.PATH_PLAIN64)
    runs = _path_runs(eng, slots, prior, lam_q, t_q, paths, args.steps, args.warmup, args.fp64_steps,
                      model=native.MODEL_GAMMA if gamma else native.MODEL_BASE, cat_probs=freq if gamma else None)
    n1, n2 = tree_flop_factors(w.tree)
    L = B + 1
    flops = F * K * 2.0 * L * (n1 * L + n2 * R)
    head = native.PATH_BATCHED32
    r = runs[head]
    name = (f"gamma-34taxa: seeded Yule tree (34 taxa, height 100), {F} birth-death families with gamma-distributed "
            f"rates, size bound 100, K=4 categories at alpha=0.5, lambda 0.003, {len(lam_q)} matrices of 101x101 "
            f"per step" if gamma else
            f"base-12taxa: seeded Yule tree (12 taxa, height 100), {F} birth-death families (root 1 + Poisson(4)), "
            f"size bound 100, lambda 0.002, {len(lam_q)} matrices of 101x101 per step")
    out = {
        "metric": ("family likelihoods/sec, batched FP32-MFMA gamma model (K=4, posteriors) vs FP64 on GPU at "
                   "|ΔlnL_total|≤1e-6" if gamma else
                   "family likelihoods/sec, batched FP32-MFMA base model vs FP64 on GPU at |ΔlnL_total|≤1e-6"),
        "value": F / (r["ms"] * 1e-3), "unit": "family likelihoods/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r["ms"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": name, "path": "batched32 (FP32 MFMA + certified FP64 recompute)",
                   "graph": bool(r["graph"]), "families": F, "trials": 1},
        "paths": {native.PATH_NAMES[p]: {"value": F / (runs[p]["ms"] * 1e-3), "ms_per_step": runs[p]["ms"],
                                         "phases_ms": np.round(runs[p]["phases"], 4).tolist()} for p in paths},
        "speedup_vs_plain64": runs[native.PATH_PLAIN64]["ms"] / r["ms"],
        "accuracy": _accuracy(runs, head, native.PATH_BATCHED64),
        "accuracy_batched16": _accuracy(runs, native.PATH_BATCHED16, native.PATH_BATCHED64),
        "roofline": _class_roofline(r, flops, PEAK_FP32_MFMA,
                                    "K 2 L (n1 L + n2 R), K=%d n1=%d n2=%d L=%d" % (K, n1, n2, L)),
        "tuning": tuning,
        "cpu_baseline": None,
    }
    if gamma:
        p64 = runs[native.PATH_BATCHED64]["res"]["post"]
        out["posteriors"] = {native.PATH_NAMES[p]: {"max_abs_dpost_vs_batched64":
                                                    float(np.nanmax(np.abs(runs[p]["res"]["post"] - p64)))}
                             for p in (native.PATH_BATCHED32, native.PATH_BATCHED16, native.PATH_PLAIN64)}
    if not args.no_cpu:
        out["cpu_baseline"] = _ref_cpu_base(w, B, R, None, None, min(2000, F), _host_threads(),
                                            gamma=(freq, rate) if gamma else None)
        log(f"cpu baseline: {out['cpu_baseline']}")
    eng.close()
    print(json.dumps(out))


def get_gamma_rates(K, alpha):
    from cafe5_amd.model import get_gamma
    return get_gamma(K, alpha)[1]


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--families", type=int, default=20000)
    ap.add_argument("--trials", type=int, default=64)
    ap.add_argument("--fp64-steps", type=int, default=3, help="timed steps of the FP64 comparison paths")
    ap.add_argument("--cpu-lambdas", type=int, default=4)
    ap.add_argument("--cpu-families", type=int, default=4000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--kernels-only", action="store_true",
                    help="run each --kernel-paths plan eagerly --steps times plus a 256 MiB copy (counter passes)")
    ap.add_argument("--kernel-paths", default="2", help="paths for --kernels-only, e.g. 2,3,1")
    ap.add_argument("--evals", type=int, default=10, help="config 5: trial points scored per path")
    ap.add_argument("--tuning", default=os.path.join(ROOT, "cafe5_amd", "tune_cache.txt"),
                    help="tile-shape cache (cafe5_amd/tune.py) loaded into the engine; 'none' for built-in shapes")
    ap.add_argument("--config", type=int, default=2, choices=(1, 2, 3, 4, 5),
                    help="1: plain base model (2,000 families); 2: uneven-size lambda grid (the headline); 3: gamma "
                         "rate variation (50,000 families, K=4); 4: error model + lambda classes, split FP16; "
                         "5: optimizer-scale gamma run")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU design is "
              "family sharding with one all-reduce of per-trial totals and is not built", file=sys.stderr)
        sys.exit(2)

    if args.config in (1, 3):
        return main_config13(args)
    if args.config == 4:
        return main_config4(args)
    if args.config == 5:
        return main_config5(args)
    w = wl.config_uneven(seed=2, F=args.families, T=args.trials)
    if args.kernels_only:
        kernels_only(w, args.steps, [int(x) for x in args.kernel_paths.split(",") if x])
        return
    F, T = w.counts.shape[0], len(w.lambdas)
    log(f"workload: {w.name} F={F} T={T} taxa={w.tree.n_leaves}")
    eng = native.Engine(0)
    tuning = load_tuning(eng, args.tuning)
    buckets, slots, lam_q, t_q, prior, N = setup(eng, w)
    log(f"buckets (B,R,count): {buckets}; keys={len(lam_q)} N={N}; tuning={tuning}")
    use_graph = not args.no_graph

    # ---- headline: BATCHED32 ----
    p32 = eng.plan(slots, prior, lam_q, t_q, path=native.PATH_BATCHED32, use_graph=use_graph)
    ms32, ph32, graph_ok, cls32 = time_plan(p32, args.steps, args.warmup, args.steps)
    r32 = p32.results()
    rc64, rcplain = p32.recompute_counts()
    log(f"batched32: {ms32:.3f} ms/step, phases(build, prune, recompute, finalize)={ph32.round(3).tolist()} "
        f"graph={graph_ok} fallback={r32['n_fallback']}")

    # ---- FP64 paths on the same inputs ----
    p64 = eng.plan(slots, prior, lam_q, t_q, path=native.PATH_BATCHED64, use_graph=use_graph)
    ms64, ph64, _, _ = time_plan(p64, args.fp64_steps, 1, 1)
    r64 = p64.results()
    log(f"batched64: {ms64:.3f} ms/step phases={ph64.round(3).tolist()}")
    p16 = eng.plan(slots, prior, lam_q, t_q, path=native.PATH_BATCHED16, use_graph=use_graph)
    ms16, ph16, _, _ = time_plan(p16, args.fp64_steps, 1, 1)
    r16 = p16.results()
    rc16 = p16.recompute_counts()
    log(f"batched16: {ms16:.3f} ms/step phases={ph16.round(3).tolist()} recomputed={rc16}")
    pplain = eng.plan(slots, prior, lam_q, t_q, path=native.PATH_PLAIN64, use_graph=use_graph)
    msplain, phplain, _, _ = time_plan(pplain, args.fp64_steps, 1, 1)
    rplain = pplain.results()
    log(f"plain64: {msplain:.3f} ms/step phases={phplain.round(3).tolist()}")

    # ---- accuracy of BATCHED32 against the plain FP64 drop-in ----
    a = r32["lnl"]
    b = rplain["lnl"]
    fin = np.isfinite(b)
    same_fin = bool(np.array_equal(np.isfinite(a), fin))
    d = np.abs(a[fin] - b[fin])
    tot_rel = np.abs(r32["total"] - rplain["total"]) / np.abs(rplain["total"])
    tot_rel = tot_rel[np.isfinite(tot_rel)]
    d64 = np.abs(r64["lnl"][fin] - b[fin]) / np.abs(b[fin])
    accuracy = {
        "reference_for_comparison": "plain64 GPU path (matches the FP64 oracle to 1e-9 in tests)",
        "max_abs_dlnL_per_family": float(d.max()) if d.size else 0.0,
        "max_rel_dlnL_per_family": float((d / np.abs(b[fin])).max()) if d.size else 0.0,
        "max_rel_dlnL_total_over_lambdas": float(tot_rel.max()) if tot_rel.size else 0.0,
        "nonfinite_families_match": same_fin,
        "nonfinite_family_likelihoods": int((~fin).sum()),
        "fp32_families_recomputed_batched64": int(rc64),
        "families_recomputed_plain64_near_subnormal": int(rcplain),
        "batched64_max_rel_dlnL": float(d64.max()) if d64.size else 0.0,
    }
    # the bars the reduced path is held to (the metric's |dlnL| <= 1e-6 is
    # the total's, relative; per family the path certifies a looser bar)
    fam_ok = bool(np.all((d <= 5e-6 + 1e-7 * np.abs(b[fin])) & (d <= 1e-5 * np.abs(b[fin])))) if d.size else True
    tot_ok = bool(tot_rel.max() <= 1e-6) if tot_rel.size else True
    accuracy["bars"] = {"total_rel": 1e-6, "per_family": "|dlnL| <= 5e-6 + 1e-7 |lnL| and <= 1e-5 |lnL|",
                        "total_within": tot_ok, "per_family_within": fam_ok,
                        "within": bool(tot_ok and fam_ok and same_fin)}
    log(f"accuracy: {accuracy}")
    if not accuracy["bars"]["within"]:
        log("WARNING: the reduced path is outside its accuracy bars on this run")

    # ---- roofline of the dominant kernel: the FP32 main-pass launch of the
    # size class that takes longest (per-class HIP events, same stream) ----
    n1, n2 = tree_flop_factors(w.tree)
    nt_cls, tiles_cls, ms_cls, launch_cls = cls32
    dom = int(np.argmax(ms_cls))
    dom_nt = int(nt_cls[dom])
    dom_flops = 0.0
    dom_buckets = []
    for B, R, cnt in buckets:
        L = B + 1
        if max(-(-L // 16), -(-(R + 1) // 16)) == dom_nt:
            dom_flops += cnt * T * 2.0 * L * (n1 * L + n2 * R)
            dom_buckets.append(B)
    dom_ms = float(ms_cls[dom])
    achieved = dom_flops / (dom_ms * 1e-3) if dom_ms > 0 else None
    phase_flops = sum(cnt * T * 2.0 * (B + 1) * (n1 * (B + 1) + n2 * R) for B, R, cnt in buckets)
    dl = launch_cls[dom]
    lane_dom = int(dl[0]) == 1
    if lane_dom:
        dom_kernel = f"void k_prune_lane<{int(dl[1])}, "
        dom_desc = (f"k_prune_lane<{int(dl[1])},*> (one family per lane, FP32 vector FMAs, no MFMA) main-pass "
                    f"launch of the {dom_nt}-tile size class (buckets B={dom_buckets}, {int(tiles_cls[dom])} tiles "
                    f"of 64 families)")
    else:
        dom_kernel = f"void k_prune_mfma<float, {int(dl[1])}, {int(dl[2])}, true, {'true' if dl[3] else 'false'}>"
        dom_desc = (f"k_prune_mfma<float,{int(dl[1])},{int(dl[2])},true,*> main-pass launch of the {dom_nt}-tile "
                    f"size class (buckets B={dom_buckets}, {int(tiles_cls[dom])} tiles of 16 families)")
    traffic = read_traffic(dom_nt, dom_kernel)
    build_bytes = len(lam_q) * N * N * (8 + 4 + 4)
    roofline = {"bound": "valu" if lane_dom else "mfma", "achieved": achieved / 1e12 if achieved else None,
                "peak": PEAK_FP32_MFMA / 1e12, "unit": "TFLOP/s",
                "frac": (achieved / PEAK_FP32_MFMA) if achieved else None,
                "frac_of_unpacked_fp32_valu": (achieved / PEAK_FP32_VALU_UNPACKED) if (achieved and lane_dom) else None,
                "traffic": traffic["bytes"] if traffic else None,
                "kernel": dom_desc,
                "algorithmic_flops_per_launch": dom_flops, "avg_ms": dom_ms,
                "flops_per_family_likelihood": "2 L (n1 L + n2 R), n1=%d n2=%d" % (n1, n2),
                "traffic_source": traffic["source"] if traffic else None,
                "classes": [{"tiles_per_vector": int(a_), "tiles": int(b_), "ms": float(c_),
                             "launch": {"kernel": KERNEL_KIND[int(l_[0])],
                                        "tiles_per_wave_or_rows_per_lane": int(l_[1]), "waves": int(l_[2]),
                                        "partials_in_lds": bool(l_[3]), "workgroups": int(l_[4]),
                                        "families_per_tile": int(l_[5])}}
                            for a_, b_, c_, l_ in zip(nt_cls, tiles_cls, ms_cls, launch_cls)],
                "prune_phase": {"ms": ph32[1], "algorithmic_flops": phase_flops,
                                "TFLOP_per_s": phase_flops / (ph32[1] * 1e-3) / 1e12 if ph32[1] > 0 else None},
                "matrix_build": {"ms": ph32[0], "bytes_written": build_bytes,
                                 "GB_per_s": build_bytes / (ph32[0] * 1e-3) / 1e9 if ph32[0] > 0 else None,
                                 "hbm_frac": build_bytes / (ph32[0] * 1e-3) / PEAK_HBM if ph32[0] > 0 else None}}

    # PLAIN64 (the drop-in): its flops (the same contractions, plus the dense
    # leaf matrix-vector products the reference makes) against the FP64 peak,
    # and the matrix bytes its per-family products read against HBM
    n_leaf = w.tree.n_leaves
    plain_flops = sum(cnt * T * 2.0 * (B + 1) * (n1 * (B + 1) + n2 * R + n_leaf * (B + 1)) for B, R, cnt in buckets)
    plain_bytes = sum(cnt * T * 8.0 * (B + 1) * (n1 * (B + 1) + n2 * R + n_leaf * (B + 1)) for B, R, cnt in buckets)
    plain_s = phplain[1] * 1e-3
    plain_roofline = {"bound": "fp64", "achieved_TFLOP_per_s": plain_flops / plain_s / 1e12,
                      "frac_of_fp64_peak": plain_flops / plain_s / PEAK_FP64,
                      "matrix_bytes_GB_per_s": plain_bytes / plain_s / 1e9,
                      "frac_of_hbm": plain_bytes / plain_s / PEAK_HBM,
                      "note": "prune phase of one profiled PLAIN64 execution; flops and matrix bytes count every "
                              "leaf branch as the reference's dense matrix-vector product"}
    value = F * T / (ms32 * 1e-3)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "family likelihoods/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms32,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "uneven-34taxa lambda grid: seeded Yule tree (34 taxa, height 100), "
                               f"{F} birth-death families, bounds 20-200 in {len(buckets)} buckets, "
                               f"{T}-point lambda grid 1e-4..1e-2, {len(lam_q)} matrices of {N}x{N} per step",
                   "path": "batched32 (FP32 MFMA + certified FP64 recompute)",
                   "graph": bool(graph_ok), "families": F, "trials": T},
        "plain64": {"value": F * T / (msplain * 1e-3), "ms_per_step": msplain,
                    "roofline": plain_roofline},
        "batched16": {"value": F * T / (ms16 * 1e-3), "ms_per_step": ms16,
                      "families_recomputed_fp64": int(rc16[0]),
                      "max_abs_dlnL_per_family": float(np.max(np.abs(r16["lnl"][fin] - b[fin]))),
                      "max_rel_dlnL_total": float(np.max(np.abs(r16["total"] - rplain["total"]) /
                                                         np.abs(rplain["total"])))},
        "batched64": {"value": F * T / (ms64 * 1e-3), "ms_per_step": ms64},
        "speedup_vs_plain64": msplain / ms32,
        "accuracy": accuracy,
        "roofline": roofline,
        "tuning": tuning,
        "cpu_baseline": None,
    }
    for p in (p32, p64, p16, pplain):
        p.close()
    # the reconstruction next to the likelihood path (an extra measurement:
    # its failure is reported, not fatal to the headline line)
    recon_states = None
    try:
        out["reconstruct"], recon_states = time_reconstruct(eng, w)
        log(f"reconstruct: {out['reconstruct']}")
    except Exception as e:  # noqa: BLE001
        out["reconstruct"] = {"error": str(e)}
    if not args.no_cpu:
        threads = _host_threads()
        log(f"cpu baseline on {threads} threads ...")
        out["cpu_baseline"] = cpu_baseline(w, args.cpu_lambdas, args.cpu_families, threads)
        log(f"cpu baseline: {out['cpu_baseline']}")
        ref_so = os.path.join(ROOT, "oracle", "_ref", "libcaferef.so")
        if recon_states is not None and os.path.exists(ref_so):
            try:
                out["reconstruct"]["cpu_reference"] = _cpu_reconstruct(ref_so, w, threads, recon_states)
                log(f"reconstruct cpu: {out['reconstruct']['cpu_reference']}")
            except Exception as e:  # noqa: BLE001
                out["reconstruct"]["cpu_reference"] = {"error": str(e)}
    eng.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
