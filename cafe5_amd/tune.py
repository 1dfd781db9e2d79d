"""Tile-shape autotuner: measures candidate launch shapes for every size class
of a workload on the GPU and writes the fastest to a cache file that
cafe_gpu_tuning_load reads (format in include/cafe_gpu.h).

    python -m cafe5_amd.tune [--workload uneven|gamma|errmodel|optimizer] [--paths 2,3,1]
                              [--out cafe5_amd/tune_cache.txt]

A size class is the set of buckets whose vectors have the same number of
16-row tiles.  For each class and path the candidates are the per-lane
kernel (where the class can take it) and every instantiated matrix-core
shape (output tiles per wave x waves >= the class's tiles; partials in LDS
where they fit, else global scratch; for BATCHED32 classes of up to 32 tiles
also 16- or 32-family tiles) at grid multipliers 4, 8 and 16 (LDS) or 1, 2
and 4 (global); each is timed as that class's launch in profiled
executions of the plan (median of 3), with the classes already tuned held at
their choices.  Cache entries are keyed by (path, tiles, leaves of the tree).
"""
from __future__ import annotations

import argparse
import os
import sys
import time
from typing import Dict, List, Tuple

import numpy as np

from . import native, priors, workloads as wl

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_CACHE = os.path.join(HERE, "tune_cache.txt")

# instantiated matrix-core shapes (tiles per wave, waves); the FP16 kernel
# stops at eight waves
MFMA_SHAPES = [(1, 1), (2, 1), (3, 1), (4, 1), (3, 2), (4, 2), (3, 4), (4, 4), (3, 8), (4, 8), (3, 16), (4, 16),
               (1, 4), (1, 8), (2, 8), (1, 16), (2, 16)]   # the last five: latency shapes
F16_SHAPES = [s for s in MFMA_SHAPES[:12] if s[1] <= 8]
# shapes of the FP32 kernel with 32-family tiles (two blocks of 16 per
# matrix operand), vectors of up to 32 tiles; the 16-wave shapes have global
# partials only
FB2_SHAPES = [(2, 4), (1, 8), (2, 8), (1, 16), (2, 16)]
FB2_MAX_TILES = 32
LANE_TILES = 3          # the per-lane kernel takes vectors of up to 33 rows (3 tiles)
MAX_SHAPE_TILES = 64    # the matrix-core shapes' largest vectors (1,024 rows)


def candidates(path: int, nt: int) -> List[Tuple[int, int, int, int, int]]:
    """(kernel, ntw, w, lds, grid_mult) candidates for a class of nt tiles."""
    out = []
    if path in (native.PATH_BATCHED32, native.PATH_BATCHED16) and nt <= LANE_TILES:
        for gm in (8, 16, 32):
            out.append((1, 0, 0, 0, gm))
    shapes = F16_SHAPES if path == native.PATH_BATCHED16 else MFMA_SHAPES
    for ntw, w in shapes:
        if ntw * w < nt or (ntw - 1) * w >= nt and ntw > 1:   # covers the class without an idle tile per wave
            continue
        for lds in (1, 0):
            if lds and (ntw, w) == (2, 16):
                continue   # instantiated with global partials only
            for gm in ((4, 8, 16) if lds else (1, 2, 4)):
                out.append((0, ntw, w, lds, gm))
    if path == native.PATH_BATCHED32 and nt <= FB2_MAX_TILES:
        for ntw, w in FB2_SHAPES:
            if ntw * w < nt or (ntw - 1) * w >= nt and ntw > 1:
                continue
            for lds in ((1, 0) if w < 16 else (0,)):
                for gm in ((4, 8, 16) if lds else (1, 2, 4)):
                    out.append((2, ntw, w, lds, gm))
    return out


def plan_key_of(nts) -> str:
    """The plan= key of a plan's size classes: their tiles per vector, up to
    64 (classes of larger vectors run the plain FP64 kernel, which has no
    launch shape to tune and no place in the engine's key)."""
    return ",".join(str(int(n)) for n in sorted(nts) if int(n) <= MAX_SHAPE_TILES)


def _class_ms(plan, nt: int, reps: int = 3) -> float:
    ms = []
    for _ in range(reps):
        plan.profile()
        nts, _, cms, _ = plan.class_times()
        idx = [i for i, x in enumerate(nts) if int(x) == nt]
        ms.append(float(cms[idx[0]]) if idx else float("inf"))
    return float(np.median(ms))


def _step_ms(plan, reps: int = 20, rounds: int = 3) -> float:
    """Median over rounds of the ms per execution of reps graph replays (the
    size classes run concurrently, as in the timed steps)."""
    for _ in range(2):
        plan.execute()
        plan.synchronize()
    ms = []
    for _ in range(rounds):
        t0 = time.perf_counter()
        for _ in range(reps):
            plan.execute()
            plan.synchronize()
        ms.append((time.perf_counter() - t0) * 1e3 / reps)
    return float(np.median(ms))


def tune_plan(eng: native.Engine, make_plan, path: int, leaves: int, log=print, objective: str = "step",
              min_gain: float = 0.005) -> List[Tuple]:
    """Measure every class of the plan make_plan(path) builds; returns the
    chosen entries (path, tiles, leaves, kernel, ntw, w, lds, grid_mult, ms,
    plan): plan is the plan's size classes ("2,3,..." tiles per vector) for
    the step objective, whose measurements hold only beside those classes,
    and "" (any plan) for the class objective.
    objective "class": the class's own launch time in profiled executions
    (classes one after another); "step": the whole execution's time with
    the classes running concurrently, as the timed steps run them (a shape
    that is fastest alone can hold LDS or matrix cores another class needs).
    A candidate replaces the current choice only when it is faster by more
    than min_gain (relative), so timing noise does not pick shapes."""
    base = make_plan(path)
    base.profile()
    nts, tiles, cms, launch = base.class_times()
    cur_ms = _step_ms(base) if objective == "step" else None
    base.close()
    plan_key = plan_key_of(nts) if objective == "step" else ""
    chosen = []
    for nt, default_ms, l in sorted(zip(nts, cms, launch), key=lambda x: -x[1]):
        nt = int(nt)
        best = (float(default_ms) if objective == "class" else cur_ms, None)
        tried = 0
        for cand in candidates(path, nt):
            kernel, ntw, w, lds, gm = cand
            eng.tuning_set(path, nt, leaves, kernel, ntw, w, bool(lds), gm)
            try:
                p = make_plan(path)
            except native.CafeGpuError:
                continue
            ms = _class_ms(p, nt) if objective == "class" else _step_ms(p)
            p.close()
            tried += 1
            if ms < best[0] * (1.0 - min_gain):
                best = (ms, cand)
        if best[1] is None:   # the built-in shape was fastest: pin it as measured
            k = int(l[0])
            kern = 2 if k == 0 and int(l[5]) == 32 else 0
            best_c = (1, 0, 0, 0, 0) if k == 1 else (kern, int(l[1]), int(l[2]), int(l[3]), 0)
        else:
            best_c = best[1]
        eng.tuning_set(path, nt, leaves, *best_c[:3], bool(best_c[3]), best_c[4])
        chosen.append((path, nt, leaves, *best_c, best[0], plan_key))
        what = "class" if objective == "class" else "step"
        log(f"path {native.PATH_NAMES[path]} class {nt} tiles: {what} {float(default_ms if objective == 'class' else cur_ms):.3f} "
            f"ms before, {best[0]:.3f} ms with {best_c} ({tried} candidates)")
        if objective == "step":
            cur_ms = best[0]
    return chosen


def write_cache(path: str, entries, header: str):
    with open(path, "w") as fh:
        fh.write("# cafe_gpu tile-shape cache (cafe5_amd/tune.py): path tiles leaves kernel ntw w lds grid_mult "
                 "[measured ms] [plan=tiles of every size class]\n")
        fh.write(f"# {header}\n")
        for e in entries:
            fh.write(" ".join(str(int(x)) for x in e[:8]) + f"  {e[8]:.4f}" + (f"  plan={e[9]}" if e[9] else "") + "\n")


def _workload(name: str):
    if name == "uneven":
        w = wl.config_uneven(seed=2, F=20000, T=64)
        lams = np.repeat(w.lambdas[:, None], w.tree.n_nodes, axis=1)
        return w, lams, w.bounds, None, 1
    if name == "gamma":
        w = wl.config_gamma(seed=3, F=50000, K=4, alpha=0.5, lam=0.003)
        from .model import get_gamma
        _, rate = get_gamma(4, 0.5)
        lams = np.stack([np.full(w.tree.n_nodes, float(w.lambdas[0]) * r) for r in rate])
        return w, lams, w.bounds, None, 4
    if name == "errmodel":
        w = wl.config_error_multi(seed=4, F=100000, bound=150)
        return w, w.lambdas[w.node_class][None, :], w.bounds, wl.error_table(0.05, 150), 1
    if name == "optimizer":
        w = wl.config_optimizer(seed=5, F=200000, K=8, bound=500, T=200)
        from .model import get_gamma
        _, rate = get_gamma(8, float(w.alphas[0]))
        lams = np.stack([np.full(w.tree.n_nodes, float(w.lambdas[0]) * r) for r in rate])
        return w, lams, w.bounds, None, 8
    raise ValueError(name)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--workload", default="uneven", choices=("uneven", "gamma", "errmodel", "optimizer"))
    ap.add_argument("--paths", default="2,3,1")
    ap.add_argument("--out", default=DEFAULT_CACHE)
    ap.add_argument("--append", action="store_true", help="keep the entries already in --out")
    ap.add_argument("--objective", default="step", choices=("step", "class"),
                    help="step: whole replayed execution (classes concurrent); class: the class's launch alone")
    args = ap.parse_args()
    w, node_lams, bounds, err, K = _workload(args.workload)
    eng = native.Engine(0)
    eng.set_tree(w.tree)
    eng.set_families(w.counts, bound_B=bounds, bound_R=bounds)
    eng.set_error_model(err)
    buckets = eng.buckets()
    keys = native.KeyTable()
    slots = keys.node_slots(w.tree, node_lams)
    lam_q, t_q = keys.arrays()
    eng.matrix_reset(int(bounds.max()) + 1, len(lam_q))
    stride = max(r for _, r, _ in buckets)
    prior = np.zeros((len(buckets), stride))
    for i, (_, r, _) in enumerate(buckets):
        prior[i, :r] = priors.uniform(r, r)
    model = native.MODEL_GAMMA if K > 1 else native.MODEL_BASE
    cat = np.full(K, 1.0 / K) if K > 1 else None

    def make_plan(path):
        return eng.plan(slots, prior, lam_q, t_q, path=path, model=model, cat_probs=cat,
                        use_graph=args.objective == "step")

    entries = []
    if args.append and os.path.exists(args.out):
        n = eng.tuning_load(args.out)
        print(f"kept {n} entries of {args.out}", flush=True)
        with open(args.out) as fh:
            for line in fh:
                line = line.split("#", 1)[0].split()
                if len(line) >= 8:
                    plan = next((t[5:] for t in line[8:] if t.startswith("plan=")), "")
                    ms = next((float(t) for t in line[8:9] if not t.startswith("plan=")), 0.0)
                    entries.append(tuple(int(x) for x in line[:8]) + (ms, plan))
    t0 = time.time()
    for path in [int(x) for x in args.paths.split(",") if x]:
        got = tune_plan(eng, make_plan, path, w.tree.n_leaves, log=lambda m: print(m, flush=True),
                        objective=args.objective)
        keyset = {(e[0], e[1], e[2], e[9]) for e in got}
        entries = [e for e in entries if (e[0], e[1], e[2], e[9]) not in keyset] + got
    write_cache(args.out, entries, f"measured on {args.workload} ({w.name}), objective {args.objective}, in "
                                   f"{time.time() - t0:.0f} s")
    print(f"wrote {len(entries)} entries to {args.out}", flush=True)
    eng.close()


if __name__ == "__main__":
    sys.exit(main())
