"""The C++ host interface (cafe5_amd/host, mirroring the reference's classes)
on the GPU: the reference's known-answer tests run through it, and seeded
models evaluated through it match the oracle."""
import os
import subprocess

import numpy as np
import pytest

from cafe5_amd import workloads as wl
from bars import reduced_ok
from oracle_lib import Oracle, Ref

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "adapter_test")

pytestmark = pytest.mark.gpu


def run(*args, timeout=120):
    res = subprocess.run([DRIVER, *args], capture_output=True, text=True, timeout=timeout)
    return res.returncode, res.stdout, res.stderr


def test_reference_known_answers_through_cpp_interface():
    rc, out, err = run("kats")
    lines = out.splitlines()
    fails = [l for l in lines if l.startswith("FAIL")]
    assert rc == 0 and not fails, (fails, err)
    assert sum(l.startswith("PASS") for l in lines) >= 18
    for name in ("lambda_epsilon_optimizer", "gamma_lambda_optimizer__provides_two_guesses",
                 "gamma_optimizer__creates_single_initial_guess",
                 "gamma_lambda_optimizer_updates_model_alpha_and_lambda"):
        assert "PASS " + name in lines
    gamma = [l for l in lines if l.startswith("GAMMA 4 0.5")][0].split()[3:]
    f, r = Oracle.get_gamma(4, 0.5)
    assert np.array_equal(np.array([float(x) for x in gamma]), r)


def write_model(path, tree, counts, lambdas, B, R, prior, gamma=None, eps=None, node_class=None, engine_path=0):
    with open(path, "w") as fh:
        fh.write(f"newick {tree.newick()}\n")
        if node_class is None:
            fh.write("lambda_tree -\n")
        else:
            fh.write(f"lambda_tree {tree.newick(labels=[int(c) + 1 for c in node_class])}\n")
        fh.write("lambdas " + " ".join(repr(float(x)) for x in np.atleast_1d(lambdas)) + "\n")
        fh.write(f"bounds {B} {R}\n")
        fh.write(f"prior {prior}\n")
        fh.write(f"gamma {gamma[0]} {gamma[1]}\n" if gamma else "gamma none\n")
        fh.write(f"error {eps}\n" if eps is not None else "error none\n")
        fh.write(f"path {engine_path}\n")
        fh.write("species " + " ".join(tree.leaf_names()) + "\n")
        for f, row in enumerate(counts):
            fh.write(f"family F{f} " + " ".join(str(int(c)) for c in row) + "\n")


def parse(out):
    total = None
    lnl, stash = [], []
    for line in out.splitlines():
        parts = line.split()
        if parts[0] == "TOTAL":
            total = float(parts[1])
        elif parts[0] == "LNL":
            lnl.append(float(parts[2]))
        elif parts[0] == "STASH":
            stash.append([float(x) for x in parts[2:]])
    return total, np.array(lnl), np.array(stash)


@pytest.mark.parametrize("engine_path", [0, 1, 2])
def test_base_model_through_cpp_interface(tmp_path, engine_path):
    w = wl.config_base(seed=1, F=300)
    B = R = 100
    f = tmp_path / "model.txt"
    write_model(f, w.tree, w.counts, 0.002, B, R, f"uniform {R}", engine_path=engine_path)
    rc, out, err = run("base", str(f))
    assert rc == 0, err
    total, lnl, _ = parse(out)
    ref = Oracle.evaluate(w.tree, np.full(w.tree.n_nodes, 0.002), w.counts, B, R, Oracle.prior_uniform(R, R),
                          threads=8)
    if engine_path == 2:
        assert np.all(reduced_ok(lnl, ref["lnl"]))
        assert abs(total - ref["total"]) <= 1e-6 * abs(ref["total"])
    else:
        assert np.all(np.abs(lnl - ref["lnl"]) <= 1e-9 * np.abs(ref["lnl"]))
        assert abs(total - ref["total"]) <= 1e-9 * abs(ref["total"])


def test_multi_lambda_error_model_poisson_through_cpp_interface(tmp_path):
    w = wl.config_error_multi(seed=4, F=120, bound=60)
    B = R = 60
    f = tmp_path / "model.txt"
    write_model(f, w.tree, w.counts, w.lambdas, B, R, f"poisson 3.0 {R}", eps=0.05, node_class=w.node_class)
    rc, out, err = run("base", str(f))
    assert rc == 0, err
    total, lnl, _ = parse(out)
    err_tab = wl.error_table(0.05, B)
    ref = Oracle.evaluate(w.tree, w.lambdas[w.node_class], w.counts, B, R, Oracle.prior_poisson(3.0, R, R),
                          err=err_tab, threads=8)
    assert np.all(np.abs(lnl - ref["lnl"]) <= 1e-9 * np.abs(ref["lnl"]))


def test_gamma_model_through_cpp_interface(tmp_path):
    w = wl.config_gamma(seed=3, F=150, K=4, bound=60)
    B = R = 60
    f = tmp_path / "model.txt"
    write_model(f, w.tree, w.counts, 0.003, B, R, f"uniform {R}", gamma=(4, 0.5))
    rc, out, err = run("base", str(f))
    assert rc == 0, err
    total, _, stash = parse(out)
    fq, r = Oracle.get_gamma(4, 0.5)
    ref = Oracle.evaluate(w.tree, np.full(w.tree.n_nodes, 0.003), w.counts, B, R, Oracle.prior_uniform(R, R),
                          n_cat=4, multipliers=r, cat_probs=fq, threads=8)
    # stash rows: multiplier, category likelihood, family likelihood, posterior
    stash = stash.reshape(150, 4, 4)
    assert np.allclose(stash[:, :, 0], r[None, :], rtol=0, atol=0)
    assert np.all(np.abs(stash[:, :, 1] - ref["cat"]) <= 1e-9 * np.abs(ref["cat"]))
    assert np.all(np.abs(stash[:, :, 3] - ref["post"]) <= 1e-9)
    assert abs(total - ref["total"]) <= 1e-9 * abs(ref["total"])


# --------------------------------------------------------------------------
# p-values (probability.cpp:335-581) through the C++ interface
# --------------------------------------------------------------------------

def write_pvalue_input(path, tree, counts, lam, B, R, root=1, seed=0, sims=0, inner=None, build=1):
    with open(path, "w") as fh:
        fh.write(f"newick {tree.newick()}\n")
        fh.write(f"lambdas {lam!r}\n")
        fh.write(f"bounds {B} {R}\n")
        fh.write(f"root {root}\nseed {seed}\nsims {sims}\nbuild {build}\n")
        fh.write("species " + " ".join(tree.leaf_names()) + "\n")
        for f, row in enumerate(counts):
            fh.write(f"family F{f} " + " ".join(str(int(c)) for c in row) + "\n")
            if inner is not None:
                fh.write(f"inner F{f} {int(inner[f])}\n")


@pytest.mark.parametrize("root", [1, 4, 11])
def test_family_probabilities_through_cpp_interface(tmp_path, root):
    tree = wl.yule_tree(9, 4)
    lam, B, R = 0.006, 70, 30
    rng = np.random.default_rng(root)
    F = 40
    counts = rng.integers(0, 25, (F, tree.n_leaves)).astype(np.int32)
    inner = rng.integers(1, 30, F).astype(np.int32)
    counts[0] = 0                                        # a family lost everywhere
    f = tmp_path / "cond.txt"
    write_pvalue_input(f, tree, counts, lam, B, R, root=root, inner=inner)
    rc, out, err = run("condprob", str(f))
    assert rc == 0, err
    got = np.array([float(l.split()[2]) for l in out.splitlines() if l.startswith("PROB")])
    assert len(got) == F
    # the map holds inner[i] at every interior node, the root included
    want = np.array([Oracle.family_probability(tree, counts[i], lam, B, max(int(counts[i].max()), int(inner[i])),
                                               root, N=max(B, R) + 1) for i in range(F)])
    # PLAIN64 bar: 1e-9 relative per family
    assert np.all(np.abs(got - want) <= 1e-9 * np.abs(want)), np.max(np.abs(got - want) / np.abs(want))
    if Ref.available():
        ref = Ref.family_probabilities(tree, counts, inner, lam, B, R, root)
        assert np.all(np.abs(got - ref) <= 1e-9 * np.abs(ref))


def _parse_pvalues(out):
    sims, cond, pv = [], {}, {}
    for line in out.splitlines():
        p = line.split()
        if p[0] == "SIM":
            sims.append((int(p[1]), int(p[2]), np.array([int(x) for x in p[3:]])))
        elif p[0] == "COND":
            cond[int(p[1])] = np.array([float(x) for x in p[2:]])
        elif p[0] == "PVALUE":
            pv[p[1]] = float(p[2])
    return sims, cond, pv


def test_pvalues_through_cpp_interface(tmp_path):
    tree = wl.yule_tree(6, 8)
    lam, B, R, nsim, seed = 0.01, 40, 12, 25, 1234
    rng = np.random.default_rng(2)
    F = 10
    counts = rng.integers(1, 9, (F, tree.n_leaves)).astype(np.int32)
    counts[3] = counts[3] + 1                           # sizes up to 9: rint(9 * 1.25) = 11 < R
    f = tmp_path / "pv.txt"
    write_pvalue_input(f, tree, counts, lam, B, R, seed=seed, sims=nsim, build=0)
    rc, out, err = run("pvalues", str(f), timeout=300)
    assert rc == 0, err
    sims, cond, pv = _parse_pvalues(out)
    assert len(sims) == R * nsim and sorted(cond) == list(range(1, R + 1)) and len(pv) == F
    # each simulated family's probability against the oracle
    by_root = {r: [] for r in range(1, R + 1)}
    for r, mx, leaf in sims:
        assert mx >= r and leaf.max() <= mx
        by_root[r].append(Oracle.family_probability(tree, leaf, lam, B, mx, r, N=max(B, R) + 1))
    for r in range(1, R + 1):
        want = np.sort(np.array(by_root[r]))
        assert np.all(np.abs(cond[r] - want) <= 1e-9 * np.abs(want)), r
    # p-values from the distributions and the observed root vectors
    cd = np.stack([cond[r] for r in range(1, R + 1)])
    for i in range(F):
        root = Oracle.prune(tree, counts[i], np.full(tree.n_nodes, lam), B, R)
        want = Oracle.find_best_pvalue(root, int(counts[i].max()), cd)
        assert pv[f"F{i}"] == want, (i, pv[f"F{i}"], want)
    # the draws follow the reference's sequence: the first root size's
    # families equal create_family's after the same seed
    if Ref.available():
        first = Ref.create_families(tree, lam, B, R, max(B, R) + 1, seed, 1, nsim)
        assert np.array_equal(np.stack([leaf for r, _, leaf in sims[:nsim]]), first)


# --------------------------------------------------------------------------
# lambda search (optimizer.cpp, optimizer_scorer.cpp) on a device-resident
# model through the C++ interface
# --------------------------------------------------------------------------

def _parse_optimize(out):
    start, evals, result = None, [], None
    for line in out.splitlines():
        p = line.split()
        if p[0] == "START":
            start = [float(x) for x in p[1:]]
        elif p[0] == "EVAL":
            evals.append([float(x) for x in p[1:]])
        elif p[0] == "RESULT":
            result = (int(p[1]), float(p[2]), [float(x) for x in p[3:]])
    return start, np.array(evals), result


def _optimize_case(tmp_path, path):
    w = wl.config_base(seed=5, F=120)
    B = R = 60
    counts = np.minimum(w.counts, B)
    f = tmp_path / "opt.txt"
    write_pvalue_input(f, w.tree, counts, 0.002, B, R, seed=77)
    rc, out, err = run("optimize", str(f), str(path), timeout=300)
    assert rc == 0, err
    return w, counts, B, R, _parse_optimize(out)


def test_lambda_search_through_cpp_interface(tmp_path):
    """PLAIN64: the search visits the points the oracle's Nelder-Mead visits
    with the oracle's scores, and ends at the same lambda."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import nelder_mead as nmr

    w, counts, B, R, (start, evals, result) = _optimize_case(tmp_path, 0)
    prior = Oracle.prior_uniform(R, R)

    def oracle_score(x):
        if not x[0] > 0:                                  # lambda.h: an invalid lambda scores +inf
            return float("inf")
        v = Oracle.evaluate(w.tree, np.full(w.tree.n_nodes, x[0]), counts, B, R, prior, threads=8)["total"]
        return float("inf") if np.isnan(v) else v

    seen = []
    vals, score, iters = nmr.optimize(lambda x: (seen.append((x[0], oracle_score(x))), seen[-1][1])[1], start)
    # the first evaluation checks the initial guess; the search follows
    assert evals[0][0] == start[0]
    got = evals[1:]
    assert len(got) == len(seen), (len(got), len(seen))
    want = np.array(seen)
    assert np.array_equal(got[:, 0], want[:, 0])
    fin = np.isfinite(want[:, 1])
    assert np.array_equal(np.isfinite(got[:, 1]), fin)
    assert np.all(np.abs(got[fin, 1] - want[fin, 1]) <= 1e-9 * np.abs(want[fin, 1]))
    assert result[0] == iters and result[2][0] == vals[0]
    assert abs(result[1] - score) <= 1e-9 * abs(score)


def test_lambda_search_batched32_through_cpp_interface(tmp_path):
    """BATCHED32 scores (1e-6 of the total) lead to the same optimum."""
    _, _, _, _, (_, _, r32) = _optimize_case(tmp_path, 2)
    _, _, _, _, (_, _, r64) = _optimize_case(tmp_path, 0)
    assert abs(r32[1] - r64[1]) <= 1e-6 * abs(r64[1])
    assert abs(r32[2][0] - r64[2][0]) <= 1e-3 * r64[2][0]


def test_lambda_epsilon_search_through_cpp_interface(tmp_path):
    """lambda_epsilon_optimizer (optimizer_scorer.cpp:72-108) on PLAIN64 over
    a device-resident model: the search visits the (lambda, epsilon) points
    the oracle's Nelder-Mead visits, each scored by the oracle with the error
    model replace_epsilons leaves ({0, 1-e, e} at size 0, {e, 1-2e, e} above),
    and finalize puts the best epsilon into every row."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import nelder_mead as nmr

    w = wl.config_base(seed=6, F=100)
    B = R = 50
    counts = np.minimum(w.counts, B)
    f = tmp_path / "opt_eps.txt"
    write_pvalue_input(f, w.tree, counts, 0.002, B, R, seed=31)
    rc, out, err = run("optimize_eps", str(f), "0.05", "0", timeout=300)
    assert rc == 0, err
    start, evals, result = _parse_optimize(out)
    final = [float(x) for x in [l for l in out.splitlines() if l.startswith("FINAL")][0].split()[1:]]
    prior = Oracle.prior_uniform(R, R)
    assert len(start) == 2 and start[1] == 0.05

    def oracle_score(x):
        if not x[0] > 0:                                  # lambda.h: an invalid lambda scores +inf
            return float("inf")
        v = Oracle.evaluate(w.tree, np.full(w.tree.n_nodes, x[0]), counts, B, R, prior,
                            err=wl.error_table(x[1], B), threads=8)["total"]
        return float("inf") if np.isnan(v) else v

    seen = []

    def logged(x):
        seen.append((x[0], x[1], oracle_score(x)))
        return seen[-1][2]

    vals, score, iters = nmr.optimize(logged, start)
    assert np.array_equal(evals[0][:2], start)
    got = evals[1:]
    want = np.array(seen)
    assert len(got) == len(want), (len(got), len(want))
    assert np.array_equal(got[:, :2], want[:, :2])
    fin = np.isfinite(want[:, 2])
    assert np.array_equal(np.isfinite(got[:, 2]), fin)
    assert np.all(np.abs(got[fin, 2] - want[fin, 2]) <= 1e-9 * np.abs(want[fin, 2]))
    assert result[0] == iters and result[2] == vals
    assert abs(result[1] - score) <= 1e-9 * abs(score)
    # finalize: the best lambda, one epsilon, rows {e, 1-2e, e}
    assert final[0] == vals[0] and final[1] == vals[1]
    assert final[2:] == [vals[1], 1 - 2 * vals[1], vals[1]]


def test_lambda_alpha_search_through_cpp_interface(tmp_path):
    """gamma_lambda_optimizer (optimizer_scorer.cpp:111-180) on PLAIN64 over
    a 4-category gamma model: the search visits the (lambda, alpha) points
    the oracle's Nelder-Mead visits, each scored by the oracle with the
    categories get_gamma gives for alpha (gamma_core.cpp:61-67) and the
    gamma model's validity rules (can_infer, gamma_core.cpp:123-141), and
    finalize leaves the best lambda and alpha in the model."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import nelder_mead as nmr

    K = 4
    w = wl.config_gamma(seed=3, F=100, K=K, bound=50)
    B = R = 50
    f = tmp_path / "opt_gamma.txt"
    write_pvalue_input(f, w.tree, w.counts, 0.003, B, R, seed=12)
    rc, out, err = run("optimize_gamma", str(f), str(K), "0", timeout=300)
    assert rc == 0, err
    start, evals, result = _parse_optimize(out)
    final = [float(x) for x in [l for l in out.splitlines() if l.startswith("FINAL")][0].split()[1:]]
    prior = Oracle.prior_uniform(R, R)
    longest = float(np.max(w.tree.length))

    def oracle_score(x):
        lam, alpha = x
        if not lam > 0 or alpha < 0:
            return float("inf")
        a = lam * longest / (1 + lam * longest)           # matrix_cache::is_saturated
        if 1 - 2 * a < 0:
            return float("inf")
        fq, r = Oracle.get_gamma(K, alpha)
        v = Oracle.evaluate(w.tree, np.full(w.tree.n_nodes, lam), w.counts, B, R, prior, n_cat=K, multipliers=r,
                            cat_probs=fq, threads=8)["total"]
        return float("inf") if np.isnan(v) else v

    seen = []

    def logged(x):
        seen.append((x[0], x[1], oracle_score(x)))
        return seen[-1][2]

    vals, score, iters = nmr.optimize(logged, start)
    assert np.array_equal(evals[0][:2], start)
    got = evals[1:]
    want = np.array(seen)
    assert len(got) == len(want), (len(got), len(want))
    assert np.array_equal(got[:, :2], want[:, :2])
    fin = np.isfinite(want[:, 2])
    assert np.array_equal(np.isfinite(got[:, 2]), fin)
    assert np.all(np.abs(got[fin, 2] - want[fin, 2]) <= 1e-9 * np.abs(want[fin, 2]))
    assert result[0] == iters and result[2] == vals
    assert abs(result[1] - score) <= 1e-9 * abs(score)
    assert final[:2] == vals
    assert np.array_equal(np.array(final[2:]), Oracle.get_gamma(K, vals[1])[1])


# --------------------------------------------------------------------------
# CAFE5's example files (tests/golden/examples, from the reference's
# examples/) read through the C++ readers and evaluated on PLAIN64
# --------------------------------------------------------------------------

@pytest.mark.parametrize("case", ["plain", "error_model"])
def test_example_files_through_cpp_interface(case):
    import json
    ex = os.path.join(ROOT, "tests", "golden", "examples")
    meta = json.load(open(os.path.join(ex, "expected.json")))
    args = [os.path.join(ex, "mammals_tree.txt"), os.path.join(ex, "mammal_gene_families.txt"),
            repr(meta["lambda"]), str(meta["B"]), str(meta["R"])]
    if case == "error_model":
        args.append(os.path.join(ex, "errormodel_0.1.txt"))
    rc, out, err = run("example", *args)
    assert rc == 0, err
    total, lnl, _ = parse(out)
    ids = [l.split()[1] for l in out.splitlines() if l.startswith("LNL")]
    want = meta["cases"][case]
    assert ids == meta["ids"]
    wl_ = np.array([float(v) for v in want["lnl"]])
    assert np.all(np.abs(lnl - wl_) <= 1e-9 * np.abs(wl_)), np.max(np.abs(lnl - wl_) / np.abs(wl_))
    assert abs(total - float(want["total"])) <= 1e-9 * abs(float(want["total"]))
