"""The C++ host interface (cafe5_amd/host, mirroring the reference's classes)
on the GPU: the reference's known-answer tests run through it, and seeded
models evaluated through it match the oracle."""
import os
import subprocess

import numpy as np
import pytest

from cafe5_amd import workloads as wl
from oracle_lib import Oracle

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
    assert sum(l.startswith("PASS") for l in lines) >= 14
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
        assert np.all(np.abs(lnl - ref["lnl"]) <= 1e-4 + 1e-6 * np.abs(ref["lnl"]))
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
