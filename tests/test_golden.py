"""The oracle against the compiled reference's own outputs, stored as fixtures
in tests/golden/ref_vectors.json (written by tests/golden/make_golden.py from
oracle/_ref).  These checks run wherever the oracle builds, also where the
CAFE5 sources are not available; every comparison is bitwise."""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

from cafe5_amd import workloads as wl
from oracle_lib import Oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "ref_vectors.json")))


def _tree(d) -> wl.Tree:
    return wl.Tree(np.array(d["parent"], dtype=np.int32), np.array(d["leaf_col"], dtype=np.int32),
                   np.array(d["length"], dtype=np.float64), list(d["names"]))


@pytest.mark.parametrize("case", GOLD["matrices"], ids=lambda c: f"l{c['lambda']}_t{c['t']}_N{c['N']}")
def test_golden_matrices(case):
    lq, tq = Oracle.key(case["lambda"], case["t"])
    assert np.array_equal(Oracle.matrix(lq, tq, case["N"]), np.array(case["P"]))


def test_golden_prune():
    tree = _tree(GOLD["prune"]["tree"])
    for c in GOLD["prune"]["cases"]:
        err = None if c["error_model"] is None else np.array(c["error_model"])
        got = Oracle.prune(tree, np.array(c["counts"]), np.full(tree.n_nodes, c["lambda"]), c["B"], c["R"], err)
        assert np.array_equal(got, np.array(c["root"]))


def test_golden_prune_heap_path():
    g = GOLD["prune_heap"]
    tree = _tree(g["tree"])
    got = Oracle.prune(tree, np.array(g["counts"]), np.full(tree.n_nodes, g["lambda"]), g["B"], g["R"])
    assert np.array_equal(got, np.array(g["root"]))


def test_golden_base_uniform():
    g = GOLD["base_uniform"]
    tree = _tree(g["tree"])
    prior = np.array(g["prior"])
    assert np.array_equal(prior, Oracle.prior_uniform(g["R"], g["R"]))
    res = Oracle.evaluate(tree, np.full(tree.n_nodes, g["lambda"]), np.array(g["counts"]), g["B"], g["R"], prior)
    assert np.array_equal(res["lnl"], np.array(g["lnl"]))
    assert res["total"] == g["total"]


def test_golden_base_multilambda_error_poisson():
    g = GOLD["base_multi_error_poisson"]
    tree = _tree(g["tree"])
    prior = np.array(g["prior"])
    assert np.array_equal(prior, Oracle.prior_poisson(3.5, g["R"], g["R"]))
    res = Oracle.evaluate(tree, np.array(g["node_lambda"]), np.array(g["counts"]), g["B"], g["R"], prior,
                          err=np.array(g["error_model"]))
    assert np.array_equal(res["lnl"], np.array(g["lnl"]))
    assert res["total"] == g["total"]


def test_golden_gamma():
    g = GOLD["gamma"]
    tree = _tree(g["tree"])
    f, r = Oracle.get_gamma(4, g["alpha"])
    assert np.array_equal(f, np.array(g["cat_probs"])) and np.array_equal(r, np.array(g["multipliers"]))
    res = Oracle.evaluate(tree, np.full(tree.n_nodes, g["lambda"]), np.array(g["counts"]), g["B"], g["R"],
                          np.array(g["prior"]), n_cat=4, multipliers=r, cat_probs=f)
    assert np.array_equal(res["cat"], np.array(g["cat"]))
    assert np.array_equal(res["fam"], np.array(g["fam"]))
    assert np.array_equal(res["post"], np.array(g["post"]))
    assert res["total"] == g["total"]


def test_golden_discrete_gamma():
    for c in GOLD["discrete_gamma"]:
        f, r = Oracle.get_gamma(c["K"], c["alpha"])
        assert np.array_equal(f, np.array(c["probs"])), c["K"]
        assert np.array_equal(r, np.array(c["multipliers"])), (c["K"], c["alpha"])


def test_golden_discrete_gamma_python_package():
    # the categories the Python package's gamma models use (the C++ mirror's
    # get_gamma through cafe5_amd.model)
    from cafe5_amd.model import get_gamma
    for c in GOLD["discrete_gamma"]:
        f, r = get_gamma(c["K"], c["alpha"])
        assert np.array_equal(f, np.array(c["probs"])), c["K"]
        assert np.array_equal(r, np.array(c["multipliers"])), (c["K"], c["alpha"])


def test_golden_priors():
    pr = GOLD["priors"]
    for c in pr["uniform"]:
        assert np.array_equal(Oracle.prior_uniform(c["max"], c["R"]), np.array(c["p"]))
    for c in pr["poisson"]:
        assert np.array_equal(Oracle.prior_poisson(c["lambda"], c["max"], c["R"]), np.array(c["p"]))
    m = pr["map"]
    assert np.array_equal(Oracle.prior_map(m["sizes"], m["counts"], m["R"]), np.array(m["p"]))


def test_golden_priors_python_package():
    # the priors the Python package hands the engine (cafe5_amd/priors.py)
    from cafe5_amd import priors
    pr = GOLD["priors"]
    for c in pr["uniform"]:
        assert np.array_equal(priors.uniform(c["max"], c["R"]), np.array(c["p"]))
    for c in pr["poisson"]:
        assert np.array_equal(priors.poisson(c["lambda"], c["max"], c["R"]), np.array(c["p"]))
    m = pr["map"]
    assert np.array_equal(priors.from_counts(dict(zip(m["sizes"], m["counts"])), m["R"]), np.array(m["p"]))
