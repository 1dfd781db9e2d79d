#!/usr/bin/env python3
"""Writes tests/golden/examples/: data files from the CAFE5 tree's examples/
directory (the mammal tree, a subset of its gene-family table, its error
model) and what the compiled reference (oracle/_ref/libcaferef.so) computes
from them: the base model's per-family lnL read through the reference's own
file readers.  These are data fixtures; no reference source is kept.

    python tests/golden/make_examples.py [/path/to/cafe5]
"""
from __future__ import annotations

import ctypes
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle_lib import REF, Ref  # noqa: E402

OUT = os.path.join(HERE, "examples")
LAMBDA = 0.0017
B = R = 100
N_FAMILIES = 250
MAX_COUNT = 80   # the error model's rows stop at 90


def main():
    src = os.path.join(sys.argv[1] if len(sys.argv) > 1 else "/root/reference", "examples")
    if not Ref.available():
        sys.exit("oracle/_ref/libcaferef.so is not built (make -C oracle ref)")
    os.makedirs(OUT, exist_ok=True)
    tree = open(os.path.join(src, "mammals_tree.txt")).read().strip()
    with open(os.path.join(OUT, "mammals_tree.txt"), "w") as fh:
        fh.write(tree + "\n")
    lines = open(os.path.join(src, "mammal_gene_families.txt")).read().splitlines()
    keep = [lines[0]]
    for line in lines[1:]:
        cells = line.split("\t")
        if cells and max(int(c) for c in cells[2:]) <= MAX_COUNT:
            keep.append(line)
        if len(keep) > N_FAMILIES:
            break
    fam_path = os.path.join(OUT, "mammal_gene_families.txt")
    with open(fam_path, "w") as fh:
        fh.write("\n".join(keep) + "\n")
    err_path = os.path.join(OUT, "errormodel_0.1.txt")
    with open(os.path.join(src, "errormodel_0.1.txt")) as fin, open(err_path, "w") as fout:
        fout.write(fin.read())

    REF.ref_example.restype = ctypes.c_double
    REF.ref_example.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_double, ctypes.c_int,
                                ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                ctypes.POINTER(ctypes.c_double)]
    cases = {}
    for name, ep in (("plain", b""), ("error_model", err_path.encode())):
        out = np.zeros(N_FAMILIES + 8)
        n = ctypes.c_int()
        total = REF.ref_example(tree.encode(), fam_path.encode(), ep, LAMBDA, B, R, len(out), ctypes.byref(n),
                                out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        if n.value <= 0:
            sys.exit("ref_example failed")
        cases[name] = {"total": repr(total), "lnl": [repr(float(v)) for v in out[: n.value]]}
    ids = [l.split("\t")[1] for l in keep[1:]]
    meta = {"source": "CAFE5 examples/ read and evaluated by the compiled reference (oracle/_ref)",
            "lambda": LAMBDA, "B": B, "R": R, "prior": "uniform over 1..R", "ids": ids, "cases": cases}
    with open(os.path.join(OUT, "expected.json"), "w") as fh:
        json.dump(meta, fh, indent=0)
    print(f"{len(ids)} families; totals", {k: v["total"] for k, v in cases.items()})


if __name__ == "__main__":
    main()
