import csv
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# fixture run: configuration.txt of the reference (4x4 subdomains, 8 knots, [-5,5]^2, N=5000, h0=0.005)
FIX_SW, FIX_NE, FIX_NSUB, FIX_NNODE = (-5.0, -5.0), (5.0, 5.0), 4, 8
FIX_FAC = [1.5, 1.27, 1.2, 1.13]
FIX_N, FIX_H = 5000, 0.005


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def golden():
    G, Gv = {}, {}
    with open(os.path.join(GOLDEN, "G.csv")) as f:
        for r in csv.DictReader(f):
            G[(int(r["G_i"]), int(r["G_j"]))] = float(r["G_ij"])
    with open(os.path.join(GOLDEN, "G_var.csv")) as f:
        for r in csv.DictReader(f):
            Gv[(int(r["Gvar_i"]), int(r["Gvar_j"]))] = float(r["Gvar_ij"])
    with open(os.path.join(GOLDEN, "B.csv")) as f:
        B = {int(r["B_i"]): float(r["B_ij"]) for r in csv.DictReader(f)}
    with open(os.path.join(GOLDEN, "B_var.csv")) as f:
        Bv = {int(r["B_i"]): float(r["B_ij"]) for r in csv.DictReader(f)}
    with open(os.path.join(GOLDEN, "solution.csv")) as f:
        sol = list(csv.DictReader(f))
    cond = []
    with open(os.path.join(GOLDEN, "cond.txt")) as f:
        for line in f:
            p = line.strip().split(",")
            if len(p) == 4:  # a few lines were interleaved by concurrent writers
                cond.append((p[0], int(p[1]), float(p[2]), int(p[3])))
    rows = {}
    for (i, j) in G:
        if i != j:
            rows.setdefault(i, []).append(j)
    return {"G": G, "Gvar": Gv, "B": B, "Bvar": Bv, "solution": sol, "cond": cond,
            "rows": {k: sorted(v) for k, v in rows.items()}}


def zscores(a, va, na, b, vb, nb):
    a, va, b, vb = (np.asarray(t, dtype=np.float64) for t in (a, va, b, vb))
    return (a - b) / np.sqrt(np.maximum(va, 0) / na + np.maximum(vb, 0) / nb)
