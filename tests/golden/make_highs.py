"""Writes tests/golden/highs_optima.json: optimal objectives of seeded
instances computed by HiGHS (scipy.optimize.linprog, method="highs"), a
library LP solver independent of this code. Run from the repository root:
    python tests/golden/make_highs.py
"""
import json
import os
import sys

import numpy as np
import scipy.sparse as sp
from scipy.optimize import linprog

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from nirnaya_amd.engine import gen_equality_lp, gen_mixed_lp, gen_staircase_lp  # noqa: E402

INSTANCES = {
    "equality_20260_2000x5000": lambda: gen_equality_lp(20260, 2000, 5000, 5),
    "mixed_11_600x1800": lambda: gen_mixed_lp(11, 600, 1800, 6.0, 0.01, 1e3),
    "staircase_3_6x8": lambda: gen_staircase_lp(3, 6, 8),
    "staircase_20263_12x10": lambda: gen_staircase_lp(20263, 12, 10),
}


def solve(lp):
    A = sp.csc_matrix((lp.value, lp.index, lp.start), shape=(lp.m, lp.n)).tocsr()
    eq = lp.rl == lp.ru
    A_eq, b_eq = A[eq], lp.rl[eq]
    up = np.isfinite(lp.ru) & ~eq
    lo = np.isfinite(lp.rl) & ~eq
    A_ub = sp.vstack([A[up], -A[lo]]).tocsr()
    b_ub = np.concatenate([lp.ru[up], -lp.rl[lo]])
    bounds = list(zip(np.where(np.isfinite(lp.l), lp.l, None), np.where(np.isfinite(lp.u), lp.u, None)))
    r = linprog(lp.c, A_ub=A_ub, b_ub=b_ub, A_eq=A_eq, b_eq=b_eq, bounds=bounds, method="highs")
    assert r.status == 0, r.message
    return float(r.fun)


def main():
    out = {name: solve(make()) for name, make in INSTANCES.items()}
    with open(os.path.join(HERE, "highs_optima.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(out)


if __name__ == "__main__":
    main()
