"""Writes tests/golden/ref_pdhg_k5000.npz: the iterates the reference's own
PdhgSolver::solve (oracle/_ref/libref.so, compiled from the reference sources
by oracle/Makefile) returns after its minimum cap of 5000 iterations on a
seeded 40 x 100 equality LP. Run from the repository root:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from nirnaya_amd.engine import gen_equality_lp  # noqa: E402
from oracle_lib import Reference  # noqa: E402


def main():
    lp = gen_equality_lp(424242, 40, 100, 3)
    x, y, st, it, obj = Reference(lp).pdhg(simplex_iter_limit=50000)
    assert st == 0
    np.savez(os.path.join(HERE, "ref_pdhg_k5000.npz"), m=lp.m, n=lp.n, start=lp.start, index=lp.index,
             value=lp.value, c=lp.c, l=lp.l, u=lp.u, rl=lp.rl, ru=lp.ru, x=x, y=y, iters=it, objective=obj)
    print("iterations", it, "objective", obj)


if __name__ == "__main__":
    main()
