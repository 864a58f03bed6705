"""Presolve parity with the reference's own Presolver::presolve
(src/presolve/presolve.hpp:32-345, compiled from the reference sources into
oracle/_ref): one small LP per rule, then a seeded battery of random small
models, each compared field by field - status, counts, the values of the
removed columns, the kept rows and columns, the reduced bounds and offset -
exactly (the rules are restated with the reference's order and
tolerances, so every number is the same double)."""
import math

import numpy as np
import pytest

from nirnaya_amd.engine import LP, presolve
from oracle_lib import have_ref, reference_presolve

needs_ref = pytest.mark.skipif(not have_ref(), reason="oracle/_ref not built (reference sources absent)")
INF = math.inf


def same(a, b):
    for k in ("status", "solved", "nrow", "ncol", "rows_removed", "cols_removed", "tightenings"):
        assert a[k] == b[k], (k, a[k], b[k])
    assert a["obj_offset"] == b["obj_offset"]
    for k in ("fixed_x", "col_map", "row_map", "collo", "colup", "rowlo", "rowup"):
        assert np.array_equal(a[k], b[k]), (k, a[k], b[k])


def check(lp, integer=None):
    ours, theirs = presolve(lp, integer), reference_presolve(lp, integer)
    same(ours, theirs)
    return ours


@needs_ref
def test_r4_fixed_columns():
    # x1 fixed at 2 (l == u): removed, its column moves the row bounds.
    lp = LP.from_dense([[1.0, 1.0, 0.0], [2.0, 0.0, 1.0]], [1.0, 2.0, 3.0], [2.0, 0.0, 0.0], [2.0, 5.0, 5.0],
                       [3.0, 4.0], [8.0, 9.0])
    r = check(lp)
    assert r["cols_removed"] >= 1 and r["fixed_x"][0] == 2.0


@needs_ref
def test_r2_empty_columns_each_cost_sign():
    # Column 1 has no nonzeros: cost > 0 -> its lower bound, cost < 0 (col 2)
    # -> its upper bound, cost 0 (col 3) -> clamp(0, bounds).
    A = [[1.0, 0.0, 0.0, 0.0, 1.0], [1.0, 0.0, 0.0, 0.0, -1.0]]
    lp = LP.from_dense(A, [1.0, 3.0, -2.0, 0.0, 1.0], [0.0, 1.5, -1.0, 0.5, 0.0], [4.0, 6.0, 7.0, 3.0, 4.0],
                       [1.0, -2.0], [5.0, 2.0])
    r = check(lp)
    assert list(r["fixed_x"][1:4]) == [1.5, 7.0, 0.5]
    # Unbounded: an empty column whose cost pulls toward an infinite bound.
    lp2 = LP.from_dense([[1.0, 0.0]], [0.0, -1.0], [0.0, 0.0], [1.0, INF], [0.0], [1.0])
    assert check(lp2)["status"] == 2


@needs_ref
def test_r5_redundant_rows():
    # x1 + x2 <= 100 cannot bind with x in [0, 10]^2; x1 - x2 >= -50 neither.
    lp = LP.from_dense([[1.0, 1.0], [1.0, -1.0], [1.0, 2.0]], [1.0, 1.0], [0.0, 0.0], [10.0, 10.0],
                       [-INF, -50.0, 3.0], [100.0, INF, 25.0])
    r = check(lp)
    assert r["rows_removed"] >= 2


@needs_ref
def test_r14_forcing_rows():
    # x1 + x2 <= 0 with x >= 0: the minimum activity equals the upper bound,
    # which forces both to their lower bounds. -x3 + x4 >= 8 with x3 in
    # [0, 2], x4 in [0, 8]: the maximum activity equals the lower bound,
    # which forces x3 to its lower bound and x4 to its upper one.
    A = [[1.0, 1.0, 0.0, 0.0, 1.0], [0.0, 0.0, -1.0, 1.0, 0.0], [1.0, 0.0, 1.0, 1.0, 1.0]]
    lp = LP.from_dense(A, [1.0, -1.0, 2.0, 1.0, 1.0], [0.0, 0.0, 0.0, 0.0, 0.0], [5.0, 5.0, 2.0, 8.0, 3.0],
                       [-INF, 8.0, 0.0], [0.0, INF, 30.0])
    r = check(lp)
    assert r["fixed_x"][3] == 8.0 and r["cols_removed"] >= 4


@needs_ref
def test_r7_integer_bound_tightening():
    # x1 + 2 x2 <= 7 with integer x in [0, 10]: x1 <= 7, x2 <= 3; and
    # 3 x3 - x4 >= 5 with x3 in [0, 4], x4 in [0, 6] integer: x3 >= ceil(...).
    A = [[1.0, 2.0, 0.0, 0.0], [0.0, 0.0, 3.0, -1.0], [1.0, 1.0, 1.0, 1.0]]
    lp = LP.from_dense(A, [1.0, 1.0, 1.0, 1.0], [0.0, 0.0, 0.0, 0.0], [10.0, 10.0, 4.0, 6.0],
                       [-INF, 5.0, 1.0], [7.0, INF, 40.0])
    integer = np.array([1, 1, 1, 1], dtype=np.int8)
    r = check(lp, integer)
    assert r["tightenings"] >= 2
    # The same rows with continuous columns: R7 leaves them alone.
    rc = check(lp)
    assert rc["tightenings"] < r["tightenings"]


@needs_ref
def test_r6_and_singletons_infeasible():
    lp = LP.from_dense([[2.0, 0.0], [0.0, 1.0]], [1.0, 1.0], [0.0, 0.0], [1.0, 1.0], [4.0, 0.0], [INF, 1.0])
    assert check(lp)["status"] == 1   # 2 x1 >= 4 with x1 <= 1


def random_model(rng, m, n, density):
    A = np.where(rng.random((m, n)) < density, np.round(rng.standard_normal((m, n)) * 4) / 2, 0.0)
    A[rng.random(m) < 0.15, :] = 0.0                                   # empty rows
    for i in np.nonzero(rng.random(m) < 0.2)[0]:                        # singleton rows
        A[i, :] = 0.0
        A[i, rng.integers(n)] = float(rng.choice([-2.0, -1.0, 0.5, 1.0, 3.0]))
    A[:, rng.random(n) < 0.1] = 0.0                                     # empty columns
    l = np.round(rng.uniform(-3, 2, n))
    u = l + np.round(rng.uniform(0, 6, n))
    fixed = rng.random(n) < 0.15
    u[fixed] = l[fixed]
    l[rng.random(n) < 0.1] = -INF
    u[rng.random(n) < 0.1] = INF
    act_lo, act_hi = A @ np.where(np.isfinite(l), l, -5), A @ np.where(np.isfinite(u), u, 5)
    lo, hi = np.minimum(act_lo, act_hi), np.maximum(act_lo, act_hi)
    rl = np.round(lo + rng.uniform(-2, 2, m) * (rng.random(m) < 0.7))
    ru = np.maximum(rl, np.round(hi - rng.uniform(-2, 2, m) * (rng.random(m) < 0.7)))
    rl[rng.random(m) < 0.2] = -INF
    ru[rng.random(m) < 0.2] = INF
    c = np.round(rng.standard_normal(n) * 2) / 2
    return LP.from_dense(A, c, l, u, rl, ru), (rng.random(n) < 0.3).astype(np.int8)


@needs_ref
@pytest.mark.parametrize("seed", range(40))
def test_random_models_match_reference(seed):
    rng = np.random.default_rng(1000 + seed)
    lp, integer = random_model(rng, int(rng.integers(3, 14)), int(rng.integers(3, 14)), 0.35)
    check(lp)
    check(lp, integer)
