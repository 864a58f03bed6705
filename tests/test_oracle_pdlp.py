"""Pins of the oracle's restarted PDHG (the engine's pdlp mode, which has no
reference counterpart): it must reach the relative-KKT target as measured by
the independent double-double evaluator or_kkt, land on the optimum a library
LP solver (HiGHS) reports, and reproduce known optima in closed form."""
import json
import math
import os

import numpy as np
import pytest

from nirnaya_amd.engine import LP, gen_equality_lp, gen_mixed_lp, gen_staircase_lp
from oracle_lib import Oracle
import lpkit

HIGHS = json.load(open(os.path.join(lpkit.GOLDEN, "highs_optima.json")))
MAKE = {
    "equality_20260_2000x5000": lambda: gen_equality_lp(20260, 2000, 5000, 5),
    "mixed_11_600x1800": lambda: gen_mixed_lp(11, 600, 1800, 6.0, 0.01, 1e3),
    "staircase_3_6x8": lambda: gen_staircase_lp(3, 6, 8),
    "staircase_20263_12x10": lambda: gen_staircase_lp(20263, 12, 10),
}


@pytest.mark.parametrize("name", sorted(MAKE))
def test_pdlp_oracle_reaches_highs_optimum(name):
    lp = MAKE[name]()
    o = Oracle(lp)
    x, y, info = o.pdlp(max_iters=400000, tol=1e-7)
    assert info.status == 0
    k = o.kkt(x, y)                      # independent evaluator, dd products
    assert max(k[0], k[1], k[2]) <= 2e-7
    ref = HIGHS[name]
    assert abs(k[3] - ref) <= 1e-5 * (1 + abs(ref))


def test_pdlp_oracle_closed_form_optima():
    o = Oracle(lpkit.pdhg_unit_lp())
    x, y, info = o.pdlp(tol=1e-9)
    assert info.status == 0
    assert np.allclose(x, [3.0, 1.0], atol=1e-6) and np.allclose(y, [1.5, 0.5], atol=1e-6)
    t = lpkit.transport_lp()
    x, y, info = Oracle(t).pdlp(tol=1e-9)
    assert info.status == 0 and abs(t.c @ x - lpkit.transport_optimum()) <= 1e-6
    assert lpkit.transport_optimum() == 13.0


def test_pdlp_oracle_warm_start_at_optimum_stops_at_once():
    lp = lpkit.pdhg_unit_lp()
    x, y, info = Oracle(lp).pdlp(tol=1e-9, x0=np.array([3.0, 1.0]), y0=np.array([1.5, 0.5]))
    assert info.status == 0 and info.iterations == 0


def test_pdlp_oracle_units_follow_unscale_vectors():
    # KKT reported in caller units: feeding R A C with unscale vectors (R, C)
    # must give the same relative KKT as the unscaled problem's evaluator.
    lp = gen_equality_lp(5, 80, 200, 4)
    rng = np.random.default_rng(2)
    R = np.ldexp(1.0, rng.integers(-3, 4, lp.m))
    Cs = np.ldexp(1.0, rng.integers(-3, 4, lp.n))
    A = lp.dense()
    S = LP.from_dense(R[:, None] * A * Cs[None, :], lp.c * Cs, lp.l / Cs, lp.u / Cs, lp.rl * R, lp.ru * R)
    xs, ys, info = Oracle(S).pdlp(tol=1e-8, unscale_row=R, unscale_col=Cs)
    assert info.status == 0
    k = Oracle(lp).kkt(xs * Cs, ys * R)
    assert max(k[0], k[1], k[2]) <= 1.5e-8
    assert info.rel_p == pytest.approx(k[0], rel=1e-3, abs=1e-14)


def _linprog_status(lp):
    """scipy's HiGHS on the same LP: 0 optimal, 2 infeasible, 3 unbounded."""
    from scipy.optimize import linprog
    import scipy.sparse as sp
    A = lpkit._csc(lp).tocsr()
    fin = lambda v: np.isfinite(v) & (np.abs(v) < 1e20)
    ub_rows = [i for i in range(lp.m) if fin(lp.ru[i])]
    lb_rows = [i for i in range(lp.m) if fin(lp.rl[i])]
    A_ub = sp.vstack([A[ub_rows], -A[lb_rows]])
    b_ub = np.concatenate([lp.ru[ub_rows], -lp.rl[lb_rows]])
    bounds = [(lp.l[j] if fin(lp.l[j]) else None, lp.u[j] if fin(lp.u[j]) else None) for j in range(lp.n)]
    return linprog(lp.c, A_ub=A_ub, b_ub=b_ub, bounds=bounds, method="highs").status


@pytest.mark.parametrize("name", sorted(lpkit.infeasible_lps()))
def test_pdlp_oracle_certifies_infeasible_and_unbounded(name):
    # A dual ray (status 21) must come only from LPs scipy's HiGHS finds
    # infeasible, a primal ray (22) only from unbounded ones; a sign error in
    # either ray objective or a dropped violation term breaks the pairing.
    kind, lp = lpkit.infeasible_lps()[name]
    x, y, info = Oracle(lp).pdlp(max_iters=200000, tol=1e-6)
    want = {"dual": (21, 2), "primal": (22, 3)}[kind]
    assert info.status == want[0], (name, info.status, info.iterations)
    assert _linprog_status(lp) == want[1]
    assert info.iterations < 20000


@pytest.mark.parametrize("name", sorted(lpkit.long_line_infeasible_lps()))
def test_pdlp_oracle_certificates_through_long_lines(name):
    # The same pairing with HiGHS where the rays' products run over a row and
    # a column of more than 4,096 nonzeros (the chunked dd order).
    kind, lp = lpkit.long_line_infeasible_lps()[name]
    x, y, info = Oracle(lp).pdlp(max_iters=100000, tol=1e-6)
    want = {"dual": (21, 2), "primal": (22, 3)}[kind]
    assert info.status == want[0], (name, info.status, info.iterations)
    assert _linprog_status(lp) == want[1]


@pytest.mark.parametrize("name", ["transport", "unit11", "staircase"])
def test_pdlp_oracle_no_certificate_on_feasible_lps(name):
    lp = {"transport": lpkit.transport_lp, "unit11": lpkit.pdhg_unit_lp,
          "staircase": lambda: gen_staircase_lp(3, 6, 8)}[name]()
    assert _linprog_status(lp) == 0
    _, _, info = Oracle(lp).pdlp(max_iters=400000, tol=1e-9)
    assert info.status == 0


def _row_sum(val, idx, v):
    from oracle_lib import oracle, _p
    import ctypes as C
    val = np.ascontiguousarray(val, dtype=np.float64)
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    v = np.ascontiguousarray(v, dtype=np.float64)
    return oracle().or_pdlp_row_sum(_p(val), _p(idx, C.c_int32), val.size, _p(v))


# Row of 1.0 followed by len-1 terms of u = 2^-60. A term or a chunk sum of
# 64 u = 2^-54 added to 1.0 rounds away (half an ulp of 1.0 is 2^-53), while
# 32 chunk sums added from 0.0 give 2^-49 exactly. So the sequential order
# returns 1.0, and the chunked order returns 1 + (groups after the first)
# * 2^-49: the first group (positions 0..2047) is lost against the 1.0. A
# chunk or group of another size, chunk sums added straight into the row sum,
# or the threshold moved by one each give another value.
@pytest.mark.parametrize("length,expect", [
    (100, 1.0),
    (4096, 1.0),                       # the longest sequential row
    (4097, 1.0 + 2.0 ** -49),          # group 1 full, group 2 = one lost term
    (6144, 1.0 + 2.0 ** -48),          # groups 1 and 2 full
    (2048 * 9, 1.0 + 8 * 2.0 ** -49),
])
def test_pdlp_row_sum_order_closed_form(length, expect):
    u = 2.0 ** -60
    val = np.full(length, u)
    val[0] = 1.0
    assert _row_sum(val, np.arange(length), np.ones(length)) == expect


def test_pdlp_row_sum_gathers_by_index():
    # Integer products are exact in any order: the result is the exact dot
    # product of the gathered entries, short and long rows alike.
    rng = np.random.default_rng(5)
    v = rng.integers(-1000, 1000, 20000).astype(np.float64)
    for length in (0, 1, 63, 4096, 4097, 9000):
        idx = rng.choice(v.size, length, replace=False)
        val = rng.integers(-50, 50, length).astype(np.float64)
        assert _row_sum(val, idx, v) == float(np.dot(val.astype(np.int64), v[idx].astype(np.int64)))
