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
    A = lpkit._dense(lp)
    fin = lambda v: np.isfinite(v) & (np.abs(v) < 1e20)
    ub_rows = [i for i in range(lp.m) if fin(lp.ru[i])]
    lb_rows = [i for i in range(lp.m) if fin(lp.rl[i])]
    A_ub = np.vstack([A[ub_rows], -A[lb_rows]])
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


@pytest.mark.parametrize("name", ["transport", "unit11", "staircase"])
def test_pdlp_oracle_no_certificate_on_feasible_lps(name):
    lp = {"transport": lpkit.transport_lp, "unit11": lpkit.pdhg_unit_lp,
          "staircase": lambda: gen_staircase_lp(3, 6, 8)}[name]()
    assert _linprog_status(lp) == 0
    _, _, info = Oracle(lp).pdlp(max_iters=400000, tol=1e-9)
    assert info.status == 0
