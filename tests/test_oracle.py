"""Pins of the C oracle (CPU only).

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong sign or index, a transposed operand) fails it: reference
known-answer tests, closed forms, brute force, the reference's own fixtures,
and - where oracle/_ref was built from the reference sources - bitwise
agreement with the reference code itself.
"""
import math

import numpy as np
import pytest

from nirnaya_amd.engine import LP, gen_equality_lp, gen_mixed_lp, gen_staircase_lp
from oracle_lib import Oracle, Reference, have_ref
import lpkit

needs_ref = pytest.mark.skipif(not have_ref(), reason="oracle/_ref not built (reference sources absent)")


def test_spmv_known_answer():
    # unit_tests.cpp:45-60: duplicates (0,0)=1.5+0.5 merge to 2; A (2,3) = (16, 0).
    lp = LP.from_dense([[2.0, 4.0], [-3.0, 2.0]], [0, 0], [0, 0], [1, 1], [0, 0], [0, 0])
    o = Oracle(lp)
    assert o.spmv(np.array([2.0, 3.0])).tolist() == [16.0, 0.0]
    # transpose on the same matrix: A^T (1, 1) = (-1, 6)
    assert o.spmv_t(np.array([1.0, 1.0])).tolist() == [-1.0, 6.0]


def test_spmv_brute_force_integer_exact():
    # Integer data: every summation order is exact, so dense products pin both
    # directions bit for bit (a transposed operand or wrong index fails).
    rng = np.random.default_rng(3)
    A = rng.integers(-5, 6, size=(7, 11)).astype(float) * (rng.random((7, 11)) < 0.4)
    lp = LP.from_dense(A, np.zeros(11), np.zeros(11), np.ones(11), np.zeros(7), np.zeros(7))
    x = rng.integers(-9, 10, size=11).astype(float)
    y = rng.integers(-9, 10, size=7).astype(float)
    o = Oracle(lp)
    assert np.array_equal(o.spmv(x), A @ x)
    assert np.array_equal(o.spmv_t(y), A.T @ y)


def test_step_size_closed_form():
    # LP of unit_tests.cpp:316-331: lambda_max(A^T A) = 6 + sqrt(32), so
    # sigma_max = 2 + sqrt(2) and tau = sigma = 0.85 / (2 + sqrt(2)).
    lp = lpkit.pdhg_unit_lp()
    _, _, tau, _, _ = Oracle(lp).pdhg(1, check=False)
    assert tau == pytest.approx(0.85 / (2.0 + math.sqrt(2.0)), rel=1e-15)
    assert tau == pytest.approx(0.2489592359914346, rel=1e-15)


def test_first_step_by_hand():
    # x0 = bound midpoint (10, 10), y0 = 0. One step:
    #   x1 = (10 - 2 tau, 10 - 3 tau), xbar = 2 x1 - x0, h = A xbar,
    #   y1_i = sigma (clamp(h_i) - h_i), V = max row violation of h.
    lp = lpkit.pdhg_unit_lp()
    x, y, tau, it, V = Oracle(lp).pdhg(1, check=False)
    t = 0.85 / (2.0 + math.sqrt(2.0))
    x1 = np.array([10 - 2 * t, 10 - 3 * t])
    xbar = 2 * x1 - 10
    h = np.array([xbar[0] + xbar[1], xbar[0] + 3 * xbar[1]])
    assert np.allclose(x, x1, rtol=1e-15, atol=0)
    assert np.allclose(x, [9.502081528017131, 9.253122292025695], rtol=1e-15)
    assert np.allclose(h, [17.510407640085653, 34.522896808188435], rtol=1e-14)
    assert y[0] == 0.0 and y[1] == pytest.approx(t * (30.0 - h[1]), rel=1e-13)
    assert y[1] == pytest.approx(-1.1260169338346908, rel=1e-13)
    assert V == pytest.approx(h[1] - 30.0, rel=1e-13)


def test_optimum_is_a_fixed_point():
    # (x*, y*) = ((3,1), (1.5,0.5)) satisfies the PDHG fixed-point equations.
    lp = lpkit.pdhg_unit_lp()
    x, y, _, _, V = Oracle(lp).pdhg(5, check=False, x0=np.array([3.0, 1.0]), y0=np.array([1.5, 0.5]))
    assert np.allclose(x, [3.0, 1.0], atol=1e-14)
    assert np.allclose(y, [1.5, 0.5], atol=1e-14)
    assert V == 0.0


def test_unit_lp_converges_like_reference_test():
    # unit_tests.cpp:329: PDHG objective 9 +- 1e-2 with status Optimal.
    lp = lpkit.pdhg_unit_lp()
    x, _, _, it, V = Oracle(lp).pdhg(50000, check=True)
    assert abs(2 * x[0] + 3 * x[1] - 9.0) <= 1e-2
    assert it < 50000 and V <= 1e-6


def test_dd_checker_cancellation_and_fixture():
    o_lp = LP.from_dense([[1.0, 1.0, 1.0]], [0, 0, 0], [-1e30] * 3, [1e30] * 3, [1.0], [1.0])
    ok, rv, _, _ = Oracle(o_lp).verify(np.array([1e16, 1.0, -1e16]), 0.0)
    # dd_real KAT (unit_tests.cpp:29-42): 1e16 + 1 - 1e16 == 1 exactly.
    assert ok and rv == 0.0
    # README.md:84: the crude-blending solution has max row violation 2.55e-14
    # and objective 22171.59777424483 (.sol:2) in the user's (max) sense.
    lp, sense, cols = lpkit.read_mps_lp(lpkit.CRUDE_MPS)
    claimed, vals = lpkit.read_sol(lpkit.CRUDE_SOL)
    x = np.array([vals[c] for c in cols])
    ok, rv, bv, obj = Oracle(lp).verify(x, claimed)
    assert ok and sense == -1
    assert f"{rv:.2e}" == "2.55e-14"
    assert obj == pytest.approx(22171.59777424483, rel=1e-15)


def test_kkt_zero_at_known_optimum():
    # A.5 LP (unit_tests.cpp:116-142): x=(3,1), y=(1.5,0.5) is optimal.
    lp = LP.from_dense([[1.0, 1.0], [1.0, 3.0]], [2.0, 3.0], [0, 0], [math.inf] * 2, [4.0, 6.0], [math.inf] * 2)
    k = Oracle(lp).kkt(np.array([3.0, 1.0]), np.array([1.5, 0.5]))
    assert k[0] == 0.0 and k[1] == 0.0 and k[2] == 0.0
    assert k[3] == 9.0 and k[4] == 9.0
    # A perturbed point is not optimal: each term reacts to its own defect.
    assert Oracle(lp).kkt(np.array([3.0, 0.9]), np.array([1.5, 0.5]))[0] > 0
    assert Oracle(lp).kkt(np.array([3.0, 1.0]), np.array([1.5, 0.6]))[1] > 0
    assert Oracle(lp).kkt(np.array([3.0, 1.0]), np.array([-1.5, 0.5]))[1] > 0   # sign of a >= row dual


def test_scaling_invariants():
    lp = gen_mixed_lp(5, 300, 900, 6.0, 0.01, 1e3)
    active, R, Cs = Oracle(lp).scale()
    assert active
    for v in np.concatenate([R, Cs]):
        m, e = math.frexp(v)
        assert m == 0.5 and -29 <= e <= 31          # exact powers of two in [2^-30, 2^30]
    # Equilibration moves every row's and column's max/min geometric mean
    # toward 1 compared with the unscaled matrix.
    A = lp.dense()
    S = (R[:, None] * A) * Cs[None, :]
    def spread(M):
        a = np.abs(M)
        a[a == 0] = np.nan
        return np.nanmax(np.abs(np.log2(np.sqrt(np.nanmax(a, 1) * np.nanmin(a, 1)))))
    assert spread(S) < spread(A)
    # Inactive when max|a| / min|a| <= 10.
    small = LP.from_dense([[1.0, 2.0], [3.0, 4.0]], [0, 0], [0, 0], [1, 1], [0, 0], [1, 1])
    assert Oracle(small).scale()[0] is False


@needs_ref
@pytest.mark.parametrize("make", [
    lambda: gen_equality_lp(20260, 2000, 5000, 5),
    lambda: gen_mixed_lp(11, 400, 1200, 6.0, 0.01, 1e3),
    lambda: gen_staircase_lp(3, 4, 5),
])
def test_oracle_matches_reference_code(make):
    lp = make()
    o, r = Oracle(lp), Reference(lp)
    rng = np.random.default_rng(0)
    x = rng.standard_normal(lp.n)
    y = rng.standard_normal(lp.m)
    assert np.array_equal(o.spmv(x), r.spmv(x))
    assert np.array_equal(o.spmv_t(y), r.spmv_t(y))
    a1, R1, C1 = o.scale()
    a2, R2, C2 = r.scale()
    assert a1 == a2 and np.array_equal(R1, R2) and np.array_equal(C1, C2)
    # Full reference loop at its minimum cap (5000 iterations), stop rule on.
    xo, yo, _, it_o, _ = o.pdhg(5000, check=True)
    xr, yr, st, it_r, _ = r.pdhg(simplex_iter_limit=50000)
    assert st == 0 and it_o == it_r
    assert np.array_equal(xo, xr) and np.array_equal(yo, yr)


@needs_ref
def test_oracle_checker_matches_reference_checker():
    lp = gen_equality_lp(1, 300, 700, 5)
    x = lp.x0 + 1e-7 * np.random.default_rng(1).standard_normal(lp.n)
    ok_o, rv_o, bv_o, obj_o = Oracle(lp).verify(x, 0.0)
    ok_r, rv_r, bv_r, obj_r = Reference(lp).verify(x, 0.0)
    assert (ok_o, rv_o, bv_o, obj_o) == (ok_r, rv_r, bv_r, obj_r)


def test_golden_reference_iterates():
    """Iterates of the reference's own PdhgSolver::solve (5000 iterations on a
    seeded 40 x 100 LP), written by tests/golden/make_golden.py."""
    import os
    path = os.path.join(lpkit.GOLDEN, "ref_pdhg_k5000.npz")
    g = np.load(path, allow_pickle=False)
    lp = LP(int(g["m"]), int(g["n"]), g["start"], g["index"], g["value"], g["c"], g["l"], g["u"], g["rl"], g["ru"])
    x, y, _, it, _ = Oracle(lp).pdhg(5000, check=True)
    assert it == int(g["iters"])
    assert np.array_equal(x, g["x"]) and np.array_equal(y, g["y"])


def test_chunked_reference_steps_equal_one_run():
    # The reference loop keeps no state beyond (x, y): 64-iteration chunks
    # from the start point give the uninterrupted run's iterates bit for bit.
    lp = gen_equality_lp(20260, 2000, 5000, 5)
    o = Oracle(lp)
    x1, y1, tau, _, _ = o.pdhg(640, check=False)
    x, y, tau0, _, _ = o.pdhg(0, check=False)
    assert tau0 == tau
    for _ in range(10):
        o.steps(64, tau, x, y)
    assert np.array_equal(x, x1) and np.array_equal(y, y1)
