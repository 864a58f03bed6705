"""GPU parity of the engine's pdlp mode (restarted reflected-Halpern PDHG)
against the oracle's or_pdlp, and the mixed-precision / polish contract.

FP64: every sum of A products is sequential in ascending index order (as in
the oracle), norms are double-double and every decision uses only IEEE +, *,
/ and sqrt, so the device trajectory is expected to be bit-identical to the
oracle's; the acceptance bar is 1e-10 relative on the returned iterates with
equal iteration and restart counts.
Mixed: FP32 iterates and SpMV; the result must meet the tolerance under the
independent FP64/double-double evaluator, and its error against the FP64
oracle solution is reported."""
import json
import os

import numpy as np
import pytest

from nirnaya_amd.engine import (MODE_PDLP, PREC_MIXED, default_params, gen_equality_lp, gen_mixed_lp,
                                gen_staircase_lp)
from oracle_lib import Oracle
import lpkit

pytestmark = pytest.mark.gpu
HIGHS = json.load(open(os.path.join(lpkit.GOLDEN, "highs_optima.json")))

CASES = {
    "unit11": lpkit.pdhg_unit_lp,
    "transport": lpkit.transport_lp,
    "staircase_3_6x8": lambda: gen_staircase_lp(3, 6, 8),
    "equality_20260_2000x5000": lambda: gen_equality_lp(20260, 2000, 5000, 5),
    "mixed_11_600x1800": lambda: gen_mixed_lp(11, 600, 1800, 6.0, 0.01, 1e3),
}


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(1.0, float(np.max(np.abs(b))))) if a.size else 0.0


@pytest.mark.parametrize("name", sorted(CASES))
def test_fp64_pdlp_matches_oracle(engine, name):
    lp = CASES[name]()
    r = engine.solve(lp, default_params(mode=MODE_PDLP, max_iters=400000, tol=1e-6))
    xo, yo, info = Oracle(lp).pdlp(max_iters=400000, tol=1e-6)
    print(f"{name}: it {r.iterations}/{info.iterations} restarts {r.info.restarts}/{info.restarts} "
          f"rel x {rel(r.x, xo):.2e} y {rel(r.y, yo):.2e} exact {np.array_equal(r.x, xo) and np.array_equal(r.y, yo)}")
    assert r.status == info.status == 0
    assert r.iterations == info.iterations and r.info.restarts == info.restarts
    assert rel(r.x, xo) <= 1e-10 and rel(r.y, yo) <= 1e-10
    k = Oracle(lp).kkt(r.x, r.y)
    assert max(k[0], k[1], k[2]) <= 1.5e-6
    if name in HIGHS:
        assert abs(k[3] - HIGHS[name]) <= 1e-4 * (1 + abs(HIGHS[name]))


@pytest.mark.parametrize("name", sorted(CASES))
def test_device_scaling_factors_equal_oracle(engine, name):
    """The device's Ruiz + Pock-Chambolle factors (k_row_measure, k_factor,
    k_scale_csr) equal or_scale_pdlp's bit for bit."""
    lp = CASES[name]()
    s = engine.session(lp, default_params(mode=MODE_PDLP, max_iters=1, tol=1e-6))
    try:
        R, Cs = s.scaling()
    finally:
        s.close()
    Ro, Co = Oracle(lp).scale_pdlp()
    assert np.array_equal(R, Ro) and np.array_equal(Cs, Co)


def test_device_scaling_factors_full_size_segmented(engine):
    """The same on the 400k x 1.2M instance, with A in three column segments
    and A^T in two (the measures and the scaling walk every segment)."""
    lp = gen_mixed_lp(20261, 400_000, 1_200_000, 6.0, 0.01, 1e3)
    s = engine.session(lp, default_params(mode=MODE_PDLP, max_iters=1, tol=1e-6, tile_nnz=2048,
                                          seg_width=400_384, seg_width_t=200_704))
    try:
        R, Cs = s.scaling()
        assert s.shape()[2] == 3 and s.shape()[6] == 2
    finally:
        s.close()
    Ro, Co = Oracle(lp).scale_pdlp()
    assert np.array_equal(R, Ro) and np.array_equal(Cs, Co)


def test_segmented_pdlp_matches_oracle(engine):
    lp = gen_staircase_lp(3, 6, 8)
    r = engine.solve(lp, default_params(mode=MODE_PDLP, max_iters=400000, tol=1e-6, seg_width=37, tile_rows=64))
    xo, yo, info = Oracle(lp).pdlp(max_iters=400000, tol=1e-6)
    assert r.iterations == info.iterations and np.array_equal(r.x, xo) and np.array_equal(r.y, yo)


def test_graph_and_eager_identical(engine):
    lp = gen_staircase_lp(4, 8, 9)
    a = engine.solve(lp, default_params(mode=MODE_PDLP, max_iters=100000, use_graph=1))
    b = engine.solve(lp, default_params(mode=MODE_PDLP, max_iters=100000, use_graph=0))
    assert a.iterations == b.iterations and np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)


def test_iteration_cap_and_warm_start(engine):
    lp = gen_equality_lp(20260, 2000, 5000, 5)
    r = engine.solve(lp, default_params(mode=MODE_PDLP, max_iters=640, tol=1e-9))
    assert r.status == 6 and r.iterations == 640
    u = lpkit.pdhg_unit_lp()
    w = engine.solve(u, default_params(mode=MODE_PDLP, tol=1e-9), x0=np.array([3.0, 1.0]), y0=np.array([1.5, 0.5]))
    assert w.status == 0 and w.iterations == 0


@pytest.mark.parametrize("name", ["equality_20260_2000x5000", "staircase_3_6x8", "mixed_11_600x1800"])
def test_mixed_precision_meets_tolerance_and_polish(engine, name):
    lp = CASES[name]()
    xo, yo, info = Oracle(lp).pdlp(max_iters=400000, tol=1e-9)
    r = engine.solve(lp, default_params(mode=MODE_PDLP, precision=PREC_MIXED, max_iters=400000, tol=1e-6))
    k = Oracle(lp).kkt(r.x, r.y)
    ex = np.linalg.norm(r.x - xo) / max(1e-30, np.linalg.norm(xo))
    print(f"{name} mixed: status {r.status} it {r.iterations} (fp64 tail {r.info.polish_iterations}) "
          f"kkt {k[0]:.1e} {k[1]:.1e} {k[2]:.1e} |x-x64|/|x64| {ex:.2e} obj err {abs(k[3]-info.pobj)/(1+abs(info.pobj)):.2e}")
    assert r.status == 0
    assert max(k[0], k[1], k[2]) <= 2e-6
    p = engine.solve(lp, default_params(mode=MODE_PDLP, precision=PREC_MIXED, max_iters=400000, tol=1e-6,
                                        polish_tol=1e-8))
    kp = Oracle(lp).kkt(p.x, p.y)
    print(f"{name} polished: it {p.iterations} (fp64 tail {p.info.polish_iterations}) kkt {kp[0]:.1e} {kp[1]:.1e} {kp[2]:.1e}")
    assert p.status == 0 and p.info.polish_iterations > 0
    assert max(kp[0], kp[1], kp[2]) <= 2e-8


def test_facade_pdlp_mode_on_fixture(tmp_path):
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sol = tmp_path / "c.sol"
    rc = subprocess.run([os.path.join(root, "bin", "sovsolve"), lpkit.CRUDE_MPS, "--method=pdhg", "--pdhg-mode=pdlp",
                         "--tol=1e-9", "--log", "0", "--sol", str(sol)]).returncode
    assert rc == 0
    chk = subprocess.run([os.path.join(root, "bin", "checker"), lpkit.CRUDE_MPS, str(sol)], capture_output=True,
                         text=True)
    assert chk.returncode == 0, chk.stdout
    obj, _ = lpkit.read_sol(str(sol))
    assert obj == pytest.approx(22171.59777424483, rel=1e-6)


def test_facade_reads_binary_container(tmp_path):
    # The same solve from the binary LP container gives the same solution
    # file, bit for bit, as from the MPS text.
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sovsolve = os.path.join(root, "bin", "sovsolve")
    lpb = str(tmp_path / "crude.nlpb")
    assert subprocess.run([sovsolve, lpkit.CRUDE_MPS, "--write-lpb", lpb], capture_output=True).returncode == 0
    sols = []
    for src in (lpkit.CRUDE_MPS, lpb):
        sol = tmp_path / (os.path.basename(src) + ".sol")
        rc = subprocess.run([sovsolve, src, "--method=pdhg", "--pdhg-mode=pdlp", "--tol=1e-9", "--log", "0", "--sol",
                             str(sol)]).returncode
        assert rc == 0
        sols.append(sol.read_text())
    assert sols[0] == sols[1]


@pytest.mark.parametrize("polish", [0, 1])
def test_facade_mixed_precision_on_fixture(tmp_path, polish):
    # CLI -> facade (presolve, scaling, engine in mixed precision, unscale,
    # postsolve, dd certificate) on the reference's crude-blending LP.
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sol = tmp_path / "m.sol"
    args = [os.path.join(root, "bin", "sovsolve"), lpkit.CRUDE_MPS, "--method=pdhg", "--pdhg-mode=pdlp",
            "--precision=mixed", "--tol=1e-7", "--log", "0", "--sol", str(sol)]
    if polish:
        args += ["--set", "pdhg_polish_tol=1e-9"]
    assert subprocess.run(args).returncode == 0
    chk = subprocess.run([os.path.join(root, "bin", "checker"), lpkit.CRUDE_MPS, str(sol)], capture_output=True,
                         text=True)
    assert chk.returncode == 0, chk.stdout
    obj, _ = lpkit.read_sol(str(sol))
    assert obj == pytest.approx(22171.59777424483, rel=1e-5)


def test_python_wrapper_mixed_precision():
    from nirnaya_amd.nirnaya import NirnayaModel
    m = NirnayaModel("a5")
    a = m.add_var(2.0, 0.0, 1e20)
    b = m.add_var(3.0, 0.0, 1e20)
    m.add_constr([a, b], [1.0, 1.0], 4.0, 1e20)
    m.add_constr([a, b], [1.0, 3.0], 6.0, 1e20)
    for k, v in (("method", "pdhg"), ("pdhg_mode", "pdlp"), ("pdhg_precision", "mixed"), ("pdhg_tol", "1e-9")):
        m.set_option(k, v)
    m.optimize()
    assert m.Status == 0
    assert m.ObjVal == pytest.approx(9.0, abs=1e-6)
    assert np.allclose(m.X, [3.0, 1.0], atol=1e-5)


def _sov_status(info):
    """The oracle's ray codes in the reference's status codes."""
    if info.status == 21:
        return 1
    if info.status == 22:
        return 2 if info.rel_p <= 1e-6 else 3
    return info.status


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("name", sorted(lpkit.infeasible_lps()))
def test_infeasibility_certificates(engine, name, prec):
    kind, lp = lpkit.infeasible_lps()[name]
    p = default_params(mode=MODE_PDLP, precision=PREC_MIXED if prec == "mixed" else 0, max_iters=200000, tol=1e-6)
    r = engine.solve(lp, p)
    print(f"{name} {prec}: status {r.status} after {r.iterations} iterations, {r.info.restarts} restarts")
    assert r.status == 1 if kind == "dual" else r.status in (2, 3)
    if prec == "fp64":
        # the same trajectory and the same decision as the oracle
        _, _, info = Oracle(lp).pdlp(max_iters=200000, tol=1e-6)
        assert (r.iterations, r.info.restarts, r.status) == (info.iterations, info.restarts, _sov_status(info))


def test_python_wrapper_reports_infeasible():
    from nirnaya_amd.nirnaya import NirnayaModel
    m = NirnayaModel("infeasible")
    a = m.add_var(1.0, 0.0, 10.0)
    b = m.add_var(1.0, 0.0, 10.0)
    m.add_constr([a, b], [1.0, 1.0], 5.0, 1e20)
    m.add_constr([a, b], [1.0, 1.0], -1e20, 3.0)
    for k, v in (("method", "pdhg"), ("pdhg_mode", "pdlp"), ("presolve", "0")):
        m.set_option(k, v)
    m.optimize()
    assert m.Status == 1


def test_cfg2_mixed_precision_full_instance(engine):
    """Mixed precision on the 400k x 1.2M instance (m = 400,000, n = 1,200,000,
    ~9.6M nonzeros, seed 20261) to relative KKT 1e-6: the returned (x, y)
    is certified by the oracle's or_kkt (an independent double-double
    evaluation on the host), and its error against the FP64 engine's
    solution to the same tolerance is reported and bounded. The LP's
    optimal face is not a single point, so x may differ along it by more
    than the tolerance; the objective may differ by at most the two gaps."""
    lp = gen_mixed_lp(20261, 400_000, 1_200_000, 6.0, 0.01, 1e3)
    o = Oracle(lp)
    mx = engine.solve(lp, default_params(mode=MODE_PDLP, precision=PREC_MIXED, max_iters=2_000_000, tol=1e-6))
    km = o.kkt(mx.x, mx.y)
    print(f"mixed: status {mx.status} it {mx.iterations} (fp64 {mx.info.polish_iterations}) restarts "
          f"{mx.info.restarts} engine kkt {mx.info.rel_primal:.3e} {mx.info.rel_dual:.3e} {mx.info.rel_gap:.3e} "
          f"or_kkt {km[0]:.3e} {km[1]:.3e} {km[2]:.3e}")
    assert mx.status == 0
    assert max(km[0], km[1], km[2]) <= 1e-6
    # The engine's own evaluation agrees (the gap's numerator P - D cancels, so
    # its rounding shows at ~1e-15 of the objectives: compared absolutely).
    for a, b in ((km[0], mx.info.rel_primal), (km[1], mx.info.rel_dual), (km[2], mx.info.rel_gap)):
        assert abs(a - b) <= 1e-12
    fp = engine.solve(lp, default_params(mode=MODE_PDLP, max_iters=2_000_000, tol=1e-6))
    kf = o.kkt(fp.x, fp.y)
    assert fp.status == 0 and max(kf[0], kf[1], kf[2]) <= 1e-6
    ex = np.linalg.norm(mx.x - fp.x) / np.linalg.norm(fp.x)
    ey = np.linalg.norm(mx.y - fp.y) / np.linalg.norm(fp.y)
    eo = abs(km[3] - kf[3]) / (1 + abs(kf[3]))
    print(f"fp64: it {fp.iterations}; mixed vs fp64: |x-x64|/|x64| {ex:.2e} |y-y64|/|y64| {ey:.2e} "
          f"objective {eo:.2e}")
    assert ex <= 1e-2 and ey <= 1e-2
    assert eo <= 4e-6
