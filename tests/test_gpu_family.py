"""GPU parity of the parametric-family engine (nir_family_*) against the
single-LP engine and the oracle.

Members share A and the column bounds and have their own costs and row
bounds; each runs on a lane of a 64-member group with interleaved vectors.
Every member must return exactly what nir_pdhg_solve returns for it alone:
x and y bit for bit, the same status, iteration and restart counts and
objective, whatever lane it ran on and whenever its lane was refilled. The
single-LP FP64 trajectory is itself or_pdlp's, which one member checks
directly."""
import numpy as np
import pytest

from nirnaya_amd.engine import LP, MODE_PDLP, _STOP_CB, default_params
from nirnaya_amd.warmstart import ParametricFamily
from oracle_lib import Oracle
import lpkit

pytestmark = pytest.mark.gpu


def same(a, b):
    return (np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y) and a.status == b.status
            and a.iterations == b.iterations and a.info.restarts == b.info.restarts and a.objective == b.objective
            and a.info.primal_weight == b.info.primal_weight)


def members(fam, thetas):
    def member(i):
        lp = fam.instance(thetas[i])
        return lp.c, lp.rl, lp.ru
    return member


@pytest.mark.parametrize("lanes", [64, 128])
def test_family_members_match_single_lp(engine, lanes):
    """100 members of a 400 x 1,000 family through 64 or 128 lanes (lanes
    refilled many times over)."""
    fam = ParametricFamily(20264, 400, 1000, 5)
    th = fam.thetas(100, 11)
    params = default_params(mode=MODE_PDLP, max_iters=400000, tol=1e-6)
    F = engine.family(fam.base, params, lanes=lanes)
    try:
        res = F.solve(len(th), members(fam, th))
        st = F.stats()
    finally:
        F.close()
    assert st.instances == len(th) and (st.refills >= 2 or lanes >= len(th))
    its = []
    for i, t in enumerate(th):
        s = engine.solve(fam.instance(t), params)
        its.append(s.iterations)
        assert same(res[i], s), f"member {i}: it {res[i].iterations}/{s.iterations}"
    print(f"lanes {lanes}: 100 members equal the single-LP engine; iterations {min(its)}..{max(its)}, "
          f"blocks {st.blocks}, refills {st.refills}, solve {st.solve_seconds:.2f}s")
    xo, yo, info = Oracle(fam.instance(th[0])).pdlp(max_iters=400000, tol=1e-6)
    assert res[0].iterations == info.iterations and np.array_equal(res[0].x, xo) and np.array_equal(res[0].y, yo)


def test_family_warm_starts_cap_and_start_point_optimum(engine):
    """Per-member warm starts (x only, y only, both), a member warm-started
    at its optimum (done at iteration 0), and an iteration cap (status 6)."""
    fam = ParametricFamily(20264, 300, 700, 5)
    th = fam.thetas(12, 5)
    params = default_params(mode=MODE_PDLP, max_iters=400000, tol=1e-6, primal_weight=0.5)
    cold = [engine.solve(fam.instance(t), params) for t in th]
    rng = np.random.default_rng(2)
    starts = []
    for i, r in enumerate(cold):
        x0 = r.x * (1 + 0.01 * rng.standard_normal(r.x.size)) if i % 3 != 1 else None
        y0 = r.y * 0.9 if i % 3 != 0 else None
        starts.append((x0, y0))
    F = engine.family(fam.base, params, lanes=64)
    try:
        res = F.solve(len(th), lambda i: (*members(fam, th)(i), *starts[i]))
    finally:
        F.close()
    for i, t in enumerate(th):
        s = engine.solve(fam.instance(t), params, x0=starts[i][0], y0=starts[i][1])
        assert same(res[i], s), f"member {i}"
    unit = lpkit.pdhg_unit_lp()
    p9 = default_params(mode=MODE_PDLP, max_iters=400000, tol=1e-9)
    F = engine.family(unit, p9, lanes=64)
    try:
        r = F.solve(2, lambda i: (unit.c, unit.rl, unit.ru, np.array([3.0, 1.0]) if i == 0 else None,
                                  np.array([1.5, 0.5]) if i == 0 else None))
    finally:
        F.close()
    assert r[0].status == 0 and r[0].iterations == 0
    assert same(r[1], engine.solve(unit, p9))
    capped = default_params(mode=MODE_PDLP, max_iters=640, tol=1e-12)
    F = engine.family(fam.base, capped, lanes=64)
    try:
        rc = F.solve(3, members(fam, th))
    finally:
        F.close()
    for i in range(3):
        s = engine.solve(fam.instance(th[i]), capped)
        assert rc[i].status == 6 and rc[i].iterations == 640 and same(rc[i], s)


def test_family_device_starts_equal_host_starts(engine):
    """Warm starts decoded on the device (nir_pod_decode) and read there by
    the family engine (nir_family_set_device_starts) give every member
    exactly the result of the same start handed over from the host, which is
    the single-LP engine's with that x0 / y0."""
    fam = ParametricFamily(20264, 300, 700, 5)
    th = fam.thetas(20, 9)
    params = default_params(mode=MODE_PDLP, max_iters=400000, tol=1e-6)
    cold = [engine.solve(fam.instance(t), params) for t in th[:6]]
    rng = np.random.default_rng(4)
    # Bases from the first cold solutions plus noise; starts decoded from
    # random coefficients around them.
    Sx = np.stack([r.x for r in cold]) + 0.01 * rng.standard_normal((6, fam.base.n))
    Sy = np.stack([r.y for r in cold]) + 0.01 * rng.standard_normal((6, fam.base.m))
    px = engine.pod_basis((Sx - Sx.mean(0)).T.copy(), np.eye(6)[:, :4], Sx.mean(0))
    py = engine.pod_basis((Sy - Sy.mean(0)).T.copy(), np.eye(6)[:, :4], Sy.mean(0))
    try:
        ax, ay = 0.3 * rng.standard_normal((len(th), 4)), 0.3 * rng.standard_normal((len(th), 4))
        X0, Y0 = px.decode(ax), py.decode(ay)
        F = engine.family(fam.base, params, lanes=64)
        try:
            F.set_device_starts(px, py)
            dev = F.solve(len(th), lambda i: (*members(fam, th)(i), None, None, 0.7))
            F.set_device_starts(None, None)
            host = F.solve(len(th), lambda i: (*members(fam, th)(i), X0[i], Y0[i], 0.7))
        finally:
            F.close()
    finally:
        px.close()
        py.close()
    for i in range(len(th)):
        assert same(dev[i], host[i]), f"member {i}"
    p7 = default_params(mode=MODE_PDLP, max_iters=400000, tol=1e-6, primal_weight=0.7)
    for i in (0, 13):
        assert same(dev[i], engine.solve(fam.instance(th[i]), p7, x0=X0[i], y0=Y0[i])), f"member {i}"


def test_family_handoff_is_bit_identical(engine):
    """The last members (at most `handoff` running, none waiting) finish on
    the single-LP engine from their lanes' state: the results are the same
    bits as with the hand-off switched off, and as the single-LP engine's."""
    fam = ParametricFamily(20264, 400, 1000, 5)
    th = fam.thetas(40, 17)
    params = default_params(mode=MODE_PDLP, max_iters=400000, tol=1e-6)
    out = {}
    for h in (0, 8, 64):
        F = engine.family(fam.base, params, lanes=64, handoff=h)
        try:
            out[h] = (F.solve(len(th), members(fam, th)), F.stats())
        finally:
            F.close()
    # handoff 64 >= 40 members: every member is handed over after its start-point evaluation
    assert out[0][1].handoffs == 0 and 1 <= out[8][1].handoffs <= 8 and out[64][1].handoffs == len(th)
    for i in range(len(th)):
        assert same(out[8][0][i], out[0][0][i]) and same(out[64][0][i], out[0][0][i]), f"member {i}"
    for i in (0, 39):
        assert same(out[64][0][i], engine.solve(fam.instance(th[i]), params)), f"member {i}"
    print(f"hand-off: {out[8][1].handoffs} members in {out[8][1].handoff_seconds:.2f}s; all {len(th)} in "
          f"{out[64][1].handoff_seconds:.2f}s; family alone {out[0][1].solve_seconds:.2f}s")


def test_family_error_return_then_destroy(engine):
    """A result callback that fails mid-solve (while the next block runs on
    the device) makes nir_family_solve return an error; destroying the family
    right after must wait for the work in flight, and the next family on the
    same context solves normally."""
    fam = ParametricFamily(20264, 300, 700, 5)
    th = fam.thetas(150, 21)
    params = default_params(mode=MODE_PDLP, max_iters=400000, tol=1e-6)
    seen = []

    def fail_on_third(i, x, y, info):
        seen.append(i)
        if len(seen) == 3:
            raise RuntimeError("callback failure forced by the test")

    F = engine.family(fam.base, params, lanes=64)
    with pytest.raises(RuntimeError, match="forced by the test"):
        F.solve(len(th), members(fam, th), on_result=fail_on_third)
    F.close()
    F = engine.family(fam.base, params, lanes=64)
    try:
        res = F.solve(4, members(fam, th))
    finally:
        F.close()
    for i in range(4):
        assert same(res[i], engine.solve(fam.instance(th[i]), params)), f"member {i}"


@pytest.mark.parametrize("polish", [0.0, 1e-7])
def test_family_mixed_precision_members_match_single_lp(engine, polish):
    """Mixed precision: every member equals the single-LP engine's mixed run
    bit for bit (re-centred FP32 phase; with a polish tolerance every member
    then runs its FP64 tail on the single-LP engine from its lane's state),
    through refills, warm starts and the hand-off; the oracle's or_kkt
    certifies the tolerance."""
    from nirnaya_amd.engine import PREC_MIXED
    fam = ParametricFamily(20264, 400, 1000, 5)
    th = fam.thetas(90, 23)
    params = default_params(mode=MODE_PDLP, precision=PREC_MIXED, max_iters=400000, tol=1e-6, polish_tol=polish)
    rng = np.random.default_rng(8)
    cold = engine.solve(fam.instance(th[0]), default_params(mode=MODE_PDLP, max_iters=400000, tol=1e-6))
    starts = {i: (cold.x * (1 + 0.01 * rng.standard_normal(cold.x.size)), None) for i in range(0, len(th), 7)}
    F = engine.family(fam.base, params, lanes=64)
    try:
        res = F.solve(len(th), lambda i: (*members(fam, th)(i), *starts.get(i, (None, None))))
        st = F.stats()
    finally:
        F.close()
    assert st.instances == len(th) and st.refills >= 2 and st.handoffs >= 1
    if polish > 0:
        assert st.tails >= 1
    for i, t in enumerate(th):
        x0, y0 = starts.get(i, (None, None))
        s = engine.solve(fam.instance(t), params, x0=x0, y0=y0)
        assert same(res[i], s), f"member {i}: it {res[i].iterations}/{s.iterations} st {res[i].status}/{s.status}"
        assert res[i].status == 0
    for i in (0, 1, 7):
        k = Oracle(fam.instance(th[i])).kkt(res[i].x, res[i].y)
        assert max(k[:3]) <= (polish or 1e-6), (i, k[:3])
    print(f"mixed, polish {polish}: {len(th)} members equal the single-LP engine; tails {st.tails}, "
          f"hand-offs {st.handoffs}, blocks {st.blocks}")


def test_family_certifies_infeasible_members(engine):
    """Members of an infeasible base (costs perturbed) end with the single-LP
    engine's certificate status and iterate."""
    for name, (kind, lp) in sorted(lpkit.infeasible_lps().items()):
        params = default_params(mode=MODE_PDLP, max_iters=50000, tol=1e-6)
        rng = np.random.default_rng(7)
        cs = [lp.c * (1 + 0.05 * rng.standard_normal(lp.n)) for _ in range(5)]
        F = engine.family(lp, params, lanes=64)
        try:
            res = F.solve(len(cs), lambda i: (cs[i], lp.rl, lp.ru))
        finally:
            F.close()
        for i, c in enumerate(cs):
            one = LP(lp.m, lp.n, lp.start, lp.index, lp.value, c, lp.l, lp.u, lp.rl, lp.ru)
            s = engine.solve(one, params)
            print(f"{name} member {i}: status {res[i].status}/{s.status} it {res[i].iterations}/{s.iterations}")
            assert same(res[i], s)
            assert res[i].status == 1 if kind == "dual" else res[i].status in (2, 3)


def test_family_should_stop(engine):
    fam = ParametricFamily(20264, 400, 1000, 5)
    th = fam.thetas(130, 3)
    calls = []

    def stop(_u):
        calls.append(1)
        return 1 if len(calls) >= 4 else 0

    params = default_params(mode=MODE_PDLP, max_iters=400000, tol=1e-9)
    cb = _STOP_CB(stop)
    params.should_stop = cb
    F = engine.family(fam.base, params, lanes=64)
    try:
        got = {}
        F.solve(len(th), members(fam, th), on_result=lambda i, x, y, info: got.__setitem__(i, info.status))
        st = F.stats()
    finally:
        F.close()
    assert st.stopped == 1 and len(got) == 64 and set(got.values()) == {4}
