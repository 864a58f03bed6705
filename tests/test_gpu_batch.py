"""GPU parity of the batched persistent-kernel engine (nir_batch_*) against the
single-LP engine (nir_pdhg_solve) and the oracle.

Each instance in a batch must return exactly what the single-LP engine
returns for it alone: the same x and y bit for bit, the same iteration,
restart and FP64-tail counts and the same status, in FP64 and in mixed
precision (with and without the FP64 polish), for every workgroup size,
work-list quantum and graph setting. The single-LP FP64 trajectory is itself
bit-identical to the oracle's or_pdlp, which one case checks directly."""
import numpy as np
import pytest

from nirnaya_amd.engine import (MODE_PDLP, PREC_MIXED, default_params, gen_equality_lp, gen_planning_lp,
                                gen_staircase_lp, planning_dims)
from oracle_lib import Oracle
import lpkit

pytestmark = pytest.mark.gpu


def small_batch():
    """Uneven sizes, different structures, one LP solved at its start point."""
    lps = [gen_staircase_lp(20263000 + k, T, K) for k, (T, K) in
           enumerate([(4, 5), (6, 8), (9, 12), (12, 7), (5, 30), (3, 6)])]
    lps += [lpkit.pdhg_unit_lp(), lpkit.transport_lp(), gen_equality_lp(31, 300, 700, 5)]
    return lps


def same(a, b):
    return (np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y) and a.status == b.status
            and a.iterations == b.iterations and a.info.restarts == b.info.restarts
            and a.info.polish_iterations == b.info.polish_iterations)


def check_batch(engine, lps, params, **opts):
    singles = [engine.solve(lp, params) for lp in lps]
    b = engine.batch(lps, params, **opts)
    try:
        b.solve()
        res = [b.result(i) for i in range(len(lps))]
        st = b.stats()
    finally:
        b.close()
    for i, (r, s) in enumerate(zip(res, singles)):
        print(f"instance {i}: m={lps[i].m} n={lps[i].n} status {r.status}/{s.status} it {r.iterations}/"
              f"{s.iterations} restarts {r.info.restarts}/{s.info.restarts} tail {r.info.polish_iterations}/"
              f"{s.info.polish_iterations} exact {np.array_equal(r.x, s.x) and np.array_equal(r.y, s.y)}")
        assert same(r, s), f"instance {i} differs from the single-LP engine"
    print(f"threads {st.threads} workgroups {st.workgroups} launches {st.launches} iterations {st.total_iterations}")
    return res, st


@pytest.mark.parametrize("threads", [0, 256, 1024])
def test_batch_fp64_matches_single(engine, threads):
    params = default_params(mode=MODE_PDLP, max_iters=200000, tol=1e-6)
    res, st = check_batch(engine, small_batch(), params, threads=threads)
    assert all(r.status == 0 for r in res)
    assert st.launches >= 1


def test_batch_mixed_and_polish_match_single(engine):
    lps = small_batch()
    check_batch(engine, lps, default_params(mode=MODE_PDLP, precision=PREC_MIXED, max_iters=200000, tol=1e-6))
    res, _ = check_batch(engine, lps, default_params(mode=MODE_PDLP, precision=PREC_MIXED, max_iters=200000,
                                                     tol=1e-6, polish_tol=1e-8))
    for lp, r in zip(lps, res):
        k = Oracle(lp).kkt(r.x, r.y)
        assert r.status == 0 and max(k[0], k[1], k[2]) <= 2e-8


def test_batch_small_quantum_eager_and_iteration_cap(engine):
    # Many launches (one block per ticket, one launch per replay), no graphs,
    # and a cap that stops some instances at status 6.
    lps = small_batch()
    params = default_params(mode=MODE_PDLP, max_iters=3000, tol=1e-7, use_graph=0)
    res, st = check_batch(engine, lps, params, quantum=1, rounds_per_graph=1)
    assert any(r.status == 6 for r in res) and st.launches > 10


def test_batch_instance_matches_oracle(engine):
    lp = gen_staircase_lp(20263002, 9, 12)
    b = engine.batch([lpkit.transport_lp(), lp], default_params(mode=MODE_PDLP, max_iters=200000, tol=1e-6))
    try:
        b.solve()
        r = b.result(1)
    finally:
        b.close()
    xo, yo, info = Oracle(lp).pdlp(max_iters=200000, tol=1e-6)
    assert r.iterations == info.iterations and np.array_equal(r.x, xo) and np.array_equal(r.y, yo)


def test_planning_family_members_match_single(engine):
    # Members of the 512-instance planning batch (seeds 20263*1000 + i).
    lps = [gen_planning_lp(20263 * 1000 + i) for i in (1, 4, 9)]
    print("dims", [planning_dims(20263 * 1000 + i) for i in (1, 4, 9)])
    check_batch(engine, lps, default_params(mode=MODE_PDLP, precision=PREC_MIXED, max_iters=400000, tol=1e-6))


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_lds_staging_regions_match_single(engine, prec):
    # 1024-lane workgroups with operands in LDS. The 48 x 40 staircase
    # (m + n = 30,768) holds y and xbar in one region copied between the
    # halves in FP64 and in two regions in FP32; the 20 x 30 one (m + n =
    # 9,620) and the small ones use two regions in both precisions. Rows of
    # length up to 40 (capacity rows) take several rounds of the sliced
    # products.
    lps = [gen_staircase_lp(20263100, 48, 40), gen_staircase_lp(20263101, 20, 30), gen_staircase_lp(20263102, 3, 6),
           lpkit.transport_lp()]
    params = default_params(mode=MODE_PDLP, precision=prec, max_iters=400000, tol=1e-6)
    res, st = check_batch(engine, lps, params, threads=1024)
    assert st.threads == 1024 and all(r.status == 0 for r in res)


def test_batch_wide_instance_32bit_indices_unstaged(engine):
    # 66,000 columns: the sliced indices take 32 bits and no instance's
    # operand fits in LDS (global gathers). An iteration cap keeps it short;
    # the capped iterates must still equal the single-LP engine's.
    lps = [gen_equality_lp(41, 300, 66000, 3), gen_staircase_lp(20263103, 5, 7)]
    for prec in (0, 1):
        params = default_params(mode=MODE_PDLP, precision=prec, max_iters=2048, tol=1e-9)
        res, st = check_batch(engine, lps, params)
        assert res[0].status == 6 and res[0].iterations == 2048


def test_batch_warm_starts_and_primal_weight_match_single(engine):
    lps = [gen_staircase_lp(20263000 + k, T, K) for k, (T, K) in enumerate([(4, 5), (6, 8), (9, 12)])]
    lps.append(gen_equality_lp(31, 300, 700, 5))
    rng = np.random.default_rng(4)
    x0s = [lp.x0 * (1 + 0.1 * rng.standard_normal(lp.n)) for lp in lps]
    y0s = [lp.y0 * 0.5 for lp in lps]
    x0s[1] = None   # cold x, warm y
    y0s[2] = None   # warm x, cold y
    for prec in (0, 1):
        params = default_params(mode=MODE_PDLP, precision=prec, max_iters=200000, tol=1e-6, primal_weight=0.7)
        b = engine.batch(lps, params, x0s=x0s, y0s=y0s)
        try:
            b.solve()
            res = [b.result(i) for i in range(len(lps))]
        finally:
            b.close()
        for i, lp in enumerate(lps):
            s = engine.solve(lp, params, x0=x0s[i], y0=y0s[i])
            print(f"prec {prec} instance {i}: it {res[i].iterations}/{s.iterations} omega0 0.7 "
                  f"exact {np.array_equal(res[i].x, s.x) and np.array_equal(res[i].y, s.y)}")
            assert same(res[i], s)


def test_batch_edge_shapes_match_single(engine):
    # Free and one-sided rows, ranged rows, an empty row, an empty column,
    # free and one-sided columns; an LP without nonzeros; a 1x1 LP.
    from nirnaya_amd.engine import LP
    rng = np.random.default_rng(12)
    m, n = 40, 90
    A = np.zeros((m, n))
    for j in range(n):
        rows = rng.choice(m - 1, size=3, replace=False) + 1
        A[rows, j] = rng.standard_normal(3)
    A[:, 5] = 0.0          # empty column
    A[0, :] = 0.0          # empty row
    inf = np.inf
    lo = np.where(rng.random(n) < 0.2, -inf, 0.0)
    up = np.where(rng.random(n) < 0.3, inf, rng.uniform(1, 5, n))
    rl = rng.uniform(-1, 0, m)
    ru = rl + rng.uniform(0, 2, m)
    rl[rng.random(m) < 0.3] = -inf
    ru[rng.random(m) < 0.3] = inf
    rl[7], ru[7] = -inf, inf
    c = rng.standard_normal(n)
    edge = LP.from_dense(A, c, lo, up, rl, ru)
    empty = LP.from_dense(np.zeros((3, 4)), np.array([1.0, -1.0, 0.0, 2.0]), np.zeros(4), np.ones(4),
                          np.full(3, -1.0), np.full(3, 1.0))
    one = LP.from_dense(np.array([[2.0]]), np.array([1.0]), np.array([0.0]), np.array([10.0]),
                        np.array([4.0]), np.array([4.0]))
    lps = [edge, empty, one, gen_staircase_lp(20263000, 4, 5)]
    for prec in (0, 1):
        params = default_params(mode=MODE_PDLP, precision=prec, max_iters=20000, tol=1e-6)
        check_batch(engine, lps, params)


def test_batch_should_stop_reports_time_limit(engine):
    from nirnaya_amd.engine import _STOP_CB
    calls = []

    def stop(_user):
        calls.append(1)
        return 1 if len(calls) >= 2 else 0

    cb = _STOP_CB(stop)
    lps = [gen_planning_lp(20263 * 1000 + i) for i in (2, 3)]
    params = default_params(mode=MODE_PDLP, max_iters=400000, tol=1e-9)
    params.should_stop = cb
    b = engine.batch(lps, params, quantum=1, rounds_per_graph=1)
    try:
        b.solve()
        res = [b.result(i) for i in range(len(lps))]
    finally:
        b.close()
    assert all(r.status == 4 for r in res)
    assert all(0 < r.iterations < 400000 for r in res)


def test_batch_certifies_infeasible_members(engine):
    # Infeasible and unbounded members end with the single-LP engine's
    # certificate status and iterate; the feasible ones beside them are solved.
    import lpkit
    bad = [lp for _, (_, lp) in sorted(lpkit.infeasible_lps().items())]
    lps = bad + [lpkit.transport_lp(), gen_staircase_lp(20263001, 5, 6)]
    for prec in (0, 1):
        params = default_params(mode=MODE_PDLP, precision=prec, max_iters=50000, tol=1e-6)
        res, _ = check_batch(engine, lps, params)
        kinds = [kind for _, (kind, _) in sorted(lpkit.infeasible_lps().items())]
        for r, kind in zip(res, kinds):
            assert r.status == 1 if kind == "dual" else r.status in (2, 3)
        assert [r.status for r in res[len(bad):]] == [0, 0]


@pytest.mark.parametrize("prec", [0, 1])
def test_cfg3_all_512_planning_lps(engine, prec):
    """The batched workload as specified: 512 multi-period planning LPs
    (seeds 20263*1000 + i, T ~ U{4..48} periods x K ~ U{5..40} products) in
    one persistent batch to relative KKT 1e-6, FP64 and mixed precision.
    Every instance must end Optimal with the oracle's independent or_kkt
    (double-double, on the host) at or below the tolerance; in FP64 the two
    smallest members must equal or_pdlp's iterates bit for bit."""
    lps = [gen_planning_lp(20263 * 1000 + i) for i in range(512)]
    params = default_params(mode=MODE_PDLP, precision=prec, max_iters=400000, tol=1e-6)
    b = engine.batch(lps, params)
    try:
        b.solve()
        res = [b.result(i) for i in range(len(lps))]
        st = b.stats()
    finally:
        b.close()
    worst = 0.0
    for i, (lp, r) in enumerate(zip(lps, res)):
        assert r.status == 0, f"instance {i} status {r.status}"
        k = Oracle(lp).kkt(r.x, r.y)
        worst = max(worst, k[0], k[1], k[2])
        assert max(k[0], k[1], k[2]) <= 1e-6, (i, k[:3])
    print(f"prec {prec}: 512 optimal, worst or_kkt {worst:.3e}, iterations {st.total_iterations}, "
          f"threads {st.threads}, solve {st.solve_seconds:.2f}s")
    if prec == 0:
        small = sorted(range(len(lps)), key=lambda i: lps[i].nnz)[:2]
        for i in small:
            xo, yo, info = Oracle(lps[i]).pdlp(max_iters=400000, tol=1e-6)
            print(f"instance {i} (m={lps[i].m}, n={lps[i].n}): it {res[i].iterations}/{info.iterations}")
            assert res[i].iterations == info.iterations and res[i].info.restarts == info.restarts
            assert np.array_equal(res[i].x, xo) and np.array_equal(res[i].y, yo)
