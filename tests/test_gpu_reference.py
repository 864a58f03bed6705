"""GPU parity of the engine's reference mode against the oracle.

The engine's CSR-stream kernels sum every row in ascending column order from
0.0 with products rounded before the add (no FMA contraction), exactly as the
reference's sequential loops do, so iterates are expected to agree bit for bit;
rows longer than one LP tile are summed as a chain of tiles in the same order,
so every row length is held to bit equality.
"""
import ctypes as C

import numpy as np
import pytest

from nirnaya_amd.engine import (LP, MODE_REFERENCE, NirParams, _STOP_CB, default_params, gen_equality_lp,
                                gen_mixed_lp, gen_staircase_lp)
from oracle_lib import Oracle
import lpkit

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(1.0, float(np.max(np.abs(b))))) if a.size else 0.0


def run_ref(engine, lp, K, check=False, **kw):
    p = default_params(mode=MODE_REFERENCE, max_iters=K, check_termination=1 if check else 0, **kw)
    return engine.solve(lp, p)


@pytest.mark.parametrize("K", [1, 10, 100, 1000])
def test_iterates_match_oracle_cfg0(engine, K):
    lp = gen_equality_lp(20260, 2000, 5000, 5)
    r = run_ref(engine, lp, K)
    xo, yo, tau, it, V = Oracle(lp).pdhg(K, check=False)
    print(f"K={K} rel x {rel(r.x, xo):.3e} y {rel(r.y, yo):.3e} exact={np.array_equal(r.x, xo) and np.array_equal(r.y, yo)}")
    assert r.iterations == it == K
    assert r.tau == tau
    assert rel(r.x, xo) <= 1e-10 and rel(r.y, yo) <= 1e-10
    assert r.info.max_primal_viol == pytest.approx(V, rel=1e-10)


@pytest.mark.parametrize("tile,rows", [(1024, 256), (2048, 64), (4096, 256)])
@pytest.mark.parametrize("graph", [0, 1])
def test_tiles_and_graphs_bit_identical(engine, tile, rows, graph):
    lp = gen_mixed_lp(7, 3000, 9000, 6.0, 0.01, 1e3)
    r = run_ref(engine, lp, 77, tile_nnz=tile, tile_rows=rows, use_graph=graph)
    xo, yo, _, _, _ = Oracle(lp).pdhg(77, check=False)
    assert np.array_equal(r.x, xo) and np.array_equal(r.y, yo)


@pytest.mark.parametrize("fused", [0, 1])
@pytest.mark.parametrize("seg", [1000, 4096, 1 << 17])
def test_column_segments_bit_identical(engine, seg, fused):
    # Segment passes carry the running row sums (in memory between launches,
    # or in registers when all segments are walked in one launch), so the
    # order of every sum is unchanged: still bit-identical to the oracle.
    lp = gen_mixed_lp(17, 6000, 18000, 6.0, 0.01, 1e3)
    r = run_ref(engine, lp, 50, seg_width=seg, tile_rows=128, seg_fused=fused)
    xo, yo, _, it, V = Oracle(lp).pdhg(50, check=False)
    assert np.array_equal(r.x, xo) and np.array_equal(r.y, yo)
    assert r.info.max_primal_viol == pytest.approx(V, rel=1e-12)


def test_fused_segments_full_size_cfg1(engine):
    """m=400,000, n=1,200,000: 5 segments walked in one launch, 12 iterations."""
    lp = gen_mixed_lp(20261, 400_000, 1_200_000, 6.0, 0.01, 1e3)
    p = default_params(mode=MODE_REFERENCE, max_iters=12, check_termination=0, seg_width=1 << 18, seg_width_t=0,
                       seg_fused=1)
    s = engine.session(lp, p)
    try:
        s.solve()
        r = s.result()
        shape = s.shape()
    finally:
        s.close()
    assert shape[2] == 5 and shape[3] == 1
    xo, yo, _, _, _ = Oracle(lp).pdhg(12, check=False)
    assert np.array_equal(r.x, xo) and np.array_equal(r.y, yo)


def test_autotuned_shapes_bit_identical(engine):
    lp = gen_equality_lp(20260, 2000, 5000, 5)
    p = default_params(mode=MODE_REFERENCE, max_iters=100, check_termination=0, tile_nnz=-1)
    s = engine.session(lp, p)
    try:
        s.run(100)
        r = s.result()
        shape = s.shape()
    finally:
        s.close()
    print("autotuned (A tile, rows, segs, AT tile, rows, segs):", shape)
    xo, yo, _, _, _ = Oracle(lp).pdhg(100, check=False)
    assert np.array_equal(r.x, xo) and np.array_equal(r.y, yo)


def test_stop_rule_matches_oracle(engine):
    # Reference stop rule (pdhg_solver.hpp:97-100) decided on the device.
    for lp in (lpkit.pdhg_unit_lp(), lpkit.transport_lp()):
        r = run_ref(engine, lp, 50000, check=True)
        xo, yo, _, it, _ = Oracle(lp).pdhg(50000, check=True)
        assert r.iterations == it and it < 50000
        assert rel(r.x, xo) <= 1e-10 and rel(r.y, yo) <= 1e-10
    # unit_tests.cpp:316-331 bar on the engine itself
    r = run_ref(engine, lpkit.pdhg_unit_lp(), 50000, check=True)
    assert r.status == 0 and abs(r.objective - 9.0) <= 1e-2


def test_edge_shapes(engine):
    # Free rows, one-sided and ranged rows, free and one-sided columns, an
    # empty row, an empty column and a dense row longer than the LP tile.
    rng = np.random.default_rng(5)
    m, n = 60, 5000
    A = np.zeros((m, n))
    for j in range(n):
        rows = rng.choice(m - 2, size=3, replace=False) + 1
        A[rows, j] = rng.standard_normal(3)
    A[:, 7] = 0.0                       # empty column
    A[m - 1, :] = rng.standard_normal(n)  # 5000-long row: long-row path
    A[0, :] = 0.0                       # empty row
    inf = np.inf
    l = np.where(rng.random(n) < 0.2, -inf, 0.0)
    u = np.where(rng.random(n) < 0.3, inf, rng.uniform(1, 5, n))
    rl = rng.uniform(-1, 0, m)
    ru = rl + rng.uniform(0, 2, m)
    rl[rng.random(m) < 0.3] = -inf
    ru[rng.random(m) < 0.3] = inf
    rl[5], ru[5] = -inf, inf            # free row
    lp = LP.from_dense(A, rng.standard_normal(n), l, u, rl, ru)
    r = run_ref(engine, lp, 300)
    xo, yo, _, _, _ = Oracle(lp).pdhg(300, check=False)
    assert np.array_equal(r.x, xo) and np.array_equal(r.y, yo)
    # Without the long row as well.
    A[m - 1, :] = 0.0
    lp2 = LP.from_dense(A, lp.c, l, u, rl, ru)
    r2 = run_ref(engine, lp2, 300)
    xo2, yo2, _, _, _ = Oracle(lp2).pdhg(300, check=False)
    assert np.array_equal(r2.x, xo2) and np.array_equal(r2.y, yo2)


def _long_row_lp():
    """12,000 x 10,500 LP with one 10,500-nonzero row (a long row of A) and
    one 12,000-nonzero column (a long row of A^T); every other column has 3
    random rows."""
    import scipy.sparse as sp
    rng = np.random.default_rng(11)
    m, n = 12_000, 10_500
    rows = [rng.choice(m, size=3, replace=False) for _ in range(n)]
    ri = np.concatenate(rows)
    ci = np.repeat(np.arange(n), 3)
    dense_row, dense_col = 3, 5
    ri = np.concatenate([ri, np.full(n, dense_row), np.arange(m)])
    ci = np.concatenate([ci, np.arange(n), np.full(m, dense_col)])
    A = sp.csc_matrix((rng.standard_normal(ri.size), (ri, ci)), shape=(m, n))
    A.sum_duplicates()
    A.sort_indices()
    x0 = rng.uniform(0, 1, n)
    b = A @ x0
    c = A.T @ rng.standard_normal(m) + rng.uniform(0, 1, n)
    assert np.diff(A.indptr).max() == m and np.diff(A.tocsr().indptr).max() >= n
    return LP(m, n, A.indptr.astype(np.int64), A.indices.astype(np.int32), A.data.astype(np.float64), c,
              np.zeros(n), np.full(n, 10.0), b.copy(), b.copy())


@pytest.mark.parametrize("tile,seg,fused", [(1024, 0, 0), (2048, 0, 0), (4096, 0, 0), (2048, 3000, 0),
                                            (1024, 4096, 1)])
def test_rows_longer_than_a_tile_bit_identical(engine, tile, seg, fused):
    """Rows of 10,500 and 12,000 nonzeros (up to 12 tiles each): the chain
    of tiles keeps the sequential sum, so iterates equal the oracle bit for
    bit in every tile size and segment layout (a fused layout falls back to
    per-segment launches where a row does not fit one tile per segment)."""
    lp = _long_row_lp()
    r = run_ref(engine, lp, 40, tile_nnz=tile, seg_width=seg, seg_fused=fused)
    xo, yo, tau, _, V = Oracle(lp).pdhg(40, check=False)
    assert r.tau == tau
    assert np.array_equal(r.x, xo) and np.array_equal(r.y, yo)
    assert r.info.max_primal_viol == V


def test_warm_start_matches_oracle(engine):
    lp = gen_staircase_lp(9, 6, 8)
    x0 = lp.x0 * 0.9
    y0 = 0.1 * lp.y0
    p = default_params(mode=MODE_REFERENCE, max_iters=200, check_termination=0)
    r = engine.solve(lp, p, x0=x0, y0=y0)
    xo, yo, _, _, _ = Oracle(lp).pdhg(200, check=False, x0=x0, y0=y0)
    assert np.array_equal(r.x, xo) and np.array_equal(r.y, yo)


def test_should_stop_gives_time_limit_on_256_boundary(engine):
    lp = gen_equality_lp(3, 500, 1200, 5)
    calls = []

    def stop(_user):
        calls.append(1)
        return 1 if len(calls) >= 3 else 0

    cb = _STOP_CB(stop)
    p = default_params(mode=MODE_REFERENCE, max_iters=100000, check_termination=0)
    p.should_stop = cb
    r = engine.solve(lp, p)
    assert r.status == 4 and r.iterations == 512
    # and the iterate is exactly the oracle's after 512 iterations
    xo, yo, _, _, _ = Oracle(lp).pdhg(512, check=False)
    assert np.array_equal(r.x, xo)


def test_session_runs_in_pieces(engine):
    lp = gen_equality_lp(8, 800, 2000, 5)
    p = default_params(mode=MODE_REFERENCE, max_iters=1000, check_termination=0)
    s = engine.session(lp, p)
    try:
        for k in (1, 31, 32, 300, 17):
            s.run(k)
        r = s.result()
    finally:
        s.close()
    xo, yo, _, _, _ = Oracle(lp).pdhg(381, check=False)
    assert r.iterations == 381 and np.array_equal(r.x, xo) and np.array_equal(r.y, yo)


@pytest.mark.parametrize("seg", [0, 700])
def test_session_kkt_matches_oracle(engine, seg):
    """Relative KKT of the plain PDHG's iterate, evaluated on the device,
    against the oracle's evaluation of its own (bit-identical) iterate. The
    per-element terms are the same double-double arithmetic; only the order
    of the block sums differs, so the terms agree to rounding of the sums."""
    lp = gen_mixed_lp(23, 1500, 4500, 6.0, 0.01, 1e3)
    lp.ru[::7] = np.inf                       # some rows free above, so y may carry sign violations
    lp.l[::11] = -np.inf                      # free-below columns give dual residuals
    p = default_params(mode=MODE_REFERENCE, max_iters=10_000, check_termination=0, seg_width=seg)
    s = engine.session(lp, p)
    o = Oracle(lp)
    try:
        xo = yo = None
        done = 0
        for k in (0, 250, 1750):
            s.run(k)
            done += k
            got = s.kkt()
            r = s.result()
            xo, yo, _, _, _ = o.pdhg(done, check=False) if done else (r.x, r.y, 0, 0, 0)
            assert np.array_equal(r.x, xo) and np.array_equal(r.y, yo)
            want = o.kkt(xo, yo)
            print(f"K={done} device {got[:3]} oracle {want[:3]}")
            np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-15)
            assert got[0] > 0 and got[2] > 0
    finally:
        s.close()


def test_session_kkt_rejects_pdlp_sessions(engine):
    from nirnaya_amd.engine import EngineError, MODE_PDLP
    lp = gen_equality_lp(8, 200, 500, 5)
    s = engine.session(lp, default_params(mode=MODE_PDLP, max_iters=100))
    try:
        with pytest.raises(EngineError):
            s.kkt()
    finally:
        s.close()


def test_gather_floor_measures_both_products(engine):
    lp = gen_mixed_lp(5, 3000, 9000, 6.0, 0.01, 1e3)
    s = engine.session(lp, default_params(mode=MODE_REFERENCE, max_iters=100, check_termination=0,
                                          seg_width=2000))
    try:
        s.run(10)
        before = s.result()
        fa, ft = s.gather_floor(0, 5), s.gather_floor(1, 5)
        after = s.result()
        assert 0 < fa < 1e5 and 0 < ft < 1e5
        # it only reads the session's arrays
        assert np.array_equal(before.x, after.x) and np.array_equal(before.y, after.y)
        from nirnaya_amd.engine import EngineError
        with pytest.raises(EngineError):
            s.gather_floor(2, 5)
    finally:
        s.close()


def test_full_size_cfg1_parity(engine):
    """m=400,000, n=1,200,000, ~9.6M nonzeros, in the layout the autotuner
    picks for it (the production path): 12 iterations equal the oracle's
    bit for bit."""
    lp = gen_mixed_lp(20261, 400_000, 1_200_000, 6.0, 0.01, 1e3)
    assert 9_000_000 < lp.nnz < 10_200_000
    s = engine.session(lp, default_params(mode=MODE_REFERENCE, max_iters=12, check_termination=0))
    try:
        s.solve()
        r = s.result()
        shape = s.shape()
    finally:
        s.close()
    print("autotuned (A tile, rows, segs, fused, AT tile, rows, segs, fused):", shape)
    xo, yo, tau, _, _ = Oracle(lp).pdhg(12, check=False)
    assert r.tau == tau
    assert np.array_equal(r.x, xo) and np.array_equal(r.y, yo)
    # The layouts the autotuner picks on MI355X, set explicitly so they are
    # covered whatever the timing picks: A in three equal segments of
    # 400,384 columns and A^T in two of 200,704, with one launch per segment
    # and with all segments walked in one launch.
    for fused in (0, 1):
        p = default_params(mode=MODE_REFERENCE, max_iters=12, check_termination=0, tile_nnz=2048,
                           seg_width=400_384, seg_width_t=200_704, seg_fused=fused)
        s = engine.session(lp, p)
        try:
            s.solve()
            r = s.result()
            shape = tuple(s.shape())
        finally:
            s.close()
        assert shape == (2048, 256, 3, fused, 2048, 256, 2, fused), shape
        assert r.tau == tau
        assert np.array_equal(r.x, xo) and np.array_equal(r.y, yo), fused
