"""pdlp mode on LPs with dense rows and columns (more than 4,096 nonzeros):
such rows are summed in the chunked order (chunks of 64 products, groups of
32 chunks, each level sequential from 0.0; the oracle's or_pdlp_row_sum),
which lets a workgroup sum a tile's chunks in parallel instead of one lane
adding every product, and lets a row be split over workgroups in pieces of
whole groups, the last piece to finish adding the group sums in order. The device trajectory must equal or_pdlp's exactly; reference
mode keeps the sequential order at every length, and the batch and family
engines (one lane per row) refuse such rows."""
import numpy as np
import pytest

from nirnaya_amd.engine import LP, MODE_PDLP, MODE_REFERENCE, PREC_MIXED, EngineError, default_params
from oracle_lib import Oracle
from lpkit import dense_line_lp, long_line_infeasible_lps

pytestmark = pytest.mark.gpu


CASES = {
    "row_5000_col_6000": lambda: dense_line_lp(5, 8000, 9000, 5000, 6000),
    "row_10000": lambda: dense_line_lp(6, 3000, 12000, 10000, 100),
    "col_4097": lambda: dense_line_lp(7, 6000, 5000, 100, 4097),
    # 265 groups of 2,048: more group sums than one pass of the combining
    # workgroup fetches.
    "row_540000": lambda: dense_line_lp(8, 2000, 560000, 540000, 100, per_col=1),
    # Longer than the 1,024 and 2,048 tiles but summed sequentially (at most
    # 4,096): one-workgroup blocks in the long-row launch, not split.
    "row_3000_col_2500": lambda: dense_line_lp(9, 3000, 6000, 3000, 2500),
}
# Iteration caps: the first and last LPs are slow to converge, so their
# trajectories are compared over their first iterations (status: iteration
# limit).
CAPS = {"row_5000_col_6000": 6000, "row_10000": 200000, "col_4097": 200000, "row_540000": 400,
        "row_3000_col_2500": 200000}


_ORACLE = {}


def _oracle(name):
    if name not in _ORACLE:
        lp = CASES[name]()
        _ORACLE[name] = (lp, Oracle(lp).pdlp(max_iters=CAPS[name], tol=1e-6))
    return _ORACLE[name]


@pytest.mark.parametrize("tile", [0, 1024, 4096])
@pytest.mark.parametrize("name", sorted(CASES))
def test_pdlp_long_rows_match_oracle(engine, name, tile):
    lp, (xo, yo, info) = _oracle(name)
    r = engine.solve(lp, default_params(mode=MODE_PDLP, max_iters=CAPS[name], tol=1e-6, tile_nnz=tile))
    print(f"{name} tile {tile}: it {r.iterations}/{info.iterations} restarts {r.info.restarts}/{info.restarts}")
    assert r.status == info.status
    assert r.iterations == info.iterations and r.info.restarts == info.restarts
    assert np.array_equal(r.x, xo) and np.array_equal(r.y, yo)


@pytest.mark.parametrize("tile,pieces", [(1024, (3, 3)), (2048, (3, 3)), (4096, (2, 2))])
def test_pdlp_splits_long_rows_into_pieces(engine, tile, pieces):
    # One row of 5,000 and one column of 6,000 nonzeros: pieces of
    # max(tile, 2,048) nonzeros each; reference mode splits nothing.
    lp = CASES["row_5000_col_6000"]()
    s = engine.session(lp, default_params(mode=MODE_PDLP, max_iters=100, tol=1e-6, tile_nnz=tile))
    try:
        assert s.split() == (1, pieces[0], 1, pieces[1])
    finally:
        s.close()
    s = engine.session(lp, default_params(mode=MODE_REFERENCE, max_iters=100, tile_nnz=tile))
    try:
        assert s.split() == (0, 0, 0, 0)
    finally:
        s.close()


@pytest.mark.parametrize("name", ["row_3000_col_2500", "row_10000"])
def test_pdlp_long_rows_without_graphs(engine, name):
    # The long-row launches run on a second stream forked from the engine's
    # and joined back; without graph capture the same bits come out.
    lp, (xo, yo, info) = _oracle(name)
    r = engine.solve(lp, default_params(mode=MODE_PDLP, max_iters=CAPS[name], tol=1e-6, tile_nnz=1024, use_graph=0))
    assert r.iterations == info.iterations and r.info.restarts == info.restarts
    assert np.array_equal(r.x, xo) and np.array_equal(r.y, yo)


def test_pdlp_long_rows_ignore_segments(engine):
    # A segment width would cut the dense row: the session keeps one segment,
    # and the result is the unsegmented one.
    lp = CASES["row_5000_col_6000"]()
    s = engine.session(lp, default_params(mode=MODE_PDLP, max_iters=20000, tol=1e-6, seg_width=2048))
    try:
        assert s.shape()[2] == 1 and s.shape()[6] == 1
    finally:
        s.close()
    a = engine.solve(lp, default_params(mode=MODE_PDLP, max_iters=20000, tol=1e-6, seg_width=2048))
    b = engine.solve(lp, default_params(mode=MODE_PDLP, max_iters=20000, tol=1e-6))
    assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)


def test_mixed_precision_long_rows_meet_tolerance(engine):
    lp = CASES["row_10000"]()
    r = engine.solve(lp, default_params(mode=MODE_PDLP, precision=PREC_MIXED, max_iters=200000, tol=1e-6))
    assert r.status == 0
    k = Oracle(lp).kkt(r.x, r.y)
    assert max(k[0], k[1], k[2]) <= 1.5e-6


def test_reference_mode_keeps_the_sequential_order(engine):
    lp = CASES["row_10000"]()
    r = engine.solve(lp, default_params(mode=MODE_REFERENCE, max_iters=300, check_termination=0))
    xo, yo, _, _, _ = Oracle(lp).pdhg(300, check=False)
    assert np.array_equal(r.x, xo) and np.array_equal(r.y, yo)


def test_batch_and_family_refuse_long_rows(engine):
    lp = CASES["col_4097"]()
    p = default_params(mode=MODE_PDLP, max_iters=1000, tol=1e-6)
    with pytest.raises(EngineError, match="4096"):
        engine.batch([lp], p)
    with pytest.raises(EngineError, match="4096"):
        engine.family(lp, p, lanes=64)


@pytest.mark.parametrize("name", sorted(long_line_infeasible_lps()))
def test_pdlp_long_lines_certificates_match_oracle(engine, name):
    # Rays through a dense row and a dense column: the certificate's dd
    # products of those lines are split over workgroups in the chunked order,
    # and the status, iteration and iterates equal or_pdlp's.
    # Engine statuses: 1 infeasible (the oracle's 21); 2 or 3 unbounded (22),
    # 2 when the primal residual also meets the tolerance.
    kind, lp = long_line_infeasible_lps()[name]
    xo, yo, info = Oracle(lp).pdlp(max_iters=100000, tol=1e-6)
    r = engine.solve(lp, default_params(mode=MODE_PDLP, max_iters=100000, tol=1e-6))
    assert info.status == {"dual": 21, "primal": 22}[kind]
    want = 1 if info.status == 21 else (2 if info.rel_p <= 1e-6 else 3)
    assert r.status == want and r.iterations == info.iterations and r.info.restarts == info.restarts
    assert np.array_equal(r.x, xo) and np.array_equal(r.y, yo)
    # Mixed precision (FP32 iterations, FP64 certificate products) reaches
    # the same verdict.
    mx = engine.solve(lp, default_params(mode=MODE_PDLP, precision=PREC_MIXED, max_iters=100000, tol=1e-6))
    assert mx.status == 1 if kind == "dual" else mx.status in (2, 3)


def test_drop_in_boundary_solves_a_dense_row_lp():
    # The whole sov_* pipeline in pdlp mode (validate, presolve, engine with
    # the split row, unscale, dd certificate) on a 1,000 x 6,000 LP with a row
    # of 5,000 nonzeros: Optimal with the certificate passing, at HiGHS's
    # objective.
    import scipy.sparse as sp
    from scipy.optimize import linprog
    from nirnaya_amd.nirnaya import NirnayaModel
    lp = dense_line_lp(6, 1000, 6000, 5000, 100)
    A = sp.csc_matrix((lp.value, lp.index, lp.start), shape=(lp.m, lp.n)).tocsr()
    m = NirnayaModel("dense_row")
    for k, v in (("log_level", "0"), ("method", "pdhg"), ("pdhg_mode", "pdlp")):
        assert m.set_option(k, v) == 0
    for j in range(lp.n):
        m.add_var(float(lp.c[j]), float(lp.l[j]), float(lp.u[j]))
    for i in range(lp.m):
        a, b = A.indptr[i], A.indptr[i + 1]
        m.add_constr([int(j) for j in A.indices[a:b]], [float(v) for v in A.data[a:b]], float(lp.rl[i]),
                     float(lp.ru[i]))
    assert m.optimize() == 0
    assert m.Status == 0
    ref = linprog(lp.c, A_eq=A, b_eq=lp.rl, bounds=list(zip(lp.l, lp.u)), method="highs")
    assert ref.status == 0
    assert abs(m.ObjVal - ref.fun) <= 1e-5 * (1 + abs(ref.fun))


@pytest.mark.parametrize("seed", range(6))
def test_random_long_line_lps_match_oracle(engine, seed):
    # Random shapes, 1-3 long rows and 0-2 long columns of random lengths
    # (random piece counts and ragged last groups), mixed bound types: the
    # first 3,008 pdlp iterations in FP64 (the cap rounded up to the
    # 64-iteration evaluation) equal or_pdlp's bit for bit, and the mixed
    # precision engine runs the same LPs within the cap.
    from lpkit import random_long_line_lp
    lp = random_long_line_lp(seed)
    xo, yo, info = Oracle(lp).pdlp(max_iters=3000, tol=1e-6)
    r = engine.solve(lp, default_params(mode=MODE_PDLP, max_iters=3000, tol=1e-6))
    print(f"seed {seed}: {lp.m} x {lp.n}, {lp.nnz} nonzeros, status {r.status}, it {r.iterations}/{info.iterations}")
    assert r.iterations == info.iterations and r.info.restarts == info.restarts
    assert np.array_equal(r.x, xo) and np.array_equal(r.y, yo)
    mx = engine.solve(lp, default_params(mode=MODE_PDLP, precision=PREC_MIXED, max_iters=3000, tol=1e-6))
    assert mx.iterations <= 3000 + 64
