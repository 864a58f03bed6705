"""The allocation-free certifier (nir_verify_solution_noalloc) against
nir_verify_solution on the same LP and point: every field of the report
equal bit for bit, on two calls in a row (the second on a smaller LP, so
the shared static scratch holds the first call's partials beyond the
second's range), plus the size limit checked on the host.

The caller's device buffers are made with the HIP runtime directly
(hipMalloc / hipMemcpy through ctypes): the entry point itself allocates
nothing."""
import ctypes as C

import numpy as np
import pytest

from nirnaya_amd.engine import MODE_REFERENCE, EngineError, NirDevLp, default_params, gen_equality_lp, gen_mixed_lp
import lpkit

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return _hip


class DeviceArrays:
    """Caller-owned device copies of host arrays."""

    def __init__(self):
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert hip().hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.ptrs.append(p)
        if a.nbytes:
            assert hip().hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0   # host to device
        return p.value

    def free(self):
        for p in self.ptrs:
            hip().hipFree(p)
        self.ptrs = []


def device_lp(lp, arrs):
    import scipy.sparse as sp
    A = sp.csc_matrix((lp.value, lp.index, lp.start), shape=(lp.m, lp.n)).tocsr()
    A.sort_indices()
    d = NirDevLp()
    d.nrow, d.ncol = lp.m, lp.n
    d.row_start = arrs.put(A.indptr.astype(np.int32))
    d.col_index = arrs.put(A.indices.astype(np.int32))
    d.value = arrs.put(A.data.astype(np.float64))
    d.cost, d.collo, d.colup = arrs.put(lp.c), arrs.put(lp.l), arrs.put(lp.u)
    d.rowlo, d.rowup = arrs.put(lp.rl), arrs.put(lp.ru)
    d.obj_offset = float(lp.offset)
    return d


FIELDS = ("max_bound_viol", "max_row_viol", "computed_obj", "obj_rel_err", "worst_col", "worst_col_upper",
          "worst_row", "worst_row_upper", "nonfinite_col", "passed")


def check_same(engine, lp, x, claimed, tol=1e-6):
    arrs = DeviceArrays()
    try:
        d = device_lp(lp, arrs)
        xd = arrs.put(np.asarray(x, np.float64))
        a = engine.verify_noalloc(d, xd, claimed, tol)
    finally:
        arrs.free()
    b = engine.verify(lp, x, claimed, tol)
    got = {f: getattr(a, f) for f in FIELDS}
    want = {f: getattr(b, f) for f in FIELDS}
    print(f"m={lp.m} n={lp.n}: {got}")
    assert got == want
    return a


def test_noalloc_certifier_matches_allocating_one(engine):
    # First call: a 40,000 x 120,000 LP at a non-optimal point (violations in
    # rows and bounds); second: the crude-blending fixture (8 rows), then a
    # 2,000 x 5,000 LP with a non-finite entry.
    big = gen_mixed_lp(20261, 40_000, 120_000, 6.0, 0.01, 1e3)
    r = engine.solve(big, default_params(mode=MODE_REFERENCE, max_iters=50, check_termination=0))
    rep = check_same(engine, big, r.x, r.objective)
    assert rep.max_row_viol > 0 and rep.worst_row >= 0
    lp, _, cols = lpkit.read_mps_lp(lpkit.CRUDE_MPS)
    claimed, vals = lpkit.read_sol(lpkit.CRUDE_SOL)
    rep = check_same(engine, lp, np.array([vals[c] for c in cols]), claimed)
    assert rep.passed and f"{rep.max_row_viol:.2e}" == "2.55e-14"
    mid = gen_equality_lp(20260, 2000, 5000, 5)
    x = mid.x0.copy()
    x[17] = np.nan
    rep = check_same(engine, mid, x, 0.0)
    assert rep.nonfinite_col == 17 and not rep.passed


def test_noalloc_certifier_rejects_sizes_beyond_its_scratch(engine):
    d = NirDevLp()
    d.nrow, d.ncol = (1 << 24) + 1, 10
    with pytest.raises(EngineError, match="16,777,216"):
        engine.verify_noalloc(d, 0, 0.0)
