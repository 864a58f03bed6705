"""ctypes access to the oracle (test infrastructure only).

  oracle/liboracle.so  - the C restatement of the reference PDHG path
  oracle/_ref/libref.so - the reference's own sources compiled by
                          oracle/Makefile (present only where /root/reference
                          was available at build time)

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, and
only as the checker or the CPU timing baseline.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref.so")

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)


class OrPdlpParams(C.Structure):
    _fields_ = [("max_iters", C.c_int64), ("tol", C.c_double), ("eval_freq", C.c_int32), ("unscale_row", _dp),
                ("unscale_col", _dp), ("x0", _dp), ("y0", _dp)]


class OrPdlpInfo(C.Structure):
    _fields_ = [("status", C.c_int32), ("restarts", C.c_int32), ("iterations", C.c_int64), ("omega", C.c_double),
                ("eta", C.c_double), ("rel_p", C.c_double), ("rel_d", C.c_double), ("rel_g", C.c_double),
                ("pobj", C.c_double), ("dobj", C.c_double)]


class OrLp(C.Structure):
    _fields_ = [("m", C.c_int32), ("n", C.c_int32), ("start", _i64p), ("index", _i32p), ("value", _dp),
                ("c", _dp), ("l", _dp), ("u", _dp), ("rl", _dp), ("ru", _dp), ("offset", C.c_double)]


def _p(a, t=C.c_double):
    if a is None:
        return C.cast(None, C.POINTER(t))
    return a.ctypes.data_as(C.POINTER(t))


_oracle = None
_ref = None


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise FileNotFoundError(ORACLE_SO + " not built")
        L = C.CDLL(ORACLE_SO)
        L.or_spmv.argtypes = [C.POINTER(OrLp), _dp, _dp]
        L.or_spmv_t.argtypes = [C.POINTER(OrLp), _dp, _dp]
        L.or_scale_geometric.argtypes = [C.POINTER(OrLp), _dp, _dp]
        L.or_pdhg_reference.argtypes = [C.POINTER(OrLp), C.c_int64, C.c_int, _dp, _dp, _dp, _dp, _dp, _i64p, _dp]
        L.or_pdhg_steps.argtypes = [C.POINTER(OrLp), C.c_int64, C.c_double, C.c_double, _dp, _dp]
        L.or_dd_verify.argtypes = [C.POINTER(OrLp), _dp, C.c_double, C.c_double, _dp, _dp, _dp]
        L.or_kkt.argtypes = [C.POINTER(OrLp), _dp, _dp, _dp]
        L.or_scale_pdlp.argtypes = [C.POINTER(OrLp), C.c_int32, C.c_int32, _dp, _dp]
        L.or_pdlp.argtypes = [C.POINTER(OrLp), C.POINTER(OrPdlpParams), _dp, _dp, C.POINTER(OrPdlpInfo)]
        L.or_pdlp_row_sum.argtypes = [_dp, _i32p, C.c_int64, _dp]
        L.or_pdlp_row_sum.restype = C.c_double
        _oracle = L
    return _oracle


def have_ref() -> bool:
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        if not have_ref():
            raise FileNotFoundError(REF_SO + " not built")
        L = C.CDLL(REF_SO)
        csc = [C.c_int32, C.c_int32, _i64p, _i32p, _dp]
        lpargs = csc + [_dp] * 5
        L.ref_pdhg.argtypes = lpargs + [C.c_int32, C.c_int32, _dp, _dp, _i64p, _dp]
        L.ref_pdhg_timed.argtypes = lpargs + [C.c_double, C.c_int32, _i64p, _dp]
        L.ref_scale.argtypes = lpargs + [_dp, _dp]
        L.ref_spmv.argtypes = csc + [_dp, _dp]
        L.ref_spmv_t.argtypes = csc + [_dp, _dp]
        L.ref_verify.argtypes = lpargs + [_dp, C.c_double, C.c_double, _dp]
        L.ref_simplex.argtypes = lpargs + [C.POINTER(C.c_int8), _i32p, _i64p, _dp, _dp]
        L.ref_presolve.argtypes = lpargs + [C.POINTER(C.c_int8), C.c_int32, _i32p, _dp, _dp, _i32p, _i32p, _dp, _dp,
                                            _dp, _dp]
        _ref = L
    return _ref


class Oracle:
    """The C oracle bound to one LP (nirnaya_amd.engine.LP)."""

    def __init__(self, lp):
        self.lp = lp
        self.s = OrLp(lp.m, lp.n, _p(lp.start, C.c_int64), _p(lp.index, C.c_int32), _p(lp.value), _p(lp.c),
                      _p(lp.l), _p(lp.u), _p(lp.rl), _p(lp.ru), float(lp.offset))
        self.L = oracle()

    def spmv(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.lp.m)
        self.L.or_spmv(C.byref(self.s), _p(x), _p(y))
        return y

    def spmv_t(self, y):
        y = np.ascontiguousarray(y, np.float64)
        g = np.zeros(self.lp.n)
        self.L.or_spmv_t(C.byref(self.s), _p(y), _p(g))
        return g

    def scale(self):
        R = np.zeros(self.lp.m)
        Cc = np.zeros(self.lp.n)
        active = self.L.or_scale_geometric(C.byref(self.s), _p(R), _p(Cc))
        return bool(active), R, Cc

    def scale_pdlp(self, ruiz=10, pc=1):
        """Row and column factors of the pdlp mode's scaling (10 Ruiz passes and
        one Pock-Chambolle pass unless asked otherwise)."""
        R = np.zeros(self.lp.m)
        Cc = np.zeros(self.lp.n)
        self.L.or_scale_pdlp(C.byref(self.s), int(ruiz), int(pc), _p(R), _p(Cc))
        return R, Cc

    def pdhg(self, K, check=True, x0=None, y0=None):
        x = np.zeros(self.lp.n)
        y = np.zeros(self.lp.m)
        tau = C.c_double()
        it = C.c_int64()
        V = C.c_double()
        x0 = None if x0 is None else np.ascontiguousarray(x0, np.float64)
        y0 = None if y0 is None else np.ascontiguousarray(y0, np.float64)
        self.L.or_pdhg_reference(C.byref(self.s), int(K), 1 if check else 0, _p(x0), _p(y0), _p(x), _p(y),
                                 C.byref(tau), C.byref(it), C.byref(V))
        return x, y, tau.value, it.value, V.value

    def steps(self, K, tau, x, y):
        """K reference iterations from (x, y) in place (step sizes tau = sigma)."""
        self.L.or_pdhg_steps(C.byref(self.s), int(K), float(tau), float(tau), _p(x), _p(y))

    def pdlp(self, max_iters=100000, tol=1e-6, eval_freq=64, unscale_row=None, unscale_col=None, x0=None, y0=None):
        keep = [None if a is None else np.ascontiguousarray(a, np.float64) for a in (unscale_row, unscale_col, x0, y0)]
        p = OrPdlpParams(int(max_iters), float(tol), int(eval_freq), *[_p(a) for a in keep])
        x = np.zeros(self.lp.n)
        y = np.zeros(self.lp.m)
        info = OrPdlpInfo()
        self.L.or_pdlp(C.byref(self.s), C.byref(p), _p(x), _p(y), C.byref(info))
        return x, y, info

    def verify(self, x, claimed, tol=1e-6):
        x = np.ascontiguousarray(x, np.float64)
        rv, bv, obj = C.c_double(), C.c_double(), C.c_double()
        ok = self.L.or_dd_verify(C.byref(self.s), _p(x), float(claimed), float(tol), C.byref(rv), C.byref(bv),
                                 C.byref(obj))
        return bool(ok), rv.value, bv.value, obj.value

    def kkt(self, x, y):
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        out = np.zeros(5)
        self.L.or_kkt(C.byref(self.s), _p(x), _p(y), _p(out))
        return out


class Reference:
    """The reference's own code (oracle/_ref/libref.so) bound to one LP."""

    def __init__(self, lp):
        self.lp = lp
        self.L = ref()
        self.csc = (lp.m, lp.n, _p(lp.start, C.c_int64), _p(lp.index, C.c_int32), _p(lp.value))
        self.vecs = (_p(lp.c), _p(lp.l), _p(lp.u), _p(lp.rl), _p(lp.ru))

    def pdhg(self, simplex_iter_limit=500000, threads=1):
        x = np.zeros(self.lp.n)
        y = np.zeros(self.lp.m)
        it = C.c_int64()
        obj = C.c_double()
        st = self.L.ref_pdhg(*self.csc, *self.vecs, int(simplex_iter_limit), int(threads), _p(x), _p(y), C.byref(it),
                             C.byref(obj))
        return x, y, st, it.value, obj.value

    def pdhg_timed(self, seconds, threads=1):
        it = C.c_int64()
        sec = C.c_double()
        st = self.L.ref_pdhg_timed(*self.csc, *self.vecs, float(seconds), int(threads), C.byref(it), C.byref(sec))
        return st, it.value, sec.value

    def scale(self):
        R = np.zeros(self.lp.m)
        Cc = np.zeros(self.lp.n)
        active = self.L.ref_scale(*self.csc, *self.vecs, _p(R), _p(Cc))
        return bool(active), R, Cc

    def spmv(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.lp.m)
        self.L.ref_spmv(*self.csc, _p(x), _p(y))
        return y

    def spmv_t(self, y):
        y = np.ascontiguousarray(y, np.float64)
        g = np.zeros(self.lp.n)
        self.L.ref_spmv_t(*self.csc, _p(y), _p(g))
        return g

    def verify(self, x, claimed, tol=1e-6):
        x = np.ascontiguousarray(x, np.float64)
        out = np.zeros(3)
        ok = self.L.ref_verify(*self.csc, *self.vecs, _p(x), float(claimed), float(tol), _p(out))
        return bool(ok), out[0], out[1], out[2]


def reference_presolve(lp, integer=None, level=2):
    """The reference's own Presolver::presolve (oracle/_ref) on one LP, in the
    form nirnaya_amd.engine.presolve returns."""
    L = ref()
    r = Reference(lp)
    integer = None if integer is None else np.ascontiguousarray(integer, dtype=np.int8)
    ints = np.zeros(7, dtype=np.int32)
    off = C.c_double()
    fixed = np.zeros(lp.n)
    cmap = np.zeros(max(lp.n, 1), dtype=np.int32)
    rmap = np.zeros(max(lp.m, 1), dtype=np.int32)
    cl, cu = np.zeros(max(lp.n, 1)), np.zeros(max(lp.n, 1))
    rl, ru = np.zeros(max(lp.m, 1)), np.zeros(max(lp.m, 1))
    ip = integer.ctypes.data_as(C.POINTER(C.c_int8)) if integer is not None else None
    L.ref_presolve(*r.csc, *r.vecs, ip, int(level), _p(ints, C.c_int32), C.byref(off), _p(fixed),
                   _p(cmap, C.c_int32), _p(rmap, C.c_int32), _p(cl), _p(cu), _p(rl), _p(ru))
    nr, nc = int(ints[2]), int(ints[3])
    return {"status": int(ints[0]), "solved": int(ints[1]), "nrow": nr, "ncol": nc, "rows_removed": int(ints[4]),
            "cols_removed": int(ints[5]), "tightenings": int(ints[6]), "obj_offset": off.value, "fixed_x": fixed,
            "col_map": cmap[:nc].copy(), "row_map": rmap[:nr].copy(), "collo": cl[:nc].copy(),
            "colup": cu[:nc].copy(), "rowlo": rl[:nr].copy(), "rowup": ru[:nr].copy()}


def reference_simplex(lp, status=None, basic=None):
    """The reference's simplex (oracle/_ref LpSolver) on one LP, cold or from a
    basis: (status, iterations, objective, x)."""
    L = ref()
    r = Reference(lp)
    it = C.c_int64()
    obj = C.c_double()
    x = np.zeros(lp.n)
    sp = None if status is None else np.ascontiguousarray(status, np.int8).ctypes.data_as(C.POINTER(C.c_int8))
    bp = None if basic is None else np.ascontiguousarray(basic, np.int32).ctypes.data_as(C.POINTER(C.c_int32))
    keep = (status, basic)
    st = L.ref_simplex(*r.csc, *r.vecs, sp, bp, C.byref(it), C.byref(obj), _p(x))
    del keep
    return st, it.value, obj.value, x
