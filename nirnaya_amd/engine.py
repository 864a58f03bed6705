"""ctypes binding of the PDHG engine C ABI (include/nirnaya/pdhg.h) exported by
bin/libsov.so.

The engine replaces the reference's internal operator
``PdhgSolver::solve(const Model&, const Options&, const Deadline&)``
(src/gpu/pdhg_solver.hpp:25); this module lets tests and the benchmark drive it
directly with numpy arrays. Everything runs in the native library: there is no
Python or CPU fallback, and a missing library raises ``EngineUnavailable``.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
import weakref
from typing import Optional

import numpy as np

_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_ROOT, "bin", "libsov.so")

MODE_REFERENCE = 0
MODE_PDLP = 1
PREC_FP64 = 0
PREC_MIXED = 1


class EngineUnavailable(RuntimeError):
    """The native library is missing or no GPU can be opened."""


class EngineError(RuntimeError):
    pass


_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_STOP_CB = C.CFUNCTYPE(C.c_int, C.c_void_p)
_FAMILY_FILL_CB = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int64, _dp, _dp, _dp, _dp, _dp, _i32p, _dp)


class NirLp(C.Structure):
    _fields_ = [
        ("nrow", C.c_int32), ("ncol", C.c_int32),
        ("col_start", _i64p), ("row_index", _i32p), ("value", _dp),
        ("cost", _dp), ("collo", _dp), ("colup", _dp), ("rowlo", _dp), ("rowup", _dp),
        ("obj_offset", C.c_double),
    ]


class NirParams(C.Structure):
    _fields_ = [
        ("mode", C.c_int32), ("precision", C.c_int32), ("max_iters", C.c_int64),
        ("tol", C.c_double), ("polish_tol", C.c_double), ("eval_freq", C.c_int32),
        ("check_termination", C.c_int32), ("use_graph", C.c_int32), ("verbose", C.c_int32),
        ("tile_nnz", C.c_int32), ("tile_rows", C.c_int32), ("seg_width", C.c_int32),
        ("seg_width_t", C.c_int32),
        ("x0", _dp), ("y0", _dp), ("unscale_row", _dp), ("unscale_col", _dp),
        ("should_stop", _STOP_CB), ("stop_user", C.c_void_p), ("primal_weight", C.c_double),
        ("seg_fused", C.c_int32),
    ]


class NirInfo(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("restarts", C.c_int32), ("iterations", C.c_int64),
        ("polish_iterations", C.c_int64), ("objective", C.c_double), ("tau", C.c_double),
        ("sigma", C.c_double), ("primal_weight", C.c_double), ("max_primal_viol", C.c_double),
        ("rel_primal", C.c_double), ("rel_dual", C.c_double), ("rel_gap", C.c_double),
        ("setup_seconds", C.c_double), ("solve_seconds", C.c_double),
    ]


class NirTiming(C.Structure):
    _fields_ = [
        ("iterations", C.c_int64), ("total_ms", C.c_double), ("spmv_ms", C.c_double),
        ("spmtv_ms", C.c_double), ("spmv_bytes", C.c_double), ("spmtv_bytes", C.c_double),
        ("iter_bytes", C.c_double), ("spmv_passes", C.c_int32), ("spmtv_passes", C.c_int32),
    ]


class NirBatchOpts(C.Structure):
    _fields_ = [("threads", C.c_int32), ("quantum", C.c_int32), ("rounds_per_graph", C.c_int32),
                ("workgroups", C.c_int32)]


class NirBatchStats(C.Structure):
    _fields_ = [
        ("count", C.c_int32), ("threads", C.c_int32), ("workgroups", C.c_int32), ("quantum", C.c_int32),
        ("launches", C.c_int64), ("total_iterations", C.c_int64), ("nnz_iterations", C.c_double),
        ("total_nnz", C.c_double), ("setup_seconds", C.c_double), ("solve_seconds", C.c_double),
    ]


class NirVerifyReport(C.Structure):
    _fields_ = [("max_bound_viol", C.c_double), ("max_row_viol", C.c_double), ("computed_obj", C.c_double),
                ("obj_rel_err", C.c_double), ("worst_col", C.c_int32), ("worst_col_upper", C.c_int32),
                ("worst_row", C.c_int32), ("worst_row_upper", C.c_int32), ("nonfinite_col", C.c_int32),
                ("passed", C.c_int32)]


_FAMILY_DONE_CB = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int64, _dp, _dp, C.POINTER(NirInfo))


class NirFamilyOpts(C.Structure):
    _fields_ = [("lanes", C.c_int32), ("handoff", C.c_int32)]


class NirFamilyStats(C.Structure):
    _fields_ = [("lanes", C.c_int32), ("stopped", C.c_int32), ("instances", C.c_int64), ("blocks", C.c_int64),
                ("refills", C.c_int64), ("total_iterations", C.c_int64), ("setup_seconds", C.c_double),
                ("solve_seconds", C.c_double), ("compactions", C.c_int64), ("group_blocks", C.c_int64),
                ("prepare_seconds", C.c_double), ("wait_seconds", C.c_double), ("handoffs", C.c_int64),
                ("handoff_seconds", C.c_double), ("tails", C.c_int64)]


class NirPresolveInfo(C.Structure):
    _fields_ = [("status", C.c_int32), ("solved", C.c_int32), ("nrow", C.c_int32), ("ncol", C.c_int32),
                ("rows_removed", C.c_int32), ("cols_removed", C.c_int32), ("tightenings", C.c_int32),
                ("reserved", C.c_int32), ("obj_offset", C.c_double)]


class NirDevLp(C.Structure):
    """Device-resident CSR LP of nir_verify_solution_noalloc (device addresses)."""
    _fields_ = [("nrow", C.c_int32), ("ncol", C.c_int32), ("row_start", C.c_void_p), ("col_index", C.c_void_p),
                ("value", C.c_void_p), ("cost", C.c_void_p), ("collo", C.c_void_p), ("colup", C.c_void_p),
                ("rowlo", C.c_void_p), ("rowup", C.c_void_p), ("obj_offset", C.c_double)]


_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load bin/libsov.so once; raise EngineUnavailable when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise EngineUnavailable(f"{path} is not built (run __graft_entry__.build())")
    lib = C.CDLL(path)
    vp = C.c_void_p
    lib.nir_params_default.argtypes = [C.POINTER(NirParams)]
    lib.nir_last_error.restype = C.c_char_p
    lib.nir_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
    lib.nir_ctx_destroy.argtypes = [vp]
    lib.nir_ctx_error.argtypes = [vp]
    lib.nir_ctx_error.restype = C.c_char_p
    lib.nir_session_create.argtypes = [vp, C.POINTER(NirLp), C.POINTER(NirParams), C.POINTER(vp)]
    lib.nir_session_run.argtypes = [vp, C.c_int64]
    lib.nir_session_solve.argtypes = [vp]
    lib.nir_session_result.argtypes = [vp, _dp, _dp, C.POINTER(NirInfo)]
    lib.nir_session_time.argtypes = [vp, C.c_int64, C.POINTER(NirTiming)]
    lib.nir_session_destroy.argtypes = [vp]
    lib.nir_session_shape.argtypes = [vp, _i32p]
    lib.nir_session_split.argtypes = [vp, _i32p]
    lib.nir_session_kkt.argtypes = [vp, _dp]
    lib.nir_session_gather_floor.argtypes = [vp, C.c_int32, C.c_int32, _dp]
    lib.nir_session_scaling.argtypes = [vp, _dp, _dp]
    lib.nir_calibrate_copy.argtypes = [vp, C.c_int64, _dp]
    lib.nir_calibrate_gather.argtypes = [vp, C.c_int64, C.c_int32, C.c_int32, C.c_int32, _dp]
    lib.nir_pdhg_solve.argtypes = [vp, C.POINTER(NirLp), C.POINTER(NirParams), _dp, _dp, C.POINTER(NirInfo)]
    for name in ("nir_gen_equality_lp", "nir_gen_mixed_lp", "nir_gen_staircase_lp"):
        getattr(lib, name).restype = vp
    lib.nir_batch_opts_default.argtypes = [C.POINTER(NirBatchOpts)]
    lib.nir_batch_create.argtypes = [vp, C.POINTER(C.POINTER(NirLp)), C.c_int32, C.POINTER(NirParams),
                                     C.POINTER(NirBatchOpts), C.POINTER(vp)]
    lib.nir_batch_create_warm.argtypes = [vp, C.POINTER(C.POINTER(NirLp)), C.c_int32, C.POINTER(_dp),
                                          C.POINTER(_dp), C.POINTER(NirParams), C.POINTER(NirBatchOpts),
                                          C.POINTER(vp)]
    lib.nir_batch_solve.argtypes = [vp]
    lib.nir_batch_result.argtypes = [vp, C.c_int32, _dp, _dp, C.POINTER(NirInfo)]
    lib.nir_batch_get_stats.argtypes = [vp, C.POINTER(NirBatchStats)]
    lib.nir_batch_destroy.argtypes = [vp]
    lib.nir_gen_planning_dims.argtypes = [C.c_uint64, _i32p, _i32p]
    _fp = C.POINTER(C.c_float)
    lib.nir_mlp_create.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_uint64, C.POINTER(vp)]
    lib.nir_mlp_num_params.argtypes = [vp]
    lib.nir_mlp_num_params.restype = C.c_int64
    lib.nir_mlp_get_params.argtypes = [vp, _fp]
    lib.nir_mlp_set_params.argtypes = [vp, _fp]
    lib.nir_mlp_forward.argtypes = [vp, _fp, C.c_int32, _fp]
    lib.nir_mlp_gradient.argtypes = [vp, _fp, _fp, C.c_int32, _fp, _dp]
    lib.nir_mlp_set_data.argtypes = [vp, _fp, _fp, C.c_int32]
    lib.nir_mlp_train.argtypes = [vp, C.c_int32, C.c_int32, C.c_float, C.c_uint64, _fp]
    lib.nir_mlp_train_stats.argtypes = [vp, C.POINTER(C.c_int64), _dp, _dp]
    lib.nir_mlp_destroy.argtypes = [vp]
    lib.nir_mlp_gemm_check.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, _fp, _fp, C.c_int32, C.c_int32, _fp]
    lib.nir_verify_solution.argtypes = [vp, C.POINTER(NirLp), _dp, C.c_double, C.c_double,
                                        C.POINTER(NirVerifyReport)]
    lib.nir_family_opts_default.argtypes = [C.POINTER(NirFamilyOpts)]
    lib.nir_family_create.argtypes = [vp, C.POINTER(NirLp), C.POINTER(NirParams), C.POINTER(NirFamilyOpts),
                                      C.POINTER(vp)]
    lib.nir_family_solve.argtypes = [vp, C.c_int64, _FAMILY_FILL_CB, _FAMILY_DONE_CB, vp]
    lib.nir_family_get_stats.argtypes = [vp, C.POINTER(NirFamilyStats)]
    lib.nir_family_destroy.argtypes = [vp]
    lib.nir_lp_presolve.argtypes = [C.POINTER(NirLp), C.POINTER(C.c_int8), C.c_int32, C.POINTER(NirPresolveInfo),
                                    _dp, _i32p, _i32p, _dp, _dp, _dp, _dp]
    lib.nir_basis_guess.argtypes = [C.POINTER(NirLp), _dp, C.c_double, C.POINTER(C.c_int8), _i32p]
    lib.nir_verify_solution_noalloc.argtypes = [vp, C.POINTER(NirDevLp), vp, C.c_double, C.c_double,
                                                C.POINTER(NirVerifyReport)]
    lib.nir_pod_gram.argtypes = [vp, _dp, C.c_int32, C.c_int32, _dp]
    lib.nir_pod_project.argtypes = [vp, _dp, C.c_int32, C.c_int32, _dp, C.c_int32, _dp]
    lib.nir_pod_create.argtypes = [vp, _dp, C.c_int32, C.c_int32, _dp, C.c_int32, _dp, C.POINTER(vp)]
    lib.nir_pod_modes.argtypes = [vp, _dp]
    lib.nir_pod_decode.argtypes = [vp, _dp, C.c_int32, _dp]
    lib.nir_pod_encode.argtypes = [vp, _dp, C.c_int32, _dp]
    lib.nir_pod_destroy.argtypes = [vp]
    lib.nir_family_set_device_starts.argtypes = [vp, vp, vp]
    lib.nir_gen_equality_lp.argtypes = [C.c_uint64, C.c_int32, C.c_int32, C.c_int32]
    lib.nir_gen_mixed_lp.argtypes = [C.c_uint64, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double]
    lib.nir_gen_staircase_lp.argtypes = [C.c_uint64, C.c_int32, C.c_int32]
    lib.nir_owned_lp_view.argtypes = [vp, C.POINTER(NirLp)]
    lib.nir_owned_lp_nnz.argtypes = [vp]
    lib.nir_owned_lp_nnz.restype = C.c_int64
    lib.nir_owned_lp_x0.argtypes = [vp]
    lib.nir_owned_lp_x0.restype = _dp
    lib.nir_owned_lp_y0.argtypes = [vp]
    lib.nir_owned_lp_y0.restype = _dp
    lib.nir_owned_lp_free.argtypes = [vp]
    lib.nir_lp_write_lpb.argtypes = [C.POINTER(NirLp), C.c_char_p]
    lib.nir_lp_read_lpb.argtypes = [C.c_char_p]
    lib.nir_lp_read_lpb.restype = vp
    _lib = lib
    return lib


def _ptr(a: Optional[np.ndarray], ctype):
    if a is None:
        return C.cast(None, C.POINTER(ctype))
    return a.ctypes.data_as(C.POINTER(ctype))


@dataclass
class LP:
    """An LP in the reference's CSC layout (numpy arrays, owned)."""
    m: int
    n: int
    start: np.ndarray   # int64, n + 1
    index: np.ndarray   # int32, nnz
    value: np.ndarray   # float64, nnz
    c: np.ndarray
    l: np.ndarray
    u: np.ndarray
    rl: np.ndarray
    ru: np.ndarray
    offset: float = 0.0
    x0: Optional[np.ndarray] = None   # generator's feasible point, if any
    y0: Optional[np.ndarray] = None

    @property
    def nnz(self) -> int:
        return int(self.start[-1]) if self.n > 0 else 0

    def struct(self) -> NirLp:
        for name in ("start", "index", "value", "c", "l", "u", "rl", "ru"):
            a = getattr(self, name)
            if not a.flags["C_CONTIGUOUS"]:
                setattr(self, name, np.ascontiguousarray(a))
        s = NirLp()
        s.nrow, s.ncol = self.m, self.n
        s.col_start = _ptr(self.start, C.c_int64)
        s.row_index = _ptr(self.index, C.c_int32)
        s.value = _ptr(self.value, C.c_double)
        s.cost = _ptr(self.c, C.c_double)
        s.collo = _ptr(self.l, C.c_double)
        s.colup = _ptr(self.u, C.c_double)
        s.rowlo = _ptr(self.rl, C.c_double)
        s.rowup = _ptr(self.ru, C.c_double)
        s.obj_offset = float(self.offset)
        return s

    @staticmethod
    def from_dense(A, c, l, u, rl, ru, offset=0.0) -> "LP":
        A = np.asarray(A, dtype=np.float64)
        m, n = A.shape
        start, index, value = [0], [], []
        for j in range(n):
            for i in range(m):
                if A[i, j] != 0.0:
                    index.append(i)
                    value.append(A[i, j])
            start.append(len(index))
        f = lambda v: np.ascontiguousarray(np.asarray(v, dtype=np.float64))
        return LP(m, n, np.asarray(start, np.int64), np.asarray(index, np.int32), f(value) if value else np.zeros(0),
                  f(c), f(l), f(u), f(rl), f(ru), offset)

    def dense(self) -> np.ndarray:
        A = np.zeros((self.m, self.n))
        for j in range(self.n):
            for p in range(self.start[j], self.start[j + 1]):
                A[self.index[p], j] = self.value[p]
        return A


def _from_owned(lib, g, what="generator rejected its arguments") -> LP:
    if not g:
        raise EngineError(what)
    v = NirLp()
    lib.nir_owned_lp_view(g, C.byref(v))
    m, n = v.nrow, v.ncol
    nnz = lib.nir_owned_lp_nnz(g)
    arr = lambda p, k, t: np.ctypeslib.as_array(p, shape=(k,)).astype(t, copy=True) if k > 0 else np.zeros(0, t)
    x0p, y0p = lib.nir_owned_lp_x0(g), lib.nir_owned_lp_y0(g)
    lp = LP(m, n, arr(v.col_start, n + 1, np.int64), arr(v.row_index, nnz, np.int32), arr(v.value, nnz, np.float64),
            arr(v.cost, n, np.float64), arr(v.collo, n, np.float64), arr(v.colup, n, np.float64),
            arr(v.rowlo, m, np.float64), arr(v.rowup, m, np.float64), float(v.obj_offset),
            arr(x0p, n, np.float64) if x0p else None, arr(y0p, m, np.float64) if y0p else None)
    lib.nir_owned_lp_free(g)
    return lp


def write_lp_binary(lp: LP, path: str) -> None:
    """Write an engine LP as a binary LP container (.nlpb)."""
    lib = load_library()
    v = lp.struct()
    if lib.nir_lp_write_lpb(C.byref(v), path.encode("utf-8")) != 0:
        raise EngineError("nir_lp_write_lpb: " + lib.nir_last_error().decode())


def read_lp_binary(path: str) -> LP:
    """Read a binary LP container (.nlpb) as an engine LP (min form)."""
    lib = load_library()
    g = lib.nir_lp_read_lpb(path.encode("utf-8"))
    if not g:
        raise EngineError("nir_lp_read_lpb: " + lib.nir_last_error().decode())
    return _from_owned(lib, g)


def gen_equality_lp(seed: int, m: int, n: int, per_col: int = 5) -> LP:
    lib = load_library()
    return _from_owned(lib, lib.nir_gen_equality_lp(seed, m, n, per_col))


def gen_mixed_lp(seed: int, m: int, n: int, mean: float = 6.0, bad_frac: float = 0.01,
                 bad_scale: float = 1e3) -> LP:
    lib = load_library()
    return _from_owned(lib, lib.nir_gen_mixed_lp(seed, m, n, mean, bad_frac, bad_scale))


def gen_staircase_lp(seed: int, T: int, K: int) -> LP:
    lib = load_library()
    return _from_owned(lib, lib.nir_gen_staircase_lp(seed, T, K))


def planning_dims(seed: int):
    """(T, K) of the planning-family member with this seed: T ~ U{4..48}, K ~ U{5..40}."""
    lib = load_library()
    T, K = C.c_int32(), C.c_int32()
    lib.nir_gen_planning_dims(seed, C.byref(T), C.byref(K))
    return T.value, K.value


def gen_planning_lp(seed: int) -> LP:
    """Planning-family member: a staircase LP of planning_dims(seed)."""
    T, K = planning_dims(seed)
    return gen_staircase_lp(seed, T, K)


@dataclass
class Result:
    x: np.ndarray
    y: np.ndarray
    status: int
    iterations: int
    objective: float
    tau: float
    sigma: float
    info: NirInfo


def default_params(**kw) -> NirParams:
    lib = load_library()
    p = NirParams()
    lib.nir_params_default(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


class Engine:
    """One device context (HIP stream) of the engine."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        if self.lib.nir_ctx_create(device, C.byref(h)) != 0:
            raise EngineUnavailable("nir_ctx_create failed: " + self.lib.nir_last_error().decode())
        self.h = h
        # Objects made on this context (sessions, batches, families, pods,
        # networks): closed before the context itself.
        self._children = weakref.WeakSet()

    def _adopt(self, child):
        self._children.add(child)

    def close(self):
        if getattr(self, "h", None):
            for child in list(getattr(self, "_children", ())):
                child.close()
            self.lib.nir_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def error(self) -> str:
        return self.lib.nir_ctx_error(self.h).decode()

    def calibrate_copy(self, nbytes: int) -> float:
        ms = C.c_double()
        if self.lib.nir_calibrate_copy(self.h, int(nbytes), C.byref(ms)) != 0:
            raise EngineError("nir_calibrate_copy: " + self.error())
        return ms.value

    def calibrate_gather(self, n: int, length: int, with_values: int = 1, blocks: int = 2048) -> float:
        """Microseconds per launch gathering an FP64 operand of `length` entries at
        n uniform random indices (with_values 1: times a streamed value each;
        2: the index and value streams only, no gather)."""
        us = C.c_double()
        if self.lib.nir_calibrate_gather(self.h, int(n), int(length), int(with_values), int(blocks),
                                         C.byref(us)) != 0:
            raise EngineError("nir_calibrate_gather: " + self.error())
        return us.value

    def gemm_check(self, A: np.ndarray, B: np.ndarray, transpose_a=False, transpose_b=False) -> np.ndarray:
        """C = A' B'^T on the device with bf16-rounded inputs, where A' is A ([M][K]) or A^T when A is
        given [K][M], and B' is B ([N][K]) or B^T when B is given [K][N]."""
        A = np.ascontiguousarray(A, dtype=np.float32)
        B = np.ascontiguousarray(B, dtype=np.float32)
        M = A.shape[1] if transpose_a else A.shape[0]
        K = A.shape[0] if transpose_a else A.shape[1]
        N = B.shape[1] if transpose_b else B.shape[0]
        out = np.zeros((M, N), dtype=np.float32)
        fp = C.POINTER(C.c_float)
        if self.lib.nir_mlp_gemm_check(self.h, M, N, K, A.ctypes.data_as(fp), B.ctypes.data_as(fp),
                                       int(transpose_a), int(transpose_b), out.ctypes.data_as(fp)) != 0:
            raise EngineError("nir_mlp_gemm_check: " + self.error())
        return out

    def verify(self, lp: LP, x: np.ndarray, claimed_obj: float, tol: float = 1e-6) -> NirVerifyReport:
        """Double-double certificate of x on the device (nir_verify_solution)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        s = lp.struct()
        rep = NirVerifyReport()
        if self.lib.nir_verify_solution(self.h, C.byref(s), _ptr(x, C.c_double), float(claimed_obj), float(tol),
                                        C.byref(rep)) != 0:
            raise EngineError("nir_verify_solution: " + self.error())
        return rep

    def verify_noalloc(self, dev_lp: "NirDevLp", x_dev: int, claimed_obj: float, tol: float = 1e-6) -> NirVerifyReport:
        """nir_verify_solution_noalloc on a device-resident LP and x (device
        addresses owned by the caller)."""
        rep = NirVerifyReport()
        if self.lib.nir_verify_solution_noalloc(self.h, C.byref(dev_lp), C.c_void_p(x_dev), float(claimed_obj),
                                                float(tol), C.byref(rep)) != 0:
            raise EngineError("nir_verify_solution_noalloc: " + self.error())
        return rep

    def pod_gram(self, X: np.ndarray) -> np.ndarray:
        """X^T X for X [len][N] on the device (FP64)."""
        X = np.ascontiguousarray(X, dtype=np.float64)
        G = np.zeros((X.shape[1], X.shape[1]))
        if self.lib.nir_pod_gram(self.h, _ptr(X, C.c_double), X.shape[0], X.shape[1], _ptr(G, C.c_double)) != 0:
            raise EngineError("nir_pod_gram: " + self.error())
        return G

    def pod_project(self, X: np.ndarray, V: np.ndarray) -> np.ndarray:
        """X V for X [len][N], V [N][R] on the device (FP64)."""
        X = np.ascontiguousarray(X, dtype=np.float64)
        V = np.ascontiguousarray(V, dtype=np.float64)
        Y = np.zeros((X.shape[0], V.shape[1]))
        if self.lib.nir_pod_project(self.h, _ptr(X, C.c_double), X.shape[0], X.shape[1], _ptr(V, C.c_double),
                                    V.shape[1], _ptr(Y, C.c_double)) != 0:
            raise EngineError("nir_pod_project: " + self.error())
        return Y

    def pod_basis(self, X: np.ndarray, V: np.ndarray, mean: np.ndarray) -> "DevicePod":
        """A POD basis kept on the device: modes X V for centred snapshots
        X [len][N] and V [N][R], and the mean (nir_pod_*)."""
        return DevicePod(self, X, V, mean)

    def mlp(self, n_in: int, hidden: int, n_out: int, max_rows: int, seed: int = 0) -> "Mlp":
        return Mlp(self, n_in, hidden, n_out, max_rows, seed)

    def batch(self, lps, params: NirParams, threads: int = 0, quantum: int = 0, rounds_per_graph: int = 0,
              workgroups: int = 0, x0s=None, y0s=None) -> "Batch":
        return Batch(self, lps, params, threads, quantum, rounds_per_graph, workgroups, x0s, y0s)

    def family(self, base: LP, params: NirParams, lanes: int = 1024, handoff: Optional[int] = None) -> "Family":
        """handoff: at most this many members still running, with none waiting,
        finish on the single-LP engine (None = the engine's default, 0 = never)."""
        return Family(self, base, params, lanes, handoff)

    def session(self, lp: LP, params: Optional[NirParams] = None, x0=None, y0=None) -> "Session":
        return Session(self, lp, params, x0, y0)

    def solve_oneshot(self, lp: LP, params: Optional[NirParams] = None, x0=None, y0=None) -> Result:
        """nir_pdhg_solve: the one-shot entry point the facade calls (small
        pdlp-mode LPs run there as a batch of one)."""
        prm = NirParams.from_buffer_copy(params) if params is not None else default_params()
        keep = []
        if x0 is not None:
            x0 = np.ascontiguousarray(x0, dtype=np.float64)
            keep.append(x0)
            prm.x0 = _ptr(x0, C.c_double)
        if y0 is not None:
            y0 = np.ascontiguousarray(y0, dtype=np.float64)
            keep.append(y0)
            prm.y0 = _ptr(y0, C.c_double)
        lps = lp.struct()
        x = np.zeros(lp.n)
        y = np.zeros(lp.m)
        info = NirInfo()
        if self.lib.nir_pdhg_solve(self.h, C.byref(lps), C.byref(prm), _ptr(x, C.c_double), _ptr(y, C.c_double),
                                   C.byref(info)) != 0:
            raise EngineError("nir_pdhg_solve: " + self.error())
        return Result(x, y, info.status, info.iterations, info.objective, info.tau, info.sigma, info)

    def solve(self, lp: LP, params: Optional[NirParams] = None, x0=None, y0=None) -> Result:
        s = self.session(lp, params, x0, y0)
        try:
            s.solve()
            return s.result()
        finally:
            s.close()


class Session:
    """An LP resident on the device (nir_session_*)."""

    def __init__(self, eng: Engine, lp: LP, params: Optional[NirParams], x0=None, y0=None):
        self.eng, self.lib, self.lp = eng, eng.lib, lp
        eng._adopt(self)
        # A private copy: the warm-start pointers set below must not leak into
        # the caller's parameters (and from there into later sessions).
        self.params = NirParams.from_buffer_copy(params) if params is not None else default_params()
        self._keep = [params]
        if x0 is not None:
            x0 = np.ascontiguousarray(x0, dtype=np.float64)
            self._keep.append(x0)
            self.params.x0 = _ptr(x0, C.c_double)
        if y0 is not None:
            y0 = np.ascontiguousarray(y0, dtype=np.float64)
            self._keep.append(y0)
            self.params.y0 = _ptr(y0, C.c_double)
        self._lps = lp.struct()
        h = C.c_void_p()
        if self.lib.nir_session_create(eng.h, C.byref(self._lps), C.byref(self.params), C.byref(h)) != 0:
            raise EngineError("nir_session_create: " + eng.error())
        self.h = h

    def run(self, iters: int):
        if self.lib.nir_session_run(self.h, int(iters)) != 0:
            raise EngineError("nir_session_run: " + self.eng.error())

    def solve(self):
        if self.lib.nir_session_solve(self.h) != 0:
            raise EngineError("nir_session_solve: " + self.eng.error())

    def result(self) -> Result:
        x = np.zeros(self.lp.n)
        y = np.zeros(self.lp.m)
        info = NirInfo()
        if self.lib.nir_session_result(self.h, _ptr(x, C.c_double), _ptr(y, C.c_double), C.byref(info)) != 0:
            raise EngineError("nir_session_result: " + self.eng.error())
        return Result(x, y, info.status, info.iterations, info.objective, info.tau, info.sigma, info)

    def time(self, iters: int) -> NirTiming:
        t = NirTiming()
        if self.lib.nir_session_time(self.h, int(iters), C.byref(t)) != 0:
            raise EngineError("nir_session_time: " + self.eng.error())
        return t

    def shape(self):
        """(A tile, A rows, A segments, A fused, A^T tile, A^T rows, A^T segments, A^T fused)."""
        out = (C.c_int32 * 8)()
        self.lib.nir_session_shape(self.h, out)
        return tuple(out)

    def split(self):
        """(split rows of A, their pieces, split rows of A^T, their pieces)."""
        out = (C.c_int32 * 4)()
        self.lib.nir_session_split(self.h, out)
        return tuple(out)

    def kkt(self) -> np.ndarray:
        """Relative KKT of a reference-mode session's current iterate, on the
        device: [rel primal, rel dual, rel gap, primal objective, dual objective]."""
        out = np.zeros(5)
        if self.lib.nir_session_kkt(self.h, out.ctypes.data_as(C.POINTER(C.c_double))) != 0:
            raise EngineError("nir_session_kkt: " + self.eng.error())
        return out

    def scaling(self):
        """(R, C): the pdlp engine's own Ruiz + Pock-Chambolle factors."""
        R = np.zeros(self.lp.m)
        Cc = np.zeros(self.lp.n)
        if self.lib.nir_session_scaling(self.h, _ptr(R, C.c_double), _ptr(Cc, C.c_double)) != 0:
            raise EngineError("nir_session_scaling: " + self.eng.error())
        return R, Cc

    def gather_floor(self, which: int, reps: int = 20) -> float:
        """Microseconds per product (0: A xbar, 1: A^T y) of the access-only
        kernel over this session's layout (reference mode)."""
        us = C.c_double()
        if self.lib.nir_session_gather_floor(self.h, int(which), int(reps), C.byref(us)) != 0:
            raise EngineError("nir_session_gather_floor: " + self.eng.error())
        return us.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.nir_session_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class DevicePod:
    """A POD basis resident on the device (nir_pod_*): decode writes start
    points on the device, where the family engine reads them."""

    def __init__(self, eng: Engine, X: np.ndarray, V: np.ndarray, mean: np.ndarray):
        self.eng, self.lib = eng, eng.lib
        eng._adopt(self)
        X = np.ascontiguousarray(X, dtype=np.float64)
        V = np.ascontiguousarray(V, dtype=np.float64)
        mean = np.ascontiguousarray(mean, dtype=np.float64)
        self.len, self.rank = X.shape[0], V.shape[1]
        if V.shape[0] != X.shape[1] or mean.shape != (self.len,):
            raise ValueError("pod_basis: X [len][N], V [N][R] and mean [len] do not match")
        h = C.c_void_p()
        if self.lib.nir_pod_create(eng.h, _ptr(X, C.c_double), X.shape[0], X.shape[1], _ptr(V, C.c_double),
                                   self.rank, _ptr(mean, C.c_double), C.byref(h)) != 0:
            raise EngineError("nir_pod_create: " + eng.error())
        self.h = h
        self.rows = 0   # rows of the last decode held on the device

    def modes(self) -> np.ndarray:
        out = np.zeros((self.len, self.rank))
        if self.lib.nir_pod_modes(self.h, _ptr(out, C.c_double)) != 0:
            raise EngineError("nir_pod_modes: " + self.eng.error())
        return out

    def decode(self, alpha: np.ndarray, to_host: bool = True):
        """Rows mean + alpha_i modes^T for alpha [count][R], computed and kept
        on the device; returned as [count][len] when to_host."""
        alpha = np.ascontiguousarray(np.atleast_2d(alpha), dtype=np.float64)
        if alpha.shape[1] != self.rank:
            raise ValueError("decode: alpha must have rank columns")
        out = np.zeros((alpha.shape[0], self.len)) if to_host else None
        if self.lib.nir_pod_decode(self.h, _ptr(alpha, C.c_double), alpha.shape[0],
                                   _ptr(out, C.c_double) if to_host else None) != 0:
            raise EngineError("nir_pod_decode: " + self.eng.error())
        self.rows = alpha.shape[0]
        return out

    def encode(self, S: np.ndarray) -> np.ndarray:
        S = np.ascontiguousarray(np.atleast_2d(S), dtype=np.float64)
        coef = np.zeros((S.shape[0], self.rank))
        if self.lib.nir_pod_encode(self.h, _ptr(S, C.c_double), S.shape[0], _ptr(coef, C.c_double)) != 0:
            raise EngineError("nir_pod_encode: " + self.eng.error())
        return coef

    def close(self):
        if getattr(self, "h", None):
            self.lib.nir_pod_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class Family:
    """Members of a parametric family (shared A and column bounds, own costs
    and row bounds) streamed through the family engine (nir_family_*)."""

    def __init__(self, eng: Engine, base: LP, params: NirParams, lanes: int = 1024, handoff: Optional[int] = None):
        self.eng, self.lib, self.base = eng, eng.lib, base
        eng._adopt(self)
        self.params = params
        self._s = base.struct()
        opts = NirFamilyOpts()
        self.lib.nir_family_opts_default(C.byref(opts))
        opts.lanes = lanes
        if handoff is not None:
            opts.handoff = handoff
        h = C.c_void_p()
        if self.lib.nir_family_create(eng.h, C.byref(self._s), C.byref(params), C.byref(opts), C.byref(h)) != 0:
            raise EngineError("nir_family_create: " + eng.error())
        self.h = h

    def set_device_starts(self, pod_x: Optional["DevicePod"], pod_y: Optional["DevicePod"]):
        """Members i below the rows of the pods' last decode start from those
        rows, read on the device (nir_family_set_device_starts)."""
        if self.lib.nir_family_set_device_starts(self.h, pod_x.h if pod_x is not None else None,
                                                 pod_y.h if pod_y is not None else None) != 0:
            raise EngineError("nir_family_set_device_starts: " + self.eng.error())

    def solve(self, count: int, member, on_result=None):
        """Solves members 0 .. count-1. member(i) returns (cost, rowlo, rowup),
        optionally followed by x0, y0 (each possibly None) and an initial
        primal weight (0 or None = the parameters' own);
        on_result(i, x, y, info) receives each solution as it finishes
        (copies); without it the results are returned as a list of Result."""
        m, n = self.base.m, self.base.n
        out = [None] * count if on_result is None else None
        err = []

        def fill(_u, idx, cost, rl, ru, x0, y0, warm, weight):
            try:
                v = member(int(idx))
                C.memmove(cost, np.ascontiguousarray(v[0], np.float64).ctypes.data, 8 * n)
                C.memmove(rl, np.ascontiguousarray(v[1], np.float64).ctypes.data, 8 * m)
                C.memmove(ru, np.ascontiguousarray(v[2], np.float64).ctypes.data, 8 * m)
                w = 0
                if len(v) > 3 and v[3] is not None:
                    C.memmove(x0, np.ascontiguousarray(v[3], np.float64).ctypes.data, 8 * n)
                    w |= 1
                if len(v) > 4 and v[4] is not None:
                    C.memmove(y0, np.ascontiguousarray(v[4], np.float64).ctypes.data, 8 * m)
                    w |= 2
                warm[0] = w
                if len(v) > 5 and v[5]:
                    weight[0] = float(v[5])
                return 0
            except Exception as e:   # reported after the call returns
                err.append(e)
                return 1

        def done(_u, idx, x, y, info):
            try:
                xs = np.ctypeslib.as_array(x, shape=(n,)).copy() if n else np.zeros(0)
                ys = np.ctypeslib.as_array(y, shape=(m,)).copy() if m else np.zeros(0)
                inf = NirInfo.from_buffer_copy(info.contents)
                if on_result is None:
                    out[int(idx)] = Result(xs, ys, inf.status, inf.iterations, inf.objective, inf.tau, inf.sigma, inf)
                else:
                    on_result(int(idx), xs, ys, inf)
                return 0
            except Exception as e:
                err.append(e)
                return 1

        fcb, dcb = _FAMILY_FILL_CB(fill), _FAMILY_DONE_CB(done)
        rc = self.lib.nir_family_solve(self.h, int(count), fcb, dcb, None)
        if err:
            raise err[0]
        if rc != 0:
            raise EngineError("nir_family_solve: " + self.eng.error())
        return out

    def stats(self) -> NirFamilyStats:
        st = NirFamilyStats()
        if self.lib.nir_family_get_stats(self.h, C.byref(st)) != 0:
            raise EngineError("nir_family_get_stats: " + self.eng.error())
        return st

    def close(self):
        if getattr(self, "h", None):
            self.lib.nir_family_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class Batch:
    """Many LPs resident on the device, solved by the persistent kernel (nir_batch_*)."""

    def __init__(self, eng: Engine, lps, params: NirParams, threads=0, quantum=0, rounds_per_graph=0, workgroups=0,
                 x0s=None, y0s=None):
        self.eng, self.lib, self.lps = eng, eng.lib, list(lps)
        eng._adopt(self)
        self.params = params
        self._structs = [lp.struct() for lp in self.lps]
        arr = (C.POINTER(NirLp) * len(self._structs))(*[C.pointer(s) for s in self._structs])
        opts = NirBatchOpts()
        self.lib.nir_batch_opts_default(C.byref(opts))
        opts.threads, opts.quantum, opts.rounds_per_graph, opts.workgroups = threads, quantum, rounds_per_graph, \
            workgroups
        h = C.c_void_p()
        if x0s is None and y0s is None:
            rc = self.lib.nir_batch_create(eng.h, arr, len(self._structs), C.byref(params), C.byref(opts), C.byref(h))
        else:
            self._keep = []

            def ptrs(vs):
                out = (_dp * len(self._structs))()
                for i, v in enumerate(vs if vs is not None else [None] * len(self._structs)):
                    if v is not None:
                        a = np.ascontiguousarray(v, dtype=np.float64)
                        self._keep.append(a)
                        out[i] = a.ctypes.data_as(_dp)
                return out
            rc = self.lib.nir_batch_create_warm(eng.h, arr, len(self._structs), ptrs(x0s), ptrs(y0s),
                                                C.byref(params), C.byref(opts), C.byref(h))
        if rc != 0:
            raise EngineError("nir_batch_create: " + eng.error())
        self.h = h

    def solve(self):
        if self.lib.nir_batch_solve(self.h) != 0:
            raise EngineError("nir_batch_solve: " + self.eng.error())

    def result(self, i: int) -> Result:
        lp = self.lps[i]
        x = np.zeros(lp.n)
        y = np.zeros(lp.m)
        info = NirInfo()
        if self.lib.nir_batch_result(self.h, i, _ptr(x, C.c_double), _ptr(y, C.c_double), C.byref(info)) != 0:
            raise EngineError("nir_batch_result: " + self.eng.error())
        return Result(x, y, info.status, info.iterations, info.objective, info.tau, info.sigma, info)

    def stats(self) -> NirBatchStats:
        st = NirBatchStats()
        if self.lib.nir_batch_get_stats(self.h, C.byref(st)) != 0:
            raise EngineError("nir_batch_get_stats: " + self.eng.error())
        return st

    def close(self):
        if getattr(self, "h", None):
            self.lib.nir_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class Mlp:
    """The warm-start network on the device (nir_mlp_*): bf16 MFMA forward and backward, Adam."""

    def __init__(self, eng: Engine, n_in: int, hidden: int, n_out: int, max_rows: int, seed: int = 0):
        self.eng, self.lib = eng, eng.lib
        eng._adopt(self)
        self.n_in, self.hidden, self.n_out = n_in, hidden, n_out
        h = C.c_void_p()
        if self.lib.nir_mlp_create(eng.h, n_in, hidden, n_out, max_rows, seed, C.byref(h)) != 0:
            raise EngineError("nir_mlp_create: " + eng.error())
        self.h = h
        self.count = int(self.lib.nir_mlp_num_params(h))

    @staticmethod
    def _f(a):
        return np.ascontiguousarray(a, dtype=np.float32)

    def _fp(self, a):
        return a.ctypes.data_as(C.POINTER(C.c_float))

    def params(self) -> np.ndarray:
        out = np.zeros(self.count, dtype=np.float32)
        if self.lib.nir_mlp_get_params(self.h, self._fp(out)) != 0:
            raise EngineError("nir_mlp_get_params: " + self.eng.error())
        return out

    def set_params(self, p: np.ndarray):
        p = self._f(p)
        if p.size != self.count or self.lib.nir_mlp_set_params(self.h, self._fp(p)) != 0:
            raise EngineError("nir_mlp_set_params: " + self.eng.error())

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = self._f(x)
        out = np.zeros((x.shape[0], self.n_out), dtype=np.float32)
        if self.lib.nir_mlp_forward(self.h, self._fp(x), x.shape[0], self._fp(out)) != 0:
            raise EngineError("nir_mlp_forward: " + self.eng.error())
        return out

    def gradient(self, x: np.ndarray, t: np.ndarray):
        x, t = self._f(x), self._f(t)
        g = np.zeros(self.count, dtype=np.float32)
        loss = C.c_double()
        if self.lib.nir_mlp_gradient(self.h, self._fp(x), self._fp(t), x.shape[0], self._fp(g), C.byref(loss)) != 0:
            raise EngineError("nir_mlp_gradient: " + self.eng.error())
        return loss.value, g

    def train(self, x: np.ndarray, t: np.ndarray, epochs: int, batch: int, lr: float, seed: int = 0) -> np.ndarray:
        x, t = self._f(x), self._f(t)
        if self.lib.nir_mlp_set_data(self.h, self._fp(x), self._fp(t), x.shape[0]) != 0:
            raise EngineError("nir_mlp_set_data: " + self.eng.error())
        losses = np.zeros(max(epochs, 1), dtype=np.float32)
        if self.lib.nir_mlp_train(self.h, epochs, batch, lr, seed, self._fp(losses)) != 0:
            raise EngineError("nir_mlp_train: " + self.eng.error())
        return losses[:epochs]

    def train_stats(self):
        rows, ms, fl = C.c_int64(), C.c_double(), C.c_double()
        self.lib.nir_mlp_train_stats(self.h, C.byref(rows), C.byref(ms), C.byref(fl))
        return rows.value, ms.value, fl.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.nir_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def presolve(lp: LP, integer=None, level: int = 2):
    """The facade's presolve on one LP (nir_lp_presolve): a dict with the
    info fields, fixed_x, col_map, row_map and the reduced bounds."""
    lib = load_library()
    integer = None if integer is None else np.ascontiguousarray(integer, dtype=np.int8)
    info = NirPresolveInfo()
    fixed = np.zeros(lp.n)
    cmap = np.zeros(max(lp.n, 1), dtype=np.int32)
    rmap = np.zeros(max(lp.m, 1), dtype=np.int32)
    cl, cu = np.zeros(max(lp.n, 1)), np.zeros(max(lp.n, 1))
    rl, ru = np.zeros(max(lp.m, 1)), np.zeros(max(lp.m, 1))
    s = lp.struct()
    ip = integer.ctypes.data_as(C.POINTER(C.c_int8)) if integer is not None else None
    if lib.nir_lp_presolve(C.byref(s), ip, int(level), C.byref(info), _ptr(fixed, C.c_double),
                           cmap.ctypes.data_as(_i32p), rmap.ctypes.data_as(_i32p), _ptr(cl, C.c_double),
                           _ptr(cu, C.c_double), _ptr(rl, C.c_double), _ptr(ru, C.c_double)) != 0:
        raise EngineError("nir_lp_presolve: " + lib.nir_last_error().decode())
    nc, nr = info.ncol, info.nrow
    return {"status": info.status, "solved": info.solved, "nrow": nr, "ncol": nc,
            "rows_removed": info.rows_removed, "cols_removed": info.cols_removed,
            "tightenings": info.tightenings, "obj_offset": info.obj_offset, "fixed_x": fixed,
            "col_map": cmap[:nc].copy(), "row_map": rmap[:nr].copy(), "collo": cl[:nc].copy(),
            "colup": cu[:nc].copy(), "rowlo": rl[:nr].copy(), "rowup": ru[:nr].copy()}


def basis_guess(lp: LP, x, tol: float = 1e-6):
    """Crossover hook (nir_basis_guess): (status[n + m], basic_index[m]) of a
    simplex starting basis guessed from the point x."""
    lib = load_library()
    x = np.ascontiguousarray(x, dtype=np.float64)
    status = np.zeros(lp.n + lp.m, dtype=np.int8)
    basic = np.zeros(max(lp.m, 1), dtype=np.int32)
    s = lp.struct()
    if lib.nir_basis_guess(C.byref(s), _ptr(x, C.c_double), float(tol), status.ctypes.data_as(C.POINTER(C.c_int8)),
                           basic.ctypes.data_as(_i32p)) != 0:
        raise EngineError("nir_basis_guess: " + lib.nir_last_error().decode())
    return status, basic[:lp.m].copy()
