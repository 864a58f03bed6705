"""NIRNAYA Python interface over the sov_* C ABI of bin/libsov.so.

Same class, methods, properties and defaults as the reference wrapper
(python/nirnaya.py:10-112): ``NirnayaModel(name, dll_path)`` with
``add_var`` / ``add_constr`` / ``set_sense`` / ``set_option`` / ``optimize``
and ``Status`` / ``ObjVal`` / ``ObjBound`` / ``X``. Set
``set_option("method", "pdhg")`` to solve on the MI355X PDHG engine.
Additions: ``Y`` (row duals), ``Iterations``, ``read_mps``, ``write_mps``,
``read_lpb`` / ``write_lpb`` (binary LP container), ``set_start`` (warm
start) and ``error_message``.
"""
import ctypes
import os
from typing import List, Optional, Sequence


class NirnayaModel:
    CONTINUOUS = 0
    INTEGER = 1
    BINARY = 2

    MINIMIZE = 1
    MAXIMIZE = -1

    def __init__(self, name: str = "nirnaya_model", dll_path: Optional[str] = None):
        if dll_path is None:
            base = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
            dll_path = os.path.join(base, "bin", "libsov.so")
        if not os.path.exists(dll_path):
            raise OSError(f"{dll_path} not found: build the library first")
        self._lib = ctypes.CDLL(dll_path)
        self._bind()
        self._env = self._lib.sov_env_create()
        self._model = self._lib.sov_model_create(self._env, name.encode("utf-8"))
        self._num_cols = 0
        self._num_rows = 0

    def _bind(self):
        L, vp, d, i = self._lib, ctypes.c_void_p, ctypes.c_double, ctypes.c_int
        dp = ctypes.POINTER(ctypes.c_double)
        L.sov_env_create.restype = vp
        L.sov_env_free.argtypes = [vp]
        L.sov_model_create.argtypes = [vp, ctypes.c_char_p]
        L.sov_model_create.restype = vp
        L.sov_model_free.argtypes = [vp]
        L.sov_add_col.argtypes = [vp, d, d, d, i, ctypes.c_char_p]
        L.sov_add_col.restype = i
        L.sov_add_row.argtypes = [vp, d, d, i, ctypes.POINTER(ctypes.c_int), dp, ctypes.c_char_p]
        L.sov_add_row.restype = i
        L.sov_set_sense.argtypes = [vp, i]
        L.sov_set_offset.argtypes = [vp, d]
        L.sov_set_option_string.argtypes = [vp, ctypes.c_char_p, ctypes.c_char_p]
        L.sov_set_option_string.restype = i
        L.sov_optimize.argtypes = [vp]
        L.sov_optimize.restype = i
        L.sov_interrupt.argtypes = [vp]
        L.sov_get_status.argtypes = [vp]
        L.sov_get_status.restype = i
        L.sov_get_obj.argtypes = [vp]
        L.sov_get_obj.restype = d
        L.sov_get_bound.argtypes = [vp]
        L.sov_get_bound.restype = d
        L.sov_get_x.argtypes = [vp, dp]
        L.sov_read_mps.argtypes = [vp, ctypes.c_char_p]
        L.sov_read_mps.restype = i
        L.sov_write_mps.argtypes = [vp, ctypes.c_char_p]
        L.sov_write_mps.restype = i
        L.sov_error_message.argtypes = [vp]
        L.sov_error_message.restype = ctypes.c_char_p
        # Entry points added by this library; absent from other sov_* builds.
        self._has_ext = hasattr(L, "sov_get_y")
        if self._has_ext:
            L.sov_get_y.argtypes = [vp, dp]
            L.sov_get_iterations.argtypes = [vp]
            L.sov_get_iterations.restype = i
            L.sov_set_start.argtypes = [vp, dp, dp]
            L.sov_get_num_cols.argtypes = [vp]
            L.sov_get_num_rows.argtypes = [vp]
            L.sov_read_lpb.argtypes = [vp, ctypes.c_char_p]
            L.sov_read_lpb.restype = i
            L.sov_write_lpb.argtypes = [vp, ctypes.c_char_p]
            L.sov_write_lpb.restype = i

    def add_var(self, cost: float = 0.0, lb: float = 0.0, ub: float = 1e20, vtype: int = 0, name: str = "") -> int:
        idx = self._lib.sov_add_col(self._model, float(cost), float(lb), float(ub), int(vtype), name.encode("utf-8"))
        self._num_cols += 1
        return idx

    def add_constr(self, cols: Sequence[int], vals: Sequence[float], lb: float = -1e20, ub: float = 1e20,
                   name: str = "") -> int:
        nnz = len(cols)
        c_idx = (ctypes.c_int * nnz)(*cols)
        c_val = (ctypes.c_double * nnz)(*vals)
        self._num_rows += 1
        return self._lib.sov_add_row(self._model, float(lb), float(ub), nnz, c_idx, c_val, name.encode("utf-8"))

    def set_sense(self, sense: int):
        self._lib.sov_set_sense(self._model, int(sense))

    def set_offset(self, offset: float):
        self._lib.sov_set_offset(self._model, float(offset))

    def set_option(self, key: str, value) -> int:
        return self._lib.sov_set_option_string(self._model, key.encode("utf-8"), str(value).encode("utf-8"))

    def read_mps(self, path: str) -> int:
        rc = self._lib.sov_read_mps(self._model, path.encode("utf-8"))
        if rc == 0 and self._has_ext:
            self._num_cols = self._lib.sov_get_num_cols(self._model)
            self._num_rows = self._lib.sov_get_num_rows(self._model)
        return rc

    def write_mps(self, path: str) -> int:
        return self._lib.sov_write_mps(self._model, path.encode("utf-8"))

    def read_lpb(self, path: str) -> int:
        rc = self._lib.sov_read_lpb(self._model, path.encode("utf-8"))
        if rc == 0 and self._has_ext:
            self._num_cols = self._lib.sov_get_num_cols(self._model)
            self._num_rows = self._lib.sov_get_num_rows(self._model)
        return rc

    def write_lpb(self, path: str) -> int:
        return self._lib.sov_write_lpb(self._model, path.encode("utf-8"))

    def set_start(self, x: Optional[Sequence[float]] = None, y: Optional[Sequence[float]] = None) -> int:
        xa = (ctypes.c_double * len(x))(*x) if x is not None else None
        ya = (ctypes.c_double * len(y))(*y) if y is not None else None
        return self._lib.sov_set_start(self._model, xa, ya)

    def optimize(self) -> int:
        return self._lib.sov_optimize(self._model)

    def interrupt(self) -> int:
        return self._lib.sov_interrupt(self._model)

    @property
    def Status(self) -> int:
        return self._lib.sov_get_status(self._model)

    @property
    def ObjVal(self) -> float:
        return self._lib.sov_get_obj(self._model)

    @property
    def ObjBound(self) -> float:
        return self._lib.sov_get_bound(self._model)

    @property
    def X(self) -> List[float]:
        arr = (ctypes.c_double * self._num_cols)()
        self._lib.sov_get_x(self._model, arr)
        return list(arr)

    @property
    def Y(self) -> List[float]:
        arr = (ctypes.c_double * self._num_rows)()
        self._lib.sov_get_y(self._model, arr)
        return list(arr)

    @property
    def Iterations(self) -> int:
        return self._lib.sov_get_iterations(self._model)

    @property
    def error_message(self) -> str:
        return self._lib.sov_error_message(self._model).decode("utf-8", "replace")

    def __del__(self):
        if getattr(self, "_model", None):
            self._lib.sov_model_free(self._model)
            self._model = None
        if getattr(self, "_env", None):
            self._lib.sov_env_free(self._env)
            self._env = None
