"""ctypes binding of libhsmu_check.so (include/hsmu_check.h): rocSPARSE's SpGEMM on the
same device inputs, and a device-CSR comparison, for cross-checking HSMU SpGEMM.

Reference counterpart: the cuSPARSE check of src/cusparse_in_tile.cu:4-272.
"""
from __future__ import annotations

import ctypes as C
import os

from . import Csr, Context, DeviceMatrix, HsmuError, _HERE, lib as _hsmu_lib

CHECK_LIB_PATH = os.path.join(os.path.dirname(_HERE), "libhsmu_check.so")
if not os.path.exists(CHECK_LIB_PATH):
    raise ImportError(f"libhsmu_check.so not built at {CHECK_LIB_PATH}; run __graft_entry__.build()")
lib = C.CDLL(CHECK_LIB_PATH)


class Diff(C.Structure):
    _fields_ = [("rowptr_mismatch", C.c_int64), ("col_mismatch", C.c_int64), ("val_mismatch", C.c_int64),
                ("first_row", C.c_int64), ("first_entry", C.c_int64), ("max_abs_err", C.c_double),
                ("max_rel_err", C.c_double)]


_P = C.c_void_p
lib.hsmu_rocsparse_spgemm.restype = C.c_int
lib.hsmu_rocsparse_spgemm.argtypes = [_P, C.POINTER(Csr), C.POINTER(Csr), C.POINTER(Csr), C.POINTER(C.c_double),
                                      C.POINTER(C.c_double)]
lib.hsmu_check_free.restype = C.c_int
lib.hsmu_check_free.argtypes = [C.POINTER(Csr)]
lib.hsmu_csr_compare.restype = C.c_int
lib.hsmu_csr_compare.argtypes = [C.POINTER(Csr), C.POINTER(Csr), C.POINTER(Csr), C.c_double, C.POINTER(Diff)]
lib.hsmu_check_last_error.restype = C.c_char_p
lib.hsmu_check_last_error.argtypes = []


def _check(status: int, where: str) -> None:
    if status != 0:
        raise HsmuError(status, f"{where}: {lib.hsmu_check_last_error().decode(errors='replace')}")


def rocsparse_spgemm(ctx: Context, A: DeviceMatrix, B: DeviceMatrix):
    """C = A * B by rocSPARSE on the context's stream: (Csr, ms of its nnz + compute
    stages, host ms of the whole call).  Release C with free()."""
    c = Csr()
    ms_stages = C.c_double(0)
    ms_wall = C.c_double(0)
    stream = _hsmu_lib.hsmu_ctx_stream(ctx.handle)
    _check(lib.hsmu_rocsparse_spgemm(stream, C.byref(A.csr), C.byref(B.csr), C.byref(c), C.byref(ms_stages),
                                     C.byref(ms_wall)), "hsmu_rocsparse_spgemm")
    return c, ms_stages.value, ms_wall.value


def free(c: Csr) -> None:
    _check(lib.hsmu_check_free(C.byref(c)), "hsmu_check_free")


def compare(X: Csr, Y: Csr, S: Csr | None = None, tol: float = 1e-12) -> dict:
    """Row pointers and columns of X and Y exactly, values within tol * scale (S's
    entries when given, else max(|x|, |y|)); returns the mismatch counts."""
    d = Diff()
    _check(lib.hsmu_csr_compare(C.byref(X), C.byref(Y), C.byref(S) if S is not None else None, tol, C.byref(d)),
           "hsmu_csr_compare")
    return {name: getattr(d, name) for name, _ in Diff._fields_}
