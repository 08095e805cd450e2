"""CPU oracle for HSMU SpGEMM -- TEST INFRASTRUCTURE ONLY.

ctypes wrapper over liboracle.so (hsmu_oracle.c, a row-wise Gustavson in C with
OpenMP) plus the comparison used by the parity tests.  Only tests/, the smoke check in
__graft_entry__.py and bench.py's cpu_baseline leg may import this package; the product
library never does.

Reference behaviour restated (file:line in wuminqaq/HSMU-SpGEMM):
  products       src/compressed/compressed_NHC_spgemm.cu:191-196
  tile format    src/compressed/compressed_mask64.cuh:61-342
  C tile OR      src/compressed/compressed_tileOR.cuh:20-104
  sorted columns src/compressed/compressed_Ccol_symbolic.cuh:38-62
  accumulation   src/compressed/compressed_Cval.cuh:51-63, src/small/numeric.cuh:233-253
How each function is pinned (golden vectors, closed forms, brute force) is listed in
tests/test_oracle.py.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
if not os.path.exists(LIB_PATH):
    raise ImportError(f"liboracle.so not built at {LIB_PATH}; run `make -C oracle`")
_lib = C.CDLL(LIB_PATH)

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i64 = C.c_int64

_lib.oracle_max_threads.restype = C.c_int
_lib.oracle_products.restype = C.c_int64
_lib.oracle_products.argtypes = [_i64, _i64, _i32p, _i32p, _i32p, C.c_void_p]
_lib.oracle_btile_counts.restype = C.c_int64
_lib.oracle_btile_counts.argtypes = [_i64, _i32p, _i32p, _i32p]
_lib.oracle_btile_fill.restype = C.c_int
_lib.oracle_btile_fill.argtypes = [_i64, _i32p, _i32p, _i64p, _i32p, _u64p]
_lib.oracle_ctile_rows.restype = C.c_int
_lib.oracle_ctile_rows.argtypes = [_i64, _i64, _i64, _i32p, _i32p, _i64p, _i32p, _u64p, _i64p, _i64p, C.c_int]
_lib.oracle_count_rows.restype = C.c_int
_lib.oracle_count_rows.argtypes = [_i64, _i64, _i64, _i32p, _i32p, _i32p, _i32p, _i64p, C.c_int]
_lib.oracle_fill_rows.restype = C.c_int
_lib.oracle_fill_rows.argtypes = [_i64, _i64, _i64, _i32p, _i32p, _f64p, _i32p, _i32p, _f64p, _i64p, _i32p,
                                  _f64p, C.c_void_p, C.c_int]


@dataclass
class Product:
    """Oracle result for rows [r0, r1): int64 row offsets relative to r0."""
    rowptr: np.ndarray
    col: np.ndarray
    val: np.ndarray
    absval: np.ndarray  # sum_j |a_ij * b_jk| per entry: the scale of the FP64 tolerance


def _arrs(m):
    return (np.ascontiguousarray(m.rowptr, dtype=np.int32), np.ascontiguousarray(m.col, dtype=np.int32),
            np.ascontiguousarray(m.val, dtype=np.float64))


def max_threads() -> int:
    return int(_lib.oracle_max_threads())


def products(A, B, row_products: bool = False):
    arp, acol, _ = _arrs(A)
    brp, _, _ = _arrs(B)
    if row_products:
        out = np.zeros(A.M, dtype=np.int64)
        tot = _lib.oracle_products(0, A.M, arp, acol, brp, out.ctypes.data)
        return int(tot), out
    return int(_lib.oracle_products(0, A.M, arp, acol, brp, None))


def btiles(B):
    """(tptr int64[M+1], tcol int32, tmask uint64) with tiles sorted per row."""
    brp, bcol, _ = _arrs(B)
    cnt = np.zeros(B.M, dtype=np.int32)
    total = _lib.oracle_btile_counts(B.M, brp, bcol, cnt)
    if total < 0:
        raise MemoryError("oracle_btile_counts")
    tptr = np.zeros(B.M + 1, dtype=np.int64)
    np.cumsum(cnt, out=tptr[1:])
    tcol = np.zeros(max(int(total), 1), dtype=np.int32)
    tmask = np.zeros(max(int(total), 1), dtype=np.uint64)
    if _lib.oracle_btile_fill(B.M, brp, bcol, tptr, tcol, tmask) != 0:
        raise MemoryError("oracle_btile_fill")
    return tptr, tcol[:total], tmask[:total]


def ctile_rows(A, B, nthreads: int = 0):
    """Per row of C: (number of 64-column tiles, nnz) by OR-ing B's tile masks."""
    arp, acol, _ = _arrs(A)
    tptr, tcol, tmask = btiles(B)
    tcol = np.ascontiguousarray(tcol if len(tcol) else np.zeros(1, np.int32))
    tmask = np.ascontiguousarray(tmask if len(tmask) else np.zeros(1, np.uint64))
    ct = np.zeros(A.M, dtype=np.int64)
    cn = np.zeros(A.M, dtype=np.int64)
    if _lib.oracle_ctile_rows(0, A.M, B.N, arp, acol, tptr, tcol, tmask, ct, cn, nthreads) != 0:
        raise MemoryError("oracle_ctile_rows")
    return ct, cn


def spgemm_rows(A, B, r0: int, r1: int, nthreads: int = 0, want_abs: bool = True) -> Product:
    """Rows [r0, r1) of C = A * B."""
    arp, acol, aval = _arrs(A)
    brp, bcol, bval = _arrs(B)
    n = r1 - r0
    cnt = np.zeros(max(n, 1), dtype=np.int64)
    if _lib.oracle_count_rows(r0, r1, B.N, arp, acol, brp, bcol, cnt, nthreads) != 0:
        raise MemoryError("oracle_count_rows")
    off = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(cnt[:n], out=off[1:])
    nnz = int(off[-1])
    col = np.zeros(max(nnz, 1), dtype=np.int32)
    val = np.zeros(max(nnz, 1), dtype=np.float64)
    absval = np.zeros(max(nnz, 1), dtype=np.float64) if want_abs else None
    rc = _lib.oracle_fill_rows(r0, r1, B.N, arp, acol, aval, brp, bcol, bval, off, col, val,
                               absval.ctypes.data if want_abs else None, nthreads)
    if rc != 0:
        raise RuntimeError(f"oracle_fill_rows failed ({rc})")
    return Product(off, col[:nnz], val[:nnz], absval[:nnz] if want_abs else None)


def spgemm(A, B, nthreads: int = 0, want_abs: bool = True) -> Product:
    return spgemm_rows(A, B, 0, A.M, nthreads, want_abs)


def compare(C, ref: Product, r0: int = 0, r1: int | None = None, tol: float = 1e-12) -> dict:
    """Check a result C (rowptr/col/val, rows r0..r1 of the full product) against the
    oracle: row pointers and columns bit-exact, values |c - c_ref| <= tol * sum|a*b|
    (an entry whose products cancel exactly must be within tol * its absolute sum).
    Returns the error summary; raises AssertionError on a mismatch."""
    r1 = C.M if r1 is None else r1
    rp = np.asarray(C.rowptr, dtype=np.int64)
    base = int(rp[r0])
    got_off = rp[r0:r1 + 1] - base
    if not np.array_equal(got_off, ref.rowptr):
        bad = int(np.nonzero(got_off != ref.rowptr)[0][0])
        raise AssertionError(f"row pointers differ first at row {r0 + bad}: {got_off[bad]} vs {ref.rowptr[bad]}")
    nnz = int(ref.rowptr[-1])
    col = np.asarray(C.col)[base:base + nnz]
    if not np.array_equal(col, ref.col):
        bad = int(np.nonzero(col != ref.col)[0][0])
        raise AssertionError(f"columns differ first at entry {base + bad}: {col[bad]} vs {ref.col[bad]}")
    val = np.asarray(C.val, dtype=np.float64)[base:base + nnz]
    err = np.abs(val - ref.val)
    scale = ref.absval if ref.absval is not None else np.abs(ref.val)
    bound = tol * scale
    viol = err > bound
    if np.any(viol):
        bad = int(np.nonzero(viol)[0][0])
        raise AssertionError(f"value at entry {base + bad} off by {err[bad]:.3e} > {bound[bad]:.3e} "
                             f"(got {val[bad]!r}, want {ref.val[bad]!r})")
    with np.errstate(divide="ignore", invalid="ignore"):
        rel = np.where(np.abs(ref.val) > 0, err / np.abs(ref.val), 0.0)
    return {"nnz": nnz, "max_abs_err": float(err.max(initial=0.0)),
            "max_rel_err": float(rel.max(initial=0.0)),
            "median_rel_err": float(np.median(rel)) if nnz else 0.0,
            "max_err_over_scale": float((err / np.where(scale > 0, scale, 1.0)).max(initial=0.0))}
