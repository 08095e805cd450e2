"""ctypes binding of libhsmu.so (the MI355X HSMU SpGEMM C ABI, include/hsmu.h).

The binding carries host arrays (numpy) to and from device memory through the library's
own helpers; it has no compute path of its own.  If libhsmu.so is missing the import
fails loudly: there is no CPU fallback.

Reference counterparts: ``Context.spgemm`` <- NHC_spgemm (src/NHC_spgemm.cuh:7-43),
``Context.btiles`` <- Initialize_compressed_mask64_matrix
(src/compressed/compressed_mask64.cuh:343-424), the ``gen_*`` helpers produce the
synthetic workloads, ``read_mtx`` <- mmio_allinone (src/mmio_highlevel.h:18-193).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libhsmu.so")
HEADER_PATH = os.path.join(os.path.dirname(os.path.dirname(_HERE)), "include", "hsmu.h")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"libhsmu.so not built at {LIB_PATH}; run __graft_entry__.build()")

lib = C.CDLL(LIB_PATH)

HSMU_OK = 0
HSMU_ERR_INVALID = 1
HSMU_ERR_UNSORTED = 2
HSMU_ERR_INDEX = 3
HSMU_ERR_OVERFLOW = 4
HSMU_ERR_NOMEM = 5
HSMU_ERR_HIP = 6
HSMU_ERR_UNSUPPORTED = 7
HSMU_ERR_PLAN = 8
HSMU_FP64 = 0
HSMU_FP32_COMP = 1
NBINS = 10


class HsmuError(RuntimeError):
    def __init__(self, status: int, where: str):
        self.status = status
        msg = lib.hsmu_status_string(status).decode()
        detail = lib.hsmu_last_error().decode()
        super().__init__(f"{where}: {msg} ({status}): {detail}")


class Csr(C.Structure):
    _fields_ = [("M", C.c_int64), ("N", C.c_int64), ("nnz", C.c_int64), ("rowptr", C.c_void_p),
                ("col", C.c_void_p), ("val", C.c_void_p), ("rowptr64", C.c_int32), ("val32", C.c_int32)]


class Tiles(C.Structure):
    _fields_ = [("M", C.c_int64), ("ntiles", C.c_int64), ("tptr", C.c_void_p), ("tcol", C.c_void_p),
                ("tmask", C.c_void_p)]


class HostCsr(C.Structure):
    _fields_ = [("M", C.c_int64), ("N", C.c_int64), ("nnz", C.c_int64), ("rowptr", C.POINTER(C.c_int32)),
                ("col", C.POINTER(C.c_int32)), ("val", C.POINTER(C.c_double))]


class Opts(C.Structure):
    _fields_ = [("precision", C.c_int32), ("rowptr64", C.c_int32), ("plain_alloc", C.c_int32),
                ("force_global", C.c_int32), ("lds_row_max", C.c_int32), ("lds_block_budget", C.c_int32),
                ("products_per_lane", C.c_int32), ("tiles_per_lane", C.c_int32), ("timing", C.c_int32),
                ("b_lanes", C.c_int32), ("t_lanes", C.c_int32), ("rebuild_dirs", C.c_int32),
                ("persistent", C.c_int32), ("min_waves", C.c_int32), ("wide_scratch_mb", C.c_int32),
                ("wide_window", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("ms_conversion", C.c_double), ("ms_form_tileCptr", C.c_double), ("ms_tileOR", C.c_double),
                ("ms_malloc", C.c_double), ("ms_get_Ccol", C.c_double), ("ms_form_cval", C.c_double),
                ("ms_time2", C.c_double), ("ms_total", C.c_double), ("ms_numeric_bin", C.c_double * NBINS),
                ("products", C.c_int64), ("tile_products", C.c_int64), ("nnz_c", C.c_int64),
                ("tiles_b", C.c_int64), ("rows_nonempty", C.c_int64), ("sym_bin_rows", C.c_int64 * NBINS),
                ("num_bin_rows", C.c_int64 * NBINS), ("compress_rate", C.c_double), ("gflops", C.c_double),
                ("space_compressed_mask_mb", C.c_double), ("alg_bytes", C.c_double),
                ("num_bin_bytes", C.c_double * NBINS), ("num_bin_products", C.c_int64 * NBINS),
                ("compulsory_bytes", C.c_double), ("lds_launched", C.c_double), ("lds_use", C.c_double)]

    def to_dict(self) -> dict:
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if hasattr(v, "__len__") else v
        return out


_P = C.c_void_p
_sig = {
    "hsmu_ctx_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "hsmu_ctx_destroy": (None, [_P]),
    "hsmu_ctx_stream": (_P, [_P]),
    "hsmu_ctx_trim": (C.c_int, [_P]),
    "hsmu_ctx_pool_bytes": (C.c_int64, [_P]),
    "hsmu_ctx_lds_max": (C.c_int, [_P]),
    "hsmu_ctx_num_cus": (C.c_int, [_P]),
    "hsmu_opts_default": (None, [C.POINTER(Opts)]),
    "hsmu_status_string": (C.c_char_p, [C.c_int]),
    "hsmu_build_info": (C.c_char_p, []),
    "hsmu_last_error": (C.c_char_p, []),
    "hsmu_spgemm": (C.c_int, [_P, C.POINTER(Csr), C.POINTER(Csr), C.POINTER(Csr), C.POINTER(Opts),
                              C.POINTER(Stats)]),
    "hsmu_csr_free": (C.c_int, [_P, C.POINTER(Csr)]),
    "hsmu_spgemm_rows": (C.c_int, [_P, C.POINTER(Csr), C.c_int64, C.c_int64, C.POINTER(Csr), C.POINTER(Csr),
                                   C.POINTER(Opts), C.POINTER(Stats)]),
    "hsmu_csr_row_sums": (C.c_int, [_P, C.POINTER(Csr), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "hsmu_spgemm_batched": (C.c_int, [_P, C.c_int64, C.POINTER(Csr), C.POINTER(Csr), C.POINTER(Csr),
                                      C.POINTER(Opts), C.POINTER(Stats)]),
    "hsmu_batch_free": (C.c_int, [_P, C.c_int64, C.POINTER(Csr)]),
    "hsmu_spgemm_each": (C.c_int, [_P, C.c_int64, C.POINTER(Csr), C.POINTER(Csr), C.POINTER(Opts), C.c_int32,
                                   C.POINTER(C.c_double)]),
    "hsmu_plan_create": (C.c_int, [_P, C.c_int64, C.POINTER(Csr), C.POINTER(Csr), C.POINTER(Csr),
                                   C.POINTER(Opts), C.POINTER(_P)]),
    "hsmu_plan_execute": (C.c_int, [_P, C.POINTER(Stats)]),
    "hsmu_plan_destroy": (None, [_P]),
    "hsmu_shape_class": (C.c_int, [C.c_int64, C.POINTER(Csr), C.POINTER(Csr), C.c_char_p, C.c_size_t]),
    "hsmu_autotune": (C.c_int, [_P, C.c_int64, C.POINTER(Csr), C.POINTER(Csr), C.POINTER(Opts), C.c_char_p,
                                C.c_int, C.POINTER(Opts), C.POINTER(C.c_int32)]),
    "hsmu_btiles": (C.c_int, [_P, C.POINTER(Csr), C.POINTER(Tiles)]),
    "hsmu_tiles_free": (C.c_int, [_P, C.POINTER(Tiles)]),
    "hsmu_device_alloc": (C.c_int, [_P, C.c_size_t, C.POINTER(_P)]),
    "hsmu_device_free": (C.c_int, [_P, _P]),
    "hsmu_copy_h2d": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "hsmu_copy_d2h": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "hsmu_sync": (C.c_int, [_P]),
    "hsmu_copy_kernel": (C.c_int, [_P, _P, _P, C.c_size_t, C.POINTER(C.c_double)]),
    "hsmu_gen_laplace2d": (C.c_int, [C.c_int32, C.POINTER(HostCsr)]),
    "hsmu_gen_laplace3d": (C.c_int, [C.c_int32, C.POINTER(HostCsr)]),
    "hsmu_gen_banded": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_uint64, C.POINTER(HostCsr)]),
    "hsmu_gen_rmat": (C.c_int, [C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_uint64,
                                C.POINTER(HostCsr)]),
    "hsmu_gen_zipf": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_uint64, C.POINTER(HostCsr)]),
    "hsmu_gen_random": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_uint64, C.POINTER(HostCsr)]),
    "hsmu_transpose": (C.c_int, [C.POINTER(HostCsr), C.POINTER(HostCsr)]),
    "hsmu_read_mtx": (C.c_int, [C.c_char_p, C.POINTER(HostCsr), C.POINTER(C.c_int32)]),
    "hsmu_host_csr_free": (None, [C.POINTER(HostCsr)]),
}
for _name, (_res, _args) in _sig.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args


def build_info() -> str:
    """The loaded library's stamp: "src_commit=<commit>[+dirty] src=<hash>"."""
    return lib.hsmu_build_info().decode()


def source_hash() -> str:
    """The hash the Makefile stamps into libhsmu.so, recomputed from the sources in this
    tree: SHA-256 of csrc/*.hip, *.cpp, *.h and include/*.h concatenated in sorted path
    order (paths as the Makefile names them), first 16 hex digits."""
    import glob
    import hashlib
    pkg = os.path.dirname(_HERE)
    rel = []
    for pat in ("csrc/*.hip", "csrc/*.cpp", "csrc/*.h", "../include/*.h"):
        rel += [os.path.relpath(f, pkg) for f in glob.glob(os.path.join(pkg, pat))]
    h = hashlib.sha256()
    for r in sorted(rel):
        with open(os.path.join(pkg, r), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def provenance() -> dict:
    """The loaded library's stamp and whether its source hash matches this tree's."""
    stamp = build_info()
    src = stamp.split("src=")[-1]
    here = source_hash()
    return {"stamp": stamp, "tree_src": here, "src_match": src == here}


def _check(status: int, where: str) -> None:
    if status != HSMU_OK:
        raise HsmuError(status, where)


@dataclass
class HostMatrix:
    """CSR on the host: int32 row pointers (int64 allowed for results), int32 columns."""
    M: int
    N: int
    rowptr: np.ndarray
    col: np.ndarray
    val: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.col.shape[0])

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val, self.col, self.rowptr), shape=(self.M, self.N))


def _take_host(h: HostCsr) -> HostMatrix:
    M, N, nnz = int(h.M), int(h.N), int(h.nnz)
    rp = np.ctypeslib.as_array(h.rowptr, shape=(M + 1,)).copy()
    col = np.ctypeslib.as_array(h.col, shape=(max(nnz, 1),))[:nnz].copy()
    val = np.ctypeslib.as_array(h.val, shape=(max(nnz, 1),))[:nnz].copy()
    lib.hsmu_host_csr_free(C.byref(h))
    return HostMatrix(M, N, rp, col, val)


def _gen(fn, *args) -> HostMatrix:
    h = HostCsr()
    _check(fn(*args, C.byref(h)), fn.__name__)
    return _take_host(h)


def gen_laplace2d(n: int) -> HostMatrix:
    return _gen(lib.hsmu_gen_laplace2d, n)


def gen_laplace3d(n: int) -> HostMatrix:
    return _gen(lib.hsmu_gen_laplace3d, n)


def gen_banded(M: int = 62451, lo: int = 40, hi: int = 88, halfwidth: int = 400, seed: int = 1) -> HostMatrix:
    return _gen(lib.hsmu_gen_banded, M, lo, hi, halfwidth, seed)


def gen_rmat(scale: int = 20, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
             seed: int = 2) -> HostMatrix:
    return _gen(lib.hsmu_gen_rmat, scale, edge_factor, a, b, c, seed)


def gen_zipf(M: int = 500000, N: int = 100000, per_row: int = 8, s: float = 1.2, seed: int = 4) -> HostMatrix:
    return _gen(lib.hsmu_gen_zipf, M, N, per_row, s, seed)


def gen_laplace_batch(count: int = 512, seed: int = 3) -> list[HostMatrix]:
    """The batched workload: count // 2 2D 5-point Laplacians with side
    floor(exp(U(ln 16, ln 256))), then count - count // 2 3D 7-point Laplacians with side
    floor(exp(U(ln 8, ln 48))), sides drawn with numpy's default_rng(seed)."""
    rng = np.random.default_rng(seed)
    n2 = count // 2
    s2 = np.floor(np.exp(rng.uniform(np.log(16), np.log(256), n2))).astype(int)
    s3 = np.floor(np.exp(rng.uniform(np.log(8), np.log(48), count - n2))).astype(int)
    return [gen_laplace2d(int(s)) for s in s2] + [gen_laplace3d(int(s)) for s in s3]


def gen_random(M: int, N: int, max_per_row: int, seed: int) -> HostMatrix:
    return _gen(lib.hsmu_gen_random, M, N, max_per_row, seed)


def _as_host_struct(m: HostMatrix) -> tuple[HostCsr, list]:
    rp = np.ascontiguousarray(m.rowptr, dtype=np.int32)
    col = np.ascontiguousarray(m.col, dtype=np.int32)
    val = np.ascontiguousarray(m.val, dtype=np.float64)
    h = HostCsr(m.M, m.N, len(col), rp.ctypes.data_as(C.POINTER(C.c_int32)),
                col.ctypes.data_as(C.POINTER(C.c_int32)), val.ctypes.data_as(C.POINTER(C.c_double)))
    return h, [rp, col, val]


def transpose(m: HostMatrix) -> HostMatrix:
    h, keep = _as_host_struct(m)
    out = HostCsr()
    _check(lib.hsmu_transpose(C.byref(h), C.byref(out)), "hsmu_transpose")
    del keep
    return _take_host(out)


def read_mtx(path: str) -> tuple[HostMatrix, bool]:
    h = HostCsr()
    sym = C.c_int32(0)
    _check(lib.hsmu_read_mtx(path.encode(), C.byref(h), C.byref(sym)), "hsmu_read_mtx")
    return _take_host(h), bool(sym.value)


def default_opts(**kw) -> Opts:
    o = Opts()
    lib.hsmu_opts_default(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(f"unknown option {k}")
        setattr(o, k, int(v))
    return o


class DeviceMatrix:
    """A CSR matrix uploaded to device memory (int32 row pointers, FP64 or FP32 values)."""

    def __init__(self, ctx: "Context", m: HostMatrix, val32: bool = False):
        self.ctx = ctx
        self.M, self.N = m.M, m.N
        rp = np.ascontiguousarray(m.rowptr, dtype=np.int32)
        col = np.ascontiguousarray(m.col, dtype=np.int32)
        val = np.ascontiguousarray(m.val, dtype=np.float32 if val32 else np.float64)
        self.nnz = len(col)
        self._ptrs = []
        self.rowptr = ctx.upload(rp)
        self.col = ctx.upload(col)
        self.val = ctx.upload(val)
        self.csr = Csr(self.M, self.N, self.nnz, self.rowptr, self.col, self.val, 0, 1 if val32 else 0)

    def free(self) -> None:
        for p in (self.rowptr, self.col, self.val):
            if p:
                self.ctx.dfree(p)
        self.rowptr = self.col = self.val = None

    def __del__(self):
        try:
            if self.rowptr is not None and self.ctx.handle:
                self.free()
        except Exception:
            pass


class Context:
    """One device context: stream, caching allocator, pinned staging (hsmu_ctx)."""

    def __init__(self, device: int = 0):
        h = _P()
        _check(lib.hsmu_ctx_create(device, C.byref(h)), "hsmu_ctx_create")
        self.handle = h

    def close(self) -> None:
        if self.handle:
            lib.hsmu_ctx_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def trim(self) -> None:
        """Give the context's cached device memory back to the device."""
        _check(lib.hsmu_ctx_trim(self.handle), "hsmu_ctx_trim")

    @property
    def lds_max(self) -> int:
        return lib.hsmu_ctx_lds_max(self.handle)

    @property
    def num_cus(self) -> int:
        return lib.hsmu_ctx_num_cus(self.handle)

    def upload(self, arr: np.ndarray) -> int:
        arr = np.ascontiguousarray(arr)
        p = _P()
        _check(lib.hsmu_device_alloc(self.handle, max(arr.nbytes, 1), C.byref(p)), "hsmu_device_alloc")
        if arr.nbytes:
            _check(lib.hsmu_copy_h2d(self.handle, p, arr.ctypes.data, arr.nbytes), "hsmu_copy_h2d")
        return p.value

    def dfree(self, p: int) -> None:
        _check(lib.hsmu_device_free(self.handle, p), "hsmu_device_free")

    def download(self, p: int, count: int, dtype) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        if count:
            _check(lib.hsmu_copy_d2h(self.handle, out.ctypes.data, p, out.nbytes), "hsmu_copy_d2h")
        return out

    def matrix(self, m: HostMatrix, val32: bool = False) -> DeviceMatrix:
        return DeviceMatrix(self, m, val32)

    def spgemm_device(self, A: DeviceMatrix, B: DeviceMatrix, opts: Opts | None = None):
        """C = A * B on the device; returns (Csr with device pointers, stats dict).
        Release C with free_result."""
        c = Csr()
        st = Stats()
        o = opts if opts is not None else default_opts()
        _check(lib.hsmu_spgemm(self.handle, C.byref(A.csr), C.byref(B.csr), C.byref(c), C.byref(o), C.byref(st)),
               "hsmu_spgemm")
        return c, st.to_dict()

    def spgemm_rows_device(self, A: DeviceMatrix, r0: int, r1: int, B: DeviceMatrix, opts: Opts | None = None):
        """C = A[r0:r1, :] * B on the device (a row panel); returns (Csr, stats dict)."""
        c = Csr()
        st = Stats()
        o = opts if opts is not None else default_opts()
        _check(lib.hsmu_spgemm_rows(self.handle, C.byref(A.csr), r0, r1, C.byref(B.csr), C.byref(c), C.byref(o),
                                    C.byref(st)), "hsmu_spgemm_rows")
        return c, st.to_dict()

    def row_sums(self, c: Csr) -> tuple[np.ndarray, np.ndarray]:
        """(sum, sum of magnitudes) of every row of a device result."""
        s = np.zeros(max(int(c.M), 1), np.float64)
        a = np.zeros(max(int(c.M), 1), np.float64)
        _check(lib.hsmu_csr_row_sums(self.handle, C.byref(c), s.ctypes.data_as(C.POINTER(C.c_double)),
                                     a.ctypes.data_as(C.POINTER(C.c_double))), "hsmu_csr_row_sums")
        return s[:int(c.M)], a[:int(c.M)]

    def free_result(self, c: Csr) -> None:
        _check(lib.hsmu_csr_free(self.handle, C.byref(c)), "hsmu_csr_free")

    def fetch(self, c: Csr) -> HostMatrix:
        rp = self.download(c.rowptr, c.M + 1, np.int64 if c.rowptr64 else np.int32)
        col = self.download(c.col, c.nnz, np.int32)
        val = self.download(c.val, c.nnz, np.float32 if c.val32 else np.float64)
        return HostMatrix(int(c.M), int(c.N), rp, col, val)

    def fetch_rows(self, c: Csr, r0: int, r1: int) -> HostMatrix:
        """Rows [r0, r1) of a device result as a matrix of r1 - r0 rows (row pointers
        rebased to 0), for checking large results a panel at a time."""
        esz = 8 if c.rowptr64 else 4
        rp = self.download(c.rowptr + r0 * esz, r1 - r0 + 1, np.int64 if c.rowptr64 else np.int32)
        rp = rp.astype(np.int64)
        base, end = int(rp[0]), int(rp[-1])
        col = self.download(c.col + base * 4, end - base, np.int32)
        vsz = 4 if c.val32 else 8
        val = self.download(c.val + base * vsz, end - base, np.float32 if c.val32 else np.float64)
        return HostMatrix(r1 - r0, int(c.N), rp - base, col, val)

    def spgemm(self, A: HostMatrix, B: HostMatrix, **opts) -> tuple[HostMatrix, dict]:
        """Upload A and B, compute C = A * B, download C, free everything."""
        val32 = int(opts.get("precision", HSMU_FP64)) == HSMU_FP32_COMP
        dA = self.matrix(A, val32)
        dB = self.matrix(B, val32)
        try:
            c, st = self.spgemm_device(dA, dB, default_opts(**opts))
            try:
                return self.fetch(c), st
            finally:
                self.free_result(c)
        finally:
            dA.free()
            dB.free()

    # ---- batched form ----

    @staticmethod
    def _pair_arrays(As, Bs):
        n = len(As)
        if n != len(Bs) or n < 1:
            raise ValueError("As and Bs must be non-empty lists of the same length")
        a = (Csr * n)(*[x.csr for x in As])
        b = (Csr * n)(*[x.csr for x in Bs])
        return n, a, b

    def spgemm_batched_device(self, As, Bs, opts: Opts | None = None):
        """C_i = A_i * B_i for every pair, as one pipeline over all instances.  Returns
        (array of Csr views, stats dict); release with free_batch."""
        n, a, b = self._pair_arrays(As, Bs)
        cs = (Csr * n)()
        st = Stats()
        o = opts if opts is not None else default_opts()
        _check(lib.hsmu_spgemm_batched(self.handle, n, a, b, cs, C.byref(o), C.byref(st)), "hsmu_spgemm_batched")
        return cs, st.to_dict()

    def spgemm_each_ms(self, As, Bs, opts: Opts | None = None, passes: int = 1) -> float:
        """ms per pass of count sequential hsmu_spgemm calls (C_i released at once),
        timed inside the library around the whole loop (no Python per call)."""
        n, a, b = self._pair_arrays(As, Bs)
        ms = C.c_double(0)
        o = opts if opts is not None else default_opts()
        _check(lib.hsmu_spgemm_each(self.handle, n, a, b, C.byref(o), passes, C.byref(ms)), "hsmu_spgemm_each")
        return ms.value

    def free_batch(self, cs) -> None:
        _check(lib.hsmu_batch_free(self.handle, len(cs), cs), "hsmu_batch_free")

    def plan(self, As, Bs, opts: Opts | None = None) -> "Plan":
        """Record the batch (or a single product: lists of one) and capture it as a HIP
        graph; Plan.execute() replays it on the current input values."""
        return Plan(self, As, Bs, opts)

    def autotune(self, As, Bs, base: Opts | None = None, table: str | None = None, search: bool = True):
        """(opts, found_in_table) for the shape class of the batch (lists of one for a
        single product)."""
        n, a, b = self._pair_arrays(As, Bs)
        out = Opts()
        found = C.c_int32(0)
        bo = base if base is not None else default_opts()
        _check(lib.hsmu_autotune(self.handle, n, a, b, C.byref(bo), table.encode() if table else None,
                                 1 if search else 0, C.byref(out), C.byref(found)), "hsmu_autotune")
        return out, bool(found.value)

    def btiles(self, B: HostMatrix) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """B -> (tptr, tcol, tmask) in the compressed 64-column tile format."""
        dB = self.matrix(B)
        t = Tiles()
        try:
            _check(lib.hsmu_btiles(self.handle, C.byref(dB.csr), C.byref(t)), "hsmu_btiles")
            tptr = self.download(t.tptr, t.M + 1, np.int32)
            tcol = self.download(t.tcol, t.ntiles, np.int32)
            tmask = self.download(t.tmask, t.ntiles, np.uint64)
            _check(lib.hsmu_tiles_free(self.handle, C.byref(t)), "hsmu_tiles_free")
            return tptr, tcol, tmask
        finally:
            dB.free()

    def copy_kernel_ms(self, nbytes: int) -> float:
        """Time a plain device-to-device copy kernel over nbytes (bandwidth probe)."""
        a = _P()
        b = _P()
        _check(lib.hsmu_device_alloc(self.handle, nbytes, C.byref(a)), "alloc")
        _check(lib.hsmu_device_alloc(self.handle, nbytes, C.byref(b)), "alloc")
        ms = C.c_double(0)
        try:
            _check(lib.hsmu_copy_kernel(self.handle, b, a, nbytes, C.byref(ms)), "hsmu_copy_kernel")
            _check(lib.hsmu_copy_kernel(self.handle, b, a, nbytes, C.byref(ms)), "hsmu_copy_kernel")
        finally:
            lib.hsmu_device_free(self.handle, a)
            lib.hsmu_device_free(self.handle, b)
        return ms.value


def shape_class(As, Bs) -> str:
    """Autotuner shape class of a batch (lists of DeviceMatrix; lists of one for a
    single product)."""
    n, a, b = Context._pair_arrays(As, Bs)
    buf = C.create_string_buffer(128)
    _check(lib.hsmu_shape_class(n, a, b, buf, 128), "hsmu_shape_class")
    return buf.value.decode()


class Plan:
    """hsmu_plan: a recorded batch replayed as one HIP graph.  ``results`` are the Csr
    views of C_i (owned by the plan, valid until destroy)."""

    def __init__(self, ctx: Context, As, Bs, opts: Opts | None = None):
        self.ctx = ctx
        n, a, b = Context._pair_arrays(As, Bs)
        self._keep = (a, b)
        self.results = (Csr * n)()
        h = _P()
        o = opts if opts is not None else default_opts()
        _check(lib.hsmu_plan_create(ctx.handle, n, a, b, self.results, C.byref(o), C.byref(h)), "hsmu_plan_create")
        self.handle = h

    def execute(self) -> dict:
        st = Stats()
        _check(lib.hsmu_plan_execute(self.handle, C.byref(st)), "hsmu_plan_execute")
        return st.to_dict()

    def destroy(self) -> None:
        if self.handle:
            lib.hsmu_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            if self.ctx.handle:
                self.destroy()
        except Exception:
            pass
