"""hsmu_spgemm_noalloc (through the C ABI) against hsmu_spgemm and the oracle.

The allocation-free entry point takes its whole workspace from one file-scope device
array and writes C into arrays the caller passes.  Two calls in a row (the second with a
smaller input, so it runs over the first one's leftovers in the array) must give what
hsmu_spgemm gives: structure bit-exact, values within the FP64 tolerance of the oracle
(1e-12 sum|a*b|) or, compensated FP32, 2^-23 |c| + 1e-10 sum|a*b|.  The context's pool
must stay empty: the entry point allocates nothing.
"""
import ctypes as C

import numpy as np
import pytest

import hsmu
import oracle

pytestmark = pytest.mark.gpu

hsmu.lib.hsmu_spgemm_noalloc.restype = C.c_int
hsmu.lib.hsmu_spgemm_noalloc.argtypes = [C.c_void_p, C.POINTER(hsmu.Csr), C.POINTER(hsmu.Csr), C.POINTER(hsmu.Csr),
                                         C.c_int64, C.POINTER(hsmu.Opts), C.POINTER(hsmu.Stats)]


def note(msg):
    print(f"[gpu-case] {msg}", flush=True)


class Outputs:
    """Device arrays for C, allocated by the test (the caller), not by the entry point."""

    def __init__(self, ctx, M, cap, rp64=False, fp32=False):
        self.ctx = ctx
        self.rp = ctx.upload(np.zeros(M + 1, np.int64 if rp64 else np.int32))
        self.col = ctx.upload(np.zeros(max(cap, 1), np.int32))
        self.val = ctx.upload(np.zeros(max(cap, 1), np.float32 if fp32 else np.float64))
        self.cap = cap

    def csr(self):
        return hsmu.Csr(0, 0, 0, self.rp, self.col, self.val, 0, 0)

    def free(self):
        for p in (self.rp, self.col, self.val):
            self.ctx.dfree(p)


def noalloc(ctx, dA, dB, out, **opts):
    c = out.csr()
    st = hsmu.Stats()
    o = hsmu.default_opts(**opts)
    rc = hsmu.lib.hsmu_spgemm_noalloc(ctx.handle, C.byref(dA.csr), C.byref(dB.csr), C.byref(c), out.cap, C.byref(o),
                                      C.byref(st))
    return rc, c, st.to_dict()


def check_same(ctx, A, B, got_c, fp32=False, **opts):
    """got_c (a device result of the allocation-free call) against hsmu_spgemm's result
    on the same inputs and against the oracle."""
    got = ctx.fetch(got_c)
    ref_c, _ = ctx.spgemm(A, B, **opts)
    assert np.array_equal(np.asarray(got.rowptr, np.int64), np.asarray(ref_c.rowptr, np.int64))
    assert np.array_equal(got.col, ref_c.col)
    ref = oracle.spgemm(A, B)
    if fp32:
        assert np.array_equal(np.asarray(got.rowptr, np.int64), ref.rowptr) and np.array_equal(got.col, ref.col)
        err = np.abs(got.val.astype(np.float64) - ref.val)
        assert np.all(err <= 2.0 ** -23 * np.abs(ref.val) + 1e-10 * ref.absval)
    else:
        oracle.compare(got, ref, tol=1e-12)
        oracle.compare(ref_c, ref, tol=1e-12)


@pytest.fixture(scope="module")
def fresh():
    """A context of its own: its pool must stay empty across the allocation-free calls."""
    c = hsmu.Context(0)
    yield c
    c.close()


def test_noalloc_two_calls_second_smaller(fresh):
    ctx = fresh
    big = hsmu.gen_banded(62451, 40, 88, 400, seed=1)
    small = hsmu.gen_laplace2d(64)
    dbig, dsmall = ctx.matrix(big), ctx.matrix(small)
    obig = Outputs(ctx, big.M, 80_000_000)
    osmall = Outputs(ctx, small.M, 60_000)
    try:
        pool0 = hsmu.lib.hsmu_ctx_pool_bytes(ctx.handle)
        note("noalloc: banded 62451 rows, then laplace2d 64")
        rc, c1, st1 = noalloc(ctx, dbig, dbig, obig)
        assert rc == hsmu.HSMU_OK, hsmu.lib.hsmu_last_error()
        rc, c2, st2 = noalloc(ctx, dsmall, dsmall, osmall)
        assert rc == hsmu.HSMU_OK, hsmu.lib.hsmu_last_error()
        assert hsmu.lib.hsmu_ctx_pool_bytes(ctx.handle) == pool0 == 0
        assert c1.nnz == st1["nnz_c"] == 79652451 and c2.nnz == st2["nnz_c"]
        check_same(ctx, big, big, c1)
        check_same(ctx, small, small, c2)
    finally:
        for x in (obig, osmall, dbig, dsmall):
            x.free()
        ctx.trim()


def test_noalloc_fp32c_rowptr64_two_calls(fresh):
    ctx = fresh
    big = hsmu.gen_banded(20000, 40, 88, 400, seed=5)
    small = hsmu.gen_random(700, 900, 20, 3)
    smallB = hsmu.gen_random(900, 5000, 30, 4)
    as32 = lambda m: hsmu.HostMatrix(m.M, m.N, m.rowptr, m.col, m.val.astype(np.float32).astype(np.float64))
    big, small, smallB = as32(big), as32(small), as32(smallB)
    dbig, dsmall, dsmallB = ctx.matrix(big, True), ctx.matrix(small, True), ctx.matrix(smallB, True)
    obig = Outputs(ctx, big.M, 30_000_000, rp64=True, fp32=True)
    osmall = Outputs(ctx, small.M, 400_000, rp64=True, fp32=True)
    try:
        note("noalloc fp32c rowptr64: banded 20000 rows, then random 700x900 * 900x5000")
        kw = dict(precision=hsmu.HSMU_FP32_COMP, rowptr64=1)
        rc, c1, _ = noalloc(ctx, dbig, dbig, obig, **kw)
        assert rc == hsmu.HSMU_OK, hsmu.lib.hsmu_last_error()
        rc, c2, _ = noalloc(ctx, dsmall, dsmallB, osmall, **kw)
        assert rc == hsmu.HSMU_OK, hsmu.lib.hsmu_last_error()
        assert c1.rowptr64 and c1.val32 and c2.rowptr64 and c2.val32
        check_same(ctx, big, big, c1, fp32=True, **kw)
        check_same(ctx, small, smallB, c2, fp32=True, **kw)
    finally:
        for x in (obig, osmall, dbig, dsmall, dsmallB):
            x.free()
        ctx.trim()


def test_noalloc_limits(fresh):
    """C larger than the caller's arrays: HSMU_ERR_NOMEM with the nnz(C) needed; a
    workspace beyond the scratch array: HSMU_ERR_NOMEM naming the limit; the next call
    works."""
    ctx = fresh
    A = hsmu.gen_laplace2d(32)
    dA = ctx.matrix(A)
    ref_nnz = int(oracle.spgemm(A, A).rowptr[-1])
    tight = Outputs(ctx, A.M, ref_nnz - 1)
    fits = Outputs(ctx, A.M, ref_nnz)
    big = hsmu.gen_banded(500000, 40, 88, 400, seed=2)  # 32M entries: B's tile workspace alone exceeds it
    dbig = ctx.matrix(big)
    obig = Outputs(ctx, big.M, 16)
    try:
        note("noalloc limits: capacity, scratch size")
        rc, c, _ = noalloc(ctx, dA, dA, tight)
        assert rc == hsmu.HSMU_ERR_NOMEM and c.nnz == ref_nnz
        assert b"capacity" in hsmu.lib.hsmu_last_error()
        rc, c, _ = noalloc(ctx, dbig, dbig, obig)
        assert rc == hsmu.HSMU_ERR_NOMEM and b"scratch" in hsmu.lib.hsmu_last_error()
        rc, c, _ = noalloc(ctx, dA, dA, fits)
        assert rc == hsmu.HSMU_OK and c.nnz == ref_nnz
        assert hsmu.lib.hsmu_ctx_pool_bytes(ctx.handle) == 0
        check_same(ctx, A, A, c)
    finally:
        for x in (tight, fits, obig, dA, dbig):
            x.free()
        ctx.trim()


def wide_pair(seed, heavy_rows, na_lo, na_hi, K, N, b_lo, b_hi):
    """A (heavy_rows + 100 rows x K) whose first rows hold na_lo..na_hi entries, B (K x N)
    with b_lo..b_hi entries per row: the first rows of C are wider than LDS and are split
    into several work items of the windowed accumulator."""
    rng = np.random.default_rng(seed)
    rows_b = [np.sort(rng.choice(N, size=int(rng.integers(b_lo, b_hi + 1)), replace=False)).astype(np.int32)
              for _ in range(K)]
    rp_b = np.zeros(K + 1, np.int32)
    rp_b[1:] = np.cumsum([len(r) for r in rows_b])
    B = hsmu.HostMatrix(K, N, rp_b, np.concatenate(rows_b), rng.uniform(-1, 1, int(rp_b[-1])))
    M = heavy_rows + 100
    rows_a = [np.sort(rng.choice(K, size=int(rng.integers(na_lo, na_hi + 1)) if i < heavy_rows
                                 else int(rng.integers(0, 31)), replace=False)).astype(np.int32) for i in range(M)]
    rp_a = np.zeros(M + 1, np.int32)
    rp_a[1:] = np.cumsum([len(r) for r in rows_a])
    A = hsmu.HostMatrix(M, K, rp_a, np.concatenate(rows_a), rng.uniform(-1, 1, int(rp_a[-1])))
    return A, B


@pytest.mark.parametrize("fp32", [False, True])
def test_noalloc_wide_rows_both_windows(fresh, fp32):
    """Wide rows split into work items, through the allocation-free call: sparse ones
    (about two columns per tile: narrow workgroups), the same through the 8,192-slot
    window (wide_window=8192), then denser ones (wide workgroups), each against
    hsmu_spgemm and the oracle."""
    ctx = fresh
    as32 = lambda m: hsmu.HostMatrix(m.M, m.N, m.rowptr, m.col, m.val.astype(np.float32).astype(np.float64))
    A1, B1 = wide_pair(11, 3, 3000, 6000, 6000, 1 << 21, 2, 2)
    A2, B2 = wide_pair(12, 4, 2000, 3000, 3000, 300000, 20, 80)
    if fp32:
        A1, B1, A2, B2 = as32(A1), as32(B1), as32(A2), as32(B2)
    kw = dict(precision=hsmu.HSMU_FP32_COMP) if fp32 else {}
    d = [ctx.matrix(m, fp32) for m in (A1, B1, A2, B2)]
    n1 = int(oracle.spgemm(A1, B1).rowptr[-1])
    n2 = int(oracle.spgemm(A2, B2).rowptr[-1])
    o1 = Outputs(ctx, A1.M, n1, fp32=fp32)
    o1b = Outputs(ctx, A1.M, n1, fp32=fp32)
    o2 = Outputs(ctx, A2.M, n2, fp32=fp32)
    try:
        note(f"noalloc wide rows ({'fp32c' if fp32 else 'fp64'}): sparse nnz(C) {n1} (default, then 8,192-slot "
             f"window), then denser nnz(C) {n2}")
        rc, c1, st1 = noalloc(ctx, d[0], d[1], o1, **kw)
        assert rc == hsmu.HSMU_OK, hsmu.lib.hsmu_last_error()
        rc, c1b, st1b = noalloc(ctx, d[0], d[1], o1b, wide_window=8192, **kw)
        assert rc == hsmu.HSMU_OK, hsmu.lib.hsmu_last_error()
        rc, c2, st2 = noalloc(ctx, d[2], d[3], o2, **kw)
        assert rc == hsmu.HSMU_OK, hsmu.lib.hsmu_last_error()
        assert hsmu.lib.hsmu_ctx_pool_bytes(ctx.handle) == 0
        for st in (st1, st1b, st2):
            assert st["num_bin_rows"][8] >= 1  # rows took the windowed path
        check_same(ctx, A1, B1, c1, fp32=fp32, **kw)
        check_same(ctx, A1, B1, c1b, fp32=fp32, **kw)
        check_same(ctx, A2, B2, c2, fp32=fp32, **kw)
    finally:
        for x in [o1, o1b, o2] + d:
            x.free()
        ctx.trim()
