"""Parity of the HIP path (through the C ABI) against the CPU oracle.

Structure (row pointers, columns) must be bit-exact; FP64 values must satisfy
|c - c_ref| <= 1e-12 * sum_j |a_ij * b_jk| (the FP64 tolerance this library promises;
LDS/L2 atomics make the summation order nondeterministic, as in the reference).
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import hsmu
import oracle

pytestmark = pytest.mark.gpu
TOL = 1e-12
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def note(msg):
    print(f"[gpu-case] {msg}", flush=True)


def check(ctx, A, B, name, tol=TOL, **opts):
    note(f"{name}: A {A.M}x{A.N} nnz {A.nnz}, B {B.M}x{B.N} nnz {B.nnz}, opts {opts}")
    Cm, st = ctx.spgemm(A, B, **opts)
    ref = oracle.spgemm(A, B)
    summary = oracle.compare(Cm, ref, tol=tol)
    assert st["nnz_c"] == Cm.nnz == int(ref.rowptr[-1])
    assert st["products"] == oracle.products(A, B)
    return Cm, st, summary


def test_golden_test_mtx(ctx):
    import json
    A, _ = hsmu.read_mtx(os.path.join(GOLDEN, "test.mtx"))
    g = json.load(open(os.path.join(GOLDEN, "test_mtx_AxA.json")))
    Cm, st, _ = check(ctx, A, A, "test.mtx")
    assert Cm.rowptr.tolist() == g["rowptr"] and Cm.col.tolist() == g["col"]
    assert Cm.val.tolist() == g["val"]  # small integers: exact
    assert st["products"] == 19


def test_laplace2d_128(ctx):
    A = hsmu.gen_laplace2d(128)
    Cm, st, _ = check(ctx, A, A, "laplace2d 128")
    assert Cm.nnz == 210436 and st["products"] == 405000
    ref = oracle.spgemm(A, A)
    assert np.array_equal(Cm.val, ref.val)  # integer-valued: exact


@pytest.mark.parametrize("n", [2, 9, 24, 40])
def test_laplace3d(ctx, n):
    A = hsmu.gen_laplace3d(n)
    Cm, _, _ = check(ctx, A, A, f"laplace3d {n}")
    assert np.array_equal(Cm.val, oracle.spgemm(A, A).val)


@pytest.mark.parametrize("M,lo,hi,hw", [(3000, 40, 88, 400), (20000, 5, 30, 2000), (9000, 100, 200, 150)])
def test_banded_small(ctx, M, lo, hi, hw):
    A = hsmu.gen_banded(M, lo, hi, hw, seed=M)
    check(ctx, A, A, f"banded {M}")


def test_banded_full(ctx):
    A = hsmu.gen_banded()
    Cm, st, s = check(ctx, A, A, "banded 62451 (full config)")
    note(f"banded full: nnz {Cm.nnz} products {st['products']} max_rel {s['max_rel_err']:.2e}")


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("shape", [(50, 70, 90), (300, 200, 5000), (1000, 3000, 70000), (7, 1, 300)])
def test_random_rectangular(ctx, shape, seed):
    m, k, n = shape
    A = hsmu.gen_random(m, k, 12, seed)
    B = hsmu.gen_random(k, n, 40, seed + 100)
    check(ctx, A, B, f"random {shape} seed {seed}")


def test_small_n_path(ctx):
    """B.N < 4096: the reference's small (dense bitmap) path; same kernels here."""
    A = hsmu.gen_random(2000, 1500, 30, 11)
    B = hsmu.gen_random(1500, 1000, 60, 12)
    check(ctx, A, B, "small N")


def test_empty_rows_and_columns(ctx):
    A = hsmu.gen_random(500, 400, 3, 5)  # many empty rows
    B = hsmu.gen_random(400, 9000, 2, 6)
    check(ctx, A, B, "sparse with empty rows")


@pytest.mark.parametrize("case", ["A_empty", "B_empty", "zero_rows"])
def test_empty_inputs(ctx, case):
    if case == "A_empty":
        A = hsmu.HostMatrix(10, 20, np.zeros(11, np.int32), np.zeros(0, np.int32), np.zeros(0))
        B = hsmu.gen_random(20, 30, 5, 1)
    elif case == "B_empty":
        A = hsmu.gen_random(10, 20, 5, 1)
        B = hsmu.HostMatrix(20, 30, np.zeros(21, np.int32), np.zeros(0, np.int32), np.zeros(0))
    else:
        A = hsmu.HostMatrix(0, 20, np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0))
        B = hsmu.gen_random(20, 30, 5, 1)
    note(f"empty input {case}")
    Cm, st = ctx.spgemm(A, B)
    assert Cm.nnz == 0 and Cm.M == A.M and np.all(Cm.rowptr == 0)


def test_duplicates_in_B_are_summed(ctx):
    B = hsmu.gen_random(300, 2000, 20, 3)
    rp, col, val = [0], [], []
    for i in range(B.M):
        c = B.col[B.rowptr[i]:B.rowptr[i + 1]]
        v = B.val[B.rowptr[i]:B.rowptr[i + 1]]
        c2 = np.sort(np.concatenate([c, c[::3]]), kind="stable")
        v2 = np.concatenate([v, v[::3]])[np.argsort(np.concatenate([c, c[::3]]), kind="stable")]
        col.extend(c2.tolist())
        val.extend(v2.tolist())
        rp.append(len(col))
    Bd = hsmu.HostMatrix(B.M, B.N, np.array(rp, np.int32), np.array(col, np.int32), np.array(val))
    A = hsmu.gen_random(200, 300, 15, 4)
    check(ctx, A, Bd, "duplicates in B")


def test_cancellation_keeps_structure(ctx):
    """Structural nonzeros stay even when the value cancels to exactly zero."""
    A = hsmu.HostMatrix(1, 2, np.array([0, 2], np.int32), np.array([0, 1], np.int32), np.array([1.0, -1.0]))
    B = hsmu.HostMatrix(2, 3, np.array([0, 2, 4], np.int32), np.array([0, 2, 0, 2], np.int32),
                        np.array([2.0, 5.0, 2.0, 5.0]))
    Cm, _, _ = check(ctx, A, B, "cancellation")
    assert Cm.col.tolist() == [0, 2] and Cm.val.tolist() == [0.0, 0.0]


@pytest.mark.parametrize("opts", [dict(force_global=1), dict(lds_row_max=256), dict(lds_block_budget=4096),
                                  dict(products_per_lane=1), dict(products_per_lane=512, tiles_per_lane=256),
                                  dict(b_lanes=1, t_lanes=1), dict(b_lanes=64, t_lanes=64), dict(rebuild_dirs=1),
                                  dict(rebuild_dirs=1, products_per_lane=4), dict(min_waves=1), dict(min_waves=64)])
def test_bins_and_fallbacks(ctx, opts):
    A = hsmu.gen_banded(4000, 10, 120, 600, seed=7)
    check(ctx, A, A, "banded under forced bins", **opts)
    B = hsmu.gen_random(800, 20000, 50, 8)
    A2 = hsmu.gen_random(600, 800, 40, 9)
    check(ctx, A2, B, "random under forced bins", **opts)


def long_rows_pair(seed):
    """B (600 x 20000) whose rows 0..4 hold 3000 entries each (the rest 0..20), and A
    (800 x 600) whose rows hit those long rows often: exercises the whole-wave walks of
    long rows of B in every kernel."""
    rng = np.random.default_rng(seed)
    K, N = 600, 20000
    rows_b = []
    for j in range(K):
        n = 3000 if j < 5 else int(rng.integers(0, 21))
        rows_b.append(np.sort(rng.choice(N, size=n, replace=False)).astype(np.int32))
    rp_b = np.zeros(K + 1, np.int32)
    rp_b[1:] = np.cumsum([len(r) for r in rows_b])
    B = hsmu.HostMatrix(K, N, rp_b, np.concatenate(rows_b), rng.uniform(-1, 1, int(rp_b[-1])))
    rows_a = []
    for i in range(800):
        cols = set(int(c) for c in rng.choice(K, size=int(rng.integers(0, 40)), replace=False))
        cols |= set(int(c) for c in range(5) if rng.random() < 0.5)
        rows_a.append(np.array(sorted(cols), np.int32))
    rp_a = np.zeros(801, np.int32)
    rp_a[1:] = np.cumsum([len(r) for r in rows_a])
    A = hsmu.HostMatrix(800, K, rp_a, np.concatenate(rows_a), rng.uniform(-1, 1, int(rp_a[-1])))
    return A, B


# force_global: every row in the global bin, whose directories fit LDS (k_symbolic_wide);
# lds_row_max=256 with rebuild_dirs: directories wider than 32 tiles in the global-memory
# workspace (k_symbolic_global)
@pytest.mark.parametrize("opts", [dict(), dict(force_global=1), dict(force_global=1, rebuild_dirs=1),
                                  dict(lds_row_max=256, rebuild_dirs=1), dict(lds_row_max=4096),
                                  dict(products_per_lane=512)])
def test_long_rows_of_B(ctx, opts):
    A, B = long_rows_pair(11)
    check(ctx, A, B, f"long rows of B {opts}", **opts)


def split_rows_pair(seed, heavy_rows, na_lo, na_hi, K, N, b_lo, b_hi):
    """A (heavy_rows + 200 rows x K) whose first heavy_rows rows hold na_lo..na_hi entries
    (the rest 0..30), B (K x N) with b_lo..b_hi entries per row: the first rows of C are
    wide and split into several work items whose sums meet in C."""
    rng = np.random.default_rng(seed)
    rows_b = [np.sort(rng.choice(N, size=int(rng.integers(b_lo, b_hi + 1)), replace=False)).astype(np.int32)
              for _ in range(K)]
    rp_b = np.zeros(K + 1, np.int32)
    rp_b[1:] = np.cumsum([len(r) for r in rows_b])
    B = hsmu.HostMatrix(K, N, rp_b, np.concatenate(rows_b), rng.uniform(-1, 1, int(rp_b[-1])))
    M = heavy_rows + 200
    rows_a = [np.sort(rng.choice(K, size=int(rng.integers(na_lo, na_hi + 1)) if i < heavy_rows
                                 else int(rng.integers(0, 31)), replace=False)).astype(np.int32) for i in range(M)]
    rp_a = np.zeros(M + 1, np.int32)
    rp_a[1:] = np.cumsum([len(r) for r in rows_a])
    A = hsmu.HostMatrix(M, K, rp_a, np.concatenate(rows_a), rng.uniform(-1, 1, int(rp_a[-1])))
    return A, B


@pytest.mark.parametrize("shape", [(4, 2000, 3000, 3000, 300000, 20, 80), (1, 20000, 20000, 40000, 1 << 20, 5, 15),
                                   (6, 1500, 2500, 5000, 60000, 100, 400), (3, 3000, 3000, 6000, 1 << 21, 2, 2),
                                   (3, 2500, 2500, 6000, 1 << 23, 2, 2)])
@pytest.mark.parametrize("opts", [dict(), dict(force_global=1), dict(wide_window=4096), dict(wide_window=8192)])
def test_split_wide_rows(ctx, shape, opts):
    A, B = split_rows_pair(5, *shape)
    check(ctx, A, B, f"split wide rows {shape} {opts}", **opts)


def test_wide_rows_repeated_entries_and_empty_B_rows(ctx):
    """Wide rows (split into several work items) whose row of A names one row of B several
    times, and rows of B that are empty: repeated entries add up, empty rows add nothing."""
    A, B = split_rows_pair(8, 2, 1500, 2500, 3000, 200000, 20, 60)
    rp_b = np.asarray(B.rowptr, np.int64)
    keep = np.ones(B.nnz, bool)
    for j in range(0, B.M, 7):  # empty every 7th row of B
        keep[rp_b[j]:rp_b[j + 1]] = False
    lens = np.diff(rp_b).copy()
    lens[::7] = 0
    rp2 = np.zeros(B.M + 1, np.int32)
    rp2[1:] = np.cumsum(lens)
    B2 = hsmu.HostMatrix(B.M, B.N, rp2, B.col[keep], B.val[keep])
    rows = []
    vals = []
    rng = np.random.default_rng(9)
    for i in range(A.M):
        c = A.col[A.rowptr[i]:A.rowptr[i + 1]]
        v = A.val[A.rowptr[i]:A.rowptr[i + 1]]
        if i < 2:  # repeat a third of the entries of the wide rows
            extra = rng.choice(len(c), size=len(c) // 3, replace=False)
            order = np.argsort(np.concatenate([c, c[extra]]), kind="stable")
            c = np.concatenate([c, c[extra]])[order]
            v = np.concatenate([v, v[extra] * 0.5])[order]
        rows.append(c)
        vals.append(v)
    rp_a = np.zeros(A.M + 1, np.int32)
    rp_a[1:] = np.cumsum([len(r) for r in rows])
    A2 = hsmu.HostMatrix(A.M, A.N, rp_a, np.concatenate(rows).astype(np.int32), np.concatenate(vals))
    for opts in (dict(), dict(force_global=1)):
        check(ctx, A2, B2, f"wide rows, repeated entries, empty rows of B {opts}", **opts)


def test_row_panel_errors(ctx):
    A = hsmu.gen_banded(300, 5, 20, 50, seed=2)
    dA = ctx.matrix(A)
    try:
        for r0, r1 in [(-1, 10), (20, 10), (0, A.M + 1)]:
            note(f"panel error rows [{r0}, {r1})")
            with pytest.raises(hsmu.HsmuError) as ei:
                ctx.spgemm_rows_device(dA, r0, r1, dA)
            assert ei.value.status == hsmu.HSMU_ERR_INVALID
        c, _ = ctx.spgemm_rows_device(dA, 0, A.M, dA)  # the context stays usable
        oracle.compare(ctx.fetch(c), oracle.spgemm(A, A), tol=TOL)
        ctx.free_result(c)
    finally:
        dA.free()


@pytest.mark.parametrize("case", ["banded", "zipf_aat"])
def test_row_panels_and_row_sums(ctx, case):
    """hsmu_spgemm_rows over consecutive panels reproduces the oracle's rows; the device
    row sums equal A * rowsum(B) (the closed form the streamed workloads check with)."""
    if case == "banded":
        A = hsmu.gen_banded(5000, 20, 60, 300, seed=6)
        B = A
    else:
        A = hsmu.gen_zipf(6000, 3000, 8, 1.2, seed=4)
        B = hsmu.transpose(A)
    dA, dB = ctx.matrix(A), ctx.matrix(B)
    rb = np.bincount(np.repeat(np.arange(B.M), np.diff(B.rowptr)), weights=B.val, minlength=B.M)
    try:
        for r0, r1 in [(0, 1000), (1000, 2500), (2500, 2500), (2500, A.M)]:
            note(f"panel {case} rows [{r0}, {r1})")
            c, st = ctx.spgemm_rows_device(dA, r0, r1, dB)
            try:
                got = ctx.fetch(c)
                ref = oracle.spgemm_rows(A, B, r0, r1)
                oracle.compare(got, ref, tol=TOL)
                sums, asums = ctx.row_sums(c)
                sub = hsmu.HostMatrix(r1 - r0, A.N, A.rowptr[r0:r1 + 1] - A.rowptr[r0],
                                      A.col[A.rowptr[r0]:A.rowptr[r1]], A.val[A.rowptr[r0]:A.rowptr[r1]])
                want = sub.to_scipy() @ rb if r1 > r0 else np.zeros(0)
                assert np.allclose(sums, want, rtol=0, atol=1e-12 * (1 + np.abs(asums).max(initial=0)))
            finally:
                ctx.free_result(c)
    finally:
        dA.free()
        dB.free()


def test_rowptr64(ctx):
    A = hsmu.gen_banded(5000, 20, 60, 300, seed=2)
    Cm, _, _ = check(ctx, A, A, "int64 row pointers", rowptr64=1)
    assert Cm.rowptr.dtype == np.int64


@pytest.mark.parametrize("scale", [10, 13])
def test_rmat(ctx, scale):
    A = hsmu.gen_rmat(scale, 16, seed=2)
    check(ctx, A, A, f"rmat scale {scale}")


def test_rmat_scale16_sampled_rows(ctx):
    """R-MAT scale 16 (nnz(C) 163.8 M, widest row 37,461 entries) computed whole, so its
    wide rows take the split work-item kernel: the 16 widest rows and 48 seeded random rows
    are compared with the oracle's, and every row sum with A * rowsum(B)."""
    A = hsmu.gen_rmat(16, 16, seed=2)
    note(f"rmat scale 16 whole product: A {A.M}x{A.N} nnz {A.nnz}")
    dA = ctx.matrix(A)
    try:
        c, st = ctx.spgemm_device(dA, dA)
        try:
            sums, asums = ctx.row_sums(c)
            rb = np.bincount(np.repeat(np.arange(A.M), np.diff(A.rowptr)), weights=A.val, minlength=A.M)
            assert np.all(np.abs(sums - A.to_scipy() @ rb) <= TOL * (1 + asums))
            got = ctx.fetch(c)
        finally:
            ctx.free_result(c)
    finally:
        dA.free()
    lens = np.diff(got.rowptr)
    assert st["nnz_c"] == int(got.rowptr[-1]) == 163791280
    rows = np.concatenate([np.argsort(-lens, kind="stable")[:16],
                           np.random.default_rng(16).integers(0, A.M, size=48)])
    for r in (int(x) for x in rows):
        oracle.compare(got, oracle.spgemm_rows(A, A, r, r + 1), r0=r, r1=r + 1, tol=TOL)


def test_zipf_aat(ctx):
    A = hsmu.gen_zipf(4000, 2000, 8, 1.2, seed=4)
    check(ctx, A, hsmu.transpose(A), "zipf A*A^T")


def test_btiles_match_oracle(ctx):
    for name, B in [("banded", hsmu.gen_banded(3000, 40, 88, 400, seed=1)),
                    ("random", hsmu.gen_random(700, 50000, 90, 3)), ("laplace", hsmu.gen_laplace2d(40))]:
        note(f"btiles {name}")
        tptr, tcol, tmask = ctx.btiles(B)
        rp, rc, rm = oracle.btiles(B)
        assert np.array_equal(tptr, rp) and np.array_equal(tcol, rc) and np.array_equal(tmask, rm)


def test_errors_fail_loudly(ctx):
    A = hsmu.gen_random(50, 60, 5, 1)
    B = hsmu.gen_random(60, 70, 5, 2)
    note("error: shape mismatch")
    with pytest.raises(hsmu.HsmuError) as e:
        ctx.spgemm(A, hsmu.gen_random(61, 70, 5, 2))
    assert e.value.status == hsmu.HSMU_ERR_INVALID
    note("error: unsorted B")
    Bu = hsmu.HostMatrix(B.M, B.N, B.rowptr, B.col.copy(), B.val)
    r = int(np.argmax(np.diff(B.rowptr) >= 2))
    s = Bu.rowptr[r]
    Bu.col[s], Bu.col[s + 1] = Bu.col[s + 1], Bu.col[s]
    with pytest.raises(hsmu.HsmuError) as e:
        ctx.spgemm(A, Bu)
    assert e.value.status == hsmu.HSMU_ERR_UNSORTED
    note("error: column out of range")
    Ab = hsmu.HostMatrix(A.M, A.N, A.rowptr, A.col.copy(), A.val)
    Ab.col[0] = 60
    with pytest.raises(hsmu.HsmuError) as e:
        ctx.spgemm(Ab, B)
    assert e.value.status == hsmu.HSMU_ERR_INDEX
    note("error: wide_window not 0, 4096 or 8192")
    with pytest.raises(hsmu.HsmuError) as e:
        ctx.spgemm(A, B, wide_window=6000)
    assert e.value.status == hsmu.HSMU_ERR_INVALID
    note("after errors the context still works")
    check(ctx, A, B, "after errors")


def test_repeated_calls_reuse_pool(ctx):
    A = hsmu.gen_banded(6000, 30, 60, 500, seed=3)
    dA = ctx.matrix(A)
    ref = oracle.spgemm(A, A)
    held = None
    for it in range(4):
        note(f"repeat {it}")
        c, st = ctx.spgemm_device(dA, dA)
        oracle.compare(ctx.fetch(c), ref)
        ctx.free_result(c)
        pool = hsmu.lib.hsmu_ctx_pool_bytes(ctx.handle)
        if held is not None:
            assert pool == held  # steady state: no new device allocations
        held = pool
    dA.free()


class NHC_CSR(ctypes.Structure):
    _fields_ = [("rowPtr", ctypes.c_void_p), ("col", ctypes.c_void_p), ("val", ctypes.c_void_p),
                ("d_ptr", ctypes.c_void_p), ("d_col", ctypes.c_void_p), ("d_val", ctypes.c_void_p),
                ("d_tem_ptr", ctypes.c_void_p), ("M", ctypes.c_int), ("N", ctypes.c_int), ("nnz", ctypes.c_int),
                ("nnz_max", ctypes.c_int), ("isSymmetric", ctypes.c_int), ("matrix_name", ctypes.c_char_p),
                ("d_tile_ptr", ctypes.c_void_p), ("d_tile_col", ctypes.c_void_p),
                ("d_tem_tile_ptr", ctypes.c_void_p), ("tile_ptr", ctypes.c_void_p),
                ("d_mask_num", ctypes.c_void_p)]


def test_nhc_spgemm_dropin(ctx):
    """The reference-signature entry point through the C ABI."""
    lib = hsmu.lib
    A = hsmu.gen_banded(4000, 30, 60, 900, seed=5)
    keep = []

    def nhc(m):
        x = NHC_CSR()
        x.M, x.N, x.nnz = m.M, m.N, m.nnz
        for f, arr in (("rowPtr", m.rowptr.astype(np.int32)), ("col", m.col.astype(np.int32)),
                       ("val", m.val.astype(np.float64))):
            arr = np.ascontiguousarray(arr)
            keep.append(arr)
            setattr(x, f, arr.ctypes.data)
        lib.HtD_value_col_ptr(ctypes.byref(x))
        return x

    a, b = nhc(A), nhc(A)
    c = NHC_CSR()
    c.M, c.N = A.M, A.N
    cb = (ctypes.c_char * 136)()
    bn = (ctypes.c_char * 32)()
    mk = (ctypes.c_char * 88)()
    note("NHC_spgemm banded 4000")
    lib.NHC_spgemm(cb, bn, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), mk)
    assert lib.NHC_last_status() == 0
    rp = np.ctypeslib.as_array(ctypes.cast(c.rowPtr, ctypes.POINTER(ctypes.c_int)), shape=(A.M + 1,)).copy()
    col = ctx.download(c.d_col, c.nnz, np.int32)
    val = ctx.download(c.d_val, c.nnz, np.float64)
    oracle.compare(hsmu.HostMatrix(A.M, A.N, rp, col, val), oracle.spgemm(A, A))
    lib.free_C_matrix(ctypes.byref(c))
    for x in (a, b):  # device arrays only: host arrays belong to numpy here
        hsmu.lib.hsmu_device_free(ctx.handle, x.d_ptr)
        hsmu.lib.hsmu_device_free(ctx.handle, x.d_col)
        hsmu.lib.hsmu_device_free(ctx.handle, x.d_val)


def test_driver_prints_reference_lines(tmp_path):
    exe = os.path.join(os.path.dirname(hsmu.LIB_PATH), "hsmu_test")
    note("driver on test.mtx")
    out = subprocess.run([exe, os.path.join(GOLDEN, "test.mtx"), "--csv-dir", str(tmp_path)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    text = out.stdout
    for line in ["the matrix is: --------------", "success open file", "input matrix A: ( 5, 5 ) nnz = 9",
                 "nums_Intermediate_product = 19", "cnnz = 11", "the total run time of NHC_spgemm is"]:
        assert line in text, (line, text)
    note("driver on gen:laplace2d:128")
    out = subprocess.run([exe, "gen:laplace2d:128"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    for line in ["the run time for Initialize_compressed_mask64_matrix is", "space_compressed_mask =",
                 "the run time for compressed_Form_tileCptr is", "the run time for h_compress_tileOR_for_Cptr is",
                 "the run time for cudaMalloc col and val is", "the run time for h_get_Ccol_from_Ctile is",
                 "the run time for h_form_cval is", "nums_Intermediate_product = 405000", "cnnz = 210436",
                 "compress_rate =", "the gflops is"]:
        assert line in out.stdout, (line, out.stdout)
    assert os.path.exists(tmp_path / "NHC_MI355X_result.csv")


def as_fp32(m):
    """The same matrix with values rounded to FP32 (kept in float64 for the oracle)."""
    return hsmu.HostMatrix(m.M, m.N, m.rowptr, m.col, m.val.astype(np.float32).astype(np.float64))


def check_fp32c(ctx, A, B, name, **opts):
    """Compensated FP32: structure bit-exact; |c - c_ref| <= 2^-23 |c_ref| + 1e-10 sum|a*b|,
    c_ref the FP64 oracle on the same FP32-rounded inputs."""
    A32, B32 = as_fp32(A), as_fp32(B)
    note(f"fp32c {name}: A {A.M}x{A.N} nnz {A.nnz}, opts {opts}")
    Cm, st = ctx.spgemm(A32, B32, precision=hsmu.HSMU_FP32_COMP, **opts)
    assert Cm.val.dtype == np.float32
    ref = oracle.spgemm(A32, B32)
    assert np.array_equal(np.asarray(Cm.rowptr, np.int64), ref.rowptr) and np.array_equal(Cm.col, ref.col)
    got = Cm.val.astype(np.float64)
    err = np.abs(got - ref.val)
    bound = 2.0 ** -23 * np.abs(ref.val) + 1e-10 * ref.absval
    assert np.all(err <= bound), f"max excess {np.max(err - bound):.3e}"
    with np.errstate(divide="ignore", invalid="ignore"):
        rel = np.where(np.abs(ref.val) > 0, err / np.abs(ref.val), 0.0)
    note(f"fp32c {name}: max rel err {rel.max(initial=0):.2e}, median {np.median(rel) if len(rel) else 0:.2e}")
    return rel


@pytest.mark.parametrize("case", ["laplace2d", "banded", "random", "rmat", "zipf_aat"])
def test_fp32_compensated(ctx, case):
    if case == "laplace2d":
        A, B = hsmu.gen_laplace2d(64), None
    elif case == "banded":
        A, B = hsmu.gen_banded(6000, 40, 88, 400, seed=3), None
    elif case == "random":
        A, B = hsmu.gen_random(800, 900, 30, 5), hsmu.gen_random(900, 30000, 60, 6)
    elif case == "rmat":
        A, B = hsmu.gen_rmat(12, 16, seed=2), None
    else:
        A = hsmu.gen_zipf(3000, 1500, 8, 1.2, seed=4)
        B = hsmu.transpose(A)
    rel = check_fp32c(ctx, A, A if B is None else B, case)
    assert np.median(rel) < 1e-6


def test_fp32_compensated_duplicate_columns(ctx):
    """Rows of B that repeat a column (summed, as in the reference): two lanes of one
    instruction may hit one slot.  Every product of the first case lands in one entry of
    C, two per row of B."""
    n = 3000
    rng = np.random.default_rng(12)
    a = rng.uniform(-1, 1, n).astype(np.float32).astype(np.float64)
    b = (rng.uniform(-1, 1, 2 * n) * 10.0 ** rng.integers(-3, 4, 2 * n)).astype(np.float32).astype(np.float64)
    A = hsmu.HostMatrix(1, n, np.array([0, n], np.int32), np.arange(n, dtype=np.int32), a)
    B = hsmu.HostMatrix(n, 1, np.arange(0, 2 * n + 1, 2, dtype=np.int32), np.zeros(2 * n, np.int32), b)
    rel = check_fp32c(ctx, A, B, "one entry, duplicated columns")
    assert rel[0] < 1e-6
    Bd = hsmu.gen_random(400, 3000, 30, 13)
    col = np.repeat(Bd.col, 2)
    val = np.repeat(Bd.val, 2) * 0.5
    Bd = hsmu.HostMatrix(Bd.M, Bd.N, (2 * np.asarray(Bd.rowptr)).astype(np.int32), col.astype(np.int32),
                         val.astype(np.float32).astype(np.float64))
    A2 = hsmu.gen_random(500, 400, 40, 14)
    check_fp32c(ctx, A2, Bd, "random with duplicated columns")


@pytest.mark.parametrize("shape", [(4, 2000, 3000, 3000, 300000, 20, 80), (6, 1500, 2500, 5000, 60000, 100, 400),
                                   (3, 3000, 3000, 6000, 1 << 21, 2, 2), "rmat12"])
@pytest.mark.parametrize("opts", [dict(), dict(force_global=1), dict(wide_scratch_mb=1), dict(wide_window=8192)])
def test_fp32_compensated_split_wide_rows(ctx, shape, opts):
    """Compensated FP32 on wide rows split into several work items: each item's window
    sums meet as (s, c) pairs in a scratch buffer (TwoSum by compare-and-swap) and are
    rounded into C at the end.  With the scratch capped at 1 MiB, rows whose scratch would
    not fit take one workgroup per row (k_numeric_wcompact) instead.  wide_window=8192
    forces the large window (items of up to 1,024 entries), whose split rows are a subset
    of those the scratch was sized for."""
    if shape == "rmat12":
        A = hsmu.gen_rmat(12, 16, seed=2)
        B = A
    else:
        A, B = split_rows_pair(5, *shape)
    check_fp32c(ctx, A, B, f"split wide rows {shape}", **opts)


@pytest.mark.parametrize("opts", [dict(force_global=1), dict(lds_row_max=512), dict(rebuild_dirs=1)])
def test_fp32_compensated_fallbacks(ctx, opts):
    A = hsmu.gen_banded(3000, 20, 90, 700, seed=9)
    check_fp32c(ctx, A, A, "banded fallback", **opts)


def test_fp32_compensated_beats_plain_fp32_sum(ctx):
    """Long sums of mixed-sign terms: the compensated result keeps FP64-level accuracy
    where a plain FP32 sum of the same FP32 products would lose about 3 digits."""
    n = 4000
    rng = np.random.default_rng(11)
    a = rng.uniform(-1, 1, n).astype(np.float32).astype(np.float64)
    b = (rng.uniform(-1, 1, n) * 10.0 ** rng.integers(-3, 4, n)).astype(np.float32).astype(np.float64)
    A = hsmu.HostMatrix(1, n, np.array([0, n], np.int32), np.arange(n, dtype=np.int32), a)
    B = hsmu.HostMatrix(n, 1, np.arange(n + 1, dtype=np.int32), np.zeros(n, np.int32), b)
    rel = check_fp32c(ctx, A, B, "long dot product")
    exact = float(np.dot(a, b))
    naive = np.float32(0)
    for x, y in zip(a.astype(np.float32), b.astype(np.float32)):
        naive = np.float32(naive + x * y)
    assert rel[0] < 1e-6 and rel[0] * 10 < abs(float(naive) - exact) / abs(exact)
