"""The large workloads at their full sizes, against the oracle.

R-MAT scale 20 (C = A*A, nnz(C) 9.7e9, about 116 GB) and Zipf A*A^T (A 500,000 x 100,000,
C about 2.7 TB, so computed in row panels of 10,000 rows as the streamed workload does)
cannot be brought to the host whole.  Their device results are checked on every row by
the row-sum identity rowsum(C) = A * rowsum(B), and entry by entry on sampled rows, read
slice by slice from the device, against the oracle's rows: structure bit-exact; FP64
values within 1e-12 of each entry's sum of magnitudes; compensated FP32 values within
2^-23 |c| + 1e-10 sum|a*b| of the FP64 oracle on the same FP32-rounded inputs.  The
512-instance batch is compared whole, every instance, exactly (integer values).
"""
import os

import numpy as np
import pytest

import hsmu
import oracle

pytestmark = pytest.mark.gpu
TOL = 1e-12
TABLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "hsmu-spgemm_amd", "tuning",
                     "gfx950.tsv")


def note(msg):
    print(f"[gpu-case] {msg}", flush=True)


def as_fp32(m):
    """The same matrix with values rounded to FP32 (kept in float64 for the oracle)."""
    return hsmu.HostMatrix(m.M, m.N, m.rowptr, m.col, m.val.astype(np.float32).astype(np.float64))


def row_of(ctx, c, rp, r):
    """Row r of a device result as a one-row HostMatrix (only that row is copied)."""
    a, b = int(rp[r]), int(rp[r + 1])
    vsz, vt = (4, np.float32) if c.val32 else (8, np.float64)
    col = ctx.download(c.col + 4 * a, b - a, np.int32)
    val = ctx.download(c.val + vsz * a, b - a, vt).astype(np.float64)
    return hsmu.HostMatrix(1, int(c.N), np.array([0, b - a], np.int64), col, val)


def fp32c_row_errors(got, ref):
    """Compensated FP32 row against the FP64 oracle row: structure bit-exact, every value
    within 2^-23 |c_ref| + 1e-10 sum|a*b|; returns the relative errors of the nonzeros."""
    assert np.array_equal(np.asarray(got.rowptr, np.int64), ref.rowptr) and np.array_equal(got.col, ref.col)
    err = np.abs(got.val - ref.val)
    bound = 2.0 ** -23 * np.abs(ref.val) + 1e-10 * ref.absval
    assert np.all(err <= bound), f"max excess {np.max(err - bound):.3e}"
    nz = ref.val != 0
    return err[nz] / np.abs(ref.val[nz])


def check_rows(ctx, c, A, B, rows, r0=0):
    """Rows of the device result c (rows r0.. of A*B) against the oracle, and every row sum
    against A * rowsum(B).  For a compensated FP32 result (A, B FP32-rounded), the row sums
    within 2^-22 of the row's sum of magnitudes plus 1e-10 of |A| * rowsum|B|; returns the
    per-row (max, median) relative errors of the sampled rows."""
    rp = ctx.download(c.rowptr, int(c.M) + 1, np.int64 if c.rowptr64 else np.int32)
    assert int(rp[-1]) == int(c.nnz)
    rows_of_b = np.repeat(np.arange(B.M), np.diff(B.rowptr))
    rb = np.bincount(rows_of_b, weights=B.val, minlength=B.M)
    a0, a1 = int(A.rowptr[r0]), int(A.rowptr[r0 + int(c.M)])
    sub = hsmu.HostMatrix(int(c.M), A.N, (A.rowptr[r0:r0 + int(c.M) + 1] - a0).astype(np.int64),
                          A.col[a0:a1], A.val[a0:a1])
    sums, asums = ctx.row_sums(c)
    want = sub.to_scipy() @ rb
    per_row = []
    if c.val32:
        rabs = abs(sub.to_scipy()) @ np.bincount(rows_of_b, weights=np.abs(B.val), minlength=B.M)
        assert np.all(np.abs(sums - want) <= 2.0 ** -22 * asums + 1e-10 * rabs + 1e-30)
    else:
        worst = np.abs(sums - want) / (1 + asums)
        assert worst.max(initial=0.0) <= TOL
    for r in rows:
        ref = oracle.spgemm_rows(A, B, r0 + r, r0 + r + 1)
        got = row_of(ctx, c, rp, r)
        if c.val32:
            rel = fp32c_row_errors(got, ref)
            per_row.append((float(rel.max(initial=0.0)), float(np.median(rel)) if rel.size else 0.0))
        else:
            oracle.compare(got, ref, tol=TOL)
    return rp, per_row


def report_rows(name, per_row):
    mx = np.array([p[0] for p in per_row])
    md = np.array([p[1] for p in per_row])
    note(f"{name}: per-row max rel err over {len(per_row)} rows: median {np.median(mx):.2e}, "
         f"90th pct {np.quantile(mx, 0.9):.2e}, max {mx.max():.2e}; per-row median rel err: "
         f"median {np.median(md):.2e}, max {md.max():.2e}")
    assert mx.max() < 1e-6


def test_rmat_scale20(ctx):
    """R-MAT scale 20, edge factor 16, seed 2, int64 row pointers: the 16 widest rows of
    C (the split work-item path) and 32 seeded random rows."""
    A = hsmu.gen_rmat(20, 16, 0.57, 0.19, 0.19, seed=2)
    note(f"rmat scale 20 whole product: A {A.M}x{A.N} nnz {A.nnz}")
    dA = ctx.matrix(A)
    try:
        c, st = ctx.spgemm_device(dA, dA, hsmu.default_opts(rowptr64=1))
        try:
            assert st["nnz_c"] == int(c.nnz) > 2 ** 31
            rp = ctx.download(c.rowptr, int(c.M) + 1, np.int64)
            lens = np.diff(rp)
            rows = np.concatenate([np.argsort(-lens, kind="stable")[:16],
                                   np.random.default_rng(20).integers(0, A.M, size=32)])
            check_rows(ctx, c, A, A, [int(r) for r in rows])
        finally:
            ctx.free_result(c)
    finally:
        dA.free()
        ctx.trim()


def test_rmat_scale20_fp32c(ctx):
    """R-MAT scale 20, compensated FP32 (values rounded to FP32): the 16 widest rows of C
    (the split work items, whose window sums meet as (s, c) pairs in the scratch) and 32
    seeded random rows against the FP64 oracle on the same rounded inputs, and every row
    sum; prints the per-row error distribution."""
    A = as_fp32(hsmu.gen_rmat(20, 16, 0.57, 0.19, 0.19, seed=2))
    note(f"rmat scale 20 fp32c whole product: A {A.M}x{A.N} nnz {A.nnz}")
    dA = ctx.matrix(A, val32=True)
    try:
        c, st = ctx.spgemm_device(dA, dA, hsmu.default_opts(rowptr64=1, precision=hsmu.HSMU_FP32_COMP))
        try:
            assert c.val32 and st["nnz_c"] == int(c.nnz) > 2 ** 31
            rp = ctx.download(c.rowptr, int(c.M) + 1, np.int64)
            lens = np.diff(rp)
            rows = np.concatenate([np.argsort(-lens, kind="stable")[:16],
                                   np.random.default_rng(20).integers(0, A.M, size=32)])
            _, per_row = check_rows(ctx, c, A, A, [int(r) for r in rows])
            report_rows("rmat scale 20 fp32c", per_row)
        finally:
            ctx.free_result(c)
    finally:
        dA.free()
        ctx.trim()


def test_zipf_aat_panel_fp32c(ctx):
    """Zipf A*A^T at full shape, compensated FP32: the panel of rows [250000, 260000), its 8
    widest rows and 24 seeded random rows against the FP64 oracle on the FP32-rounded
    inputs, every row sum, and the per-row error distribution."""
    A = as_fp32(hsmu.gen_zipf(500000, 100000, 8, 1.2, seed=4))
    B = hsmu.transpose(A)
    r0 = 250000
    note(f"zipf A*A^T fp32c panel rows [{r0}, {r0 + 10000})")
    dA, dB = ctx.matrix(A, val32=True), ctx.matrix(B, val32=True)
    try:
        c, st = ctx.spgemm_rows_device(dA, r0, r0 + 10000, dB,
                                       hsmu.default_opts(rowptr64=1, precision=hsmu.HSMU_FP32_COMP))
        try:
            rp = ctx.download(c.rowptr, int(c.M) + 1, np.int64)
            lens = np.diff(rp)
            rows = np.concatenate([np.argsort(-lens, kind="stable")[:8],
                                   np.random.default_rng(r0).integers(0, 10000, size=24)])
            _, per_row = check_rows(ctx, c, A, B, [int(r) for r in rows], r0=r0)
            report_rows("zipf panel fp32c", per_row)
        finally:
            ctx.free_result(c)
    finally:
        dA.free()
        dB.free()
        ctx.trim()


@pytest.fixture(scope="module")
def laplace_batch():
    return hsmu.gen_laplace_batch(512, seed=3)


def check_batch(ctx, mats, results, name):
    """Every instance of a batch result against the oracle, exactly (integer values)."""
    for i, m in enumerate(mats):
        got = ctx.fetch(results[i])
        ref = oracle.spgemm(m, m)
        oracle.compare(got, ref, tol=0.0)
    note(f"{name}: {len(mats)} instances identical to the oracle")


def test_batched_512_against_oracle(ctx, laplace_batch):
    """The batched workload as configured: 512 uneven Laplacians (seed 3) in one
    hsmu_spgemm_batched call, every C_i against the oracle."""
    mats = laplace_batch
    note(f"batched 512: rows {sum(m.M for m in mats)}, nnz {sum(m.nnz for m in mats)}")
    dms = [ctx.matrix(m) for m in mats]
    try:
        opts, _ = ctx.autotune(dms, dms, base=hsmu.default_opts(), table=TABLE, search=False)
        cs, st = ctx.spgemm_batched_device(dms, dms, opts)
        try:
            assert st["products"] == sum(oracle.products(m, m) for m in mats)
            check_batch(ctx, mats, cs, "batched call")
        finally:
            ctx.free_batch(cs)
    finally:
        for d in dms:
            d.free()
        ctx.trim()


@pytest.mark.parametrize("persistent", [0, 1])
def test_batched_512_plan_replay_against_oracle(ctx, laplace_batch, persistent):
    """The same batch as a plan (the autotuner table's options, and the persistent form),
    replayed twice as a HIP graph after the values were scaled on the device: every C_i
    of the last replay against the oracle of the scaled inputs."""
    mats = laplace_batch
    note(f"batched 512 plan replay, persistent={persistent}")
    dms = [ctx.matrix(m) for m in mats]
    try:
        opts, _ = ctx.autotune(dms, dms, base=hsmu.default_opts(), table=TABLE, search=False)
        opts.persistent = persistent
        plan = ctx.plan(dms, dms, opts)
        try:
            scaled = []
            for m, d in zip(mats, dms):
                s = hsmu.HostMatrix(m.M, m.N, m.rowptr, m.col, 2.0 * m.val)
                hsmu.lib.hsmu_copy_h2d(ctx.handle, d.val, s.val.ctypes.data, s.val.nbytes)
                scaled.append(s)
            plan.execute()
            plan.execute()
            check_batch(ctx, scaled, plan.results, f"plan replay persistent={persistent}")
        finally:
            plan.destroy()
    finally:
        for d in dms:
            d.free()
        ctx.trim()


@pytest.mark.parametrize("r0", [0, 250000, 490000])
def test_zipf_aat_panels(ctx, r0):
    """Zipf A*A^T, A 500,000 x 100,000, 8 columns per row Zipf(1.2), seed 4: a panel of
    10,000 rows of A, its 8 widest rows and 24 seeded random rows."""
    A = hsmu.gen_zipf(500000, 100000, 8, 1.2, seed=4)
    B = hsmu.transpose(A)
    note(f"zipf A*A^T panel rows [{r0}, {r0 + 10000})")
    dA, dB = ctx.matrix(A), ctx.matrix(B)
    try:
        c, st = ctx.spgemm_rows_device(dA, r0, r0 + 10000, dB, hsmu.default_opts(rowptr64=1))
        try:
            rp = ctx.download(c.rowptr, int(c.M) + 1, np.int64)
            lens = np.diff(rp)
            rows = np.concatenate([np.argsort(-lens, kind="stable")[:8],
                                   np.random.default_rng(r0).integers(0, 10000, size=24)])
            check_rows(ctx, c, A, B, [int(r) for r in rows], r0=r0)
        finally:
            ctx.free_result(c)
    finally:
        dA.free()
        dB.free()
        ctx.trim()


def test_zipf_panels_reuse_pool(ctx):
    """Consecutive 2,000-row panels of the Zipf A*A^T: C of each is several GB and differs
    from the last by a fraction of a percent.  Blocks of 1 GiB or more come in
    eighth-octave sizes, so after the first panel the context's pool takes no further
    block of C's size (smaller workspaces may still grow by a few MB)."""
    A = hsmu.gen_zipf(500000, 100000, 8, 1.2, seed=4)
    B = hsmu.transpose(A)
    dA, dB = ctx.matrix(A), ctx.matrix(B)

    def bucket(b):
        g = 1 << (b.bit_length() - 4) if b >= 1 << 30 else 2 << 20
        return -(-b // g) * g

    try:
        ctx.trim()
        nnz, held = [], []
        for p in range(4):
            note(f"zipf pool reuse: panel rows [{2000 * p}, {2000 * (p + 1)})")
            c, st = ctx.spgemm_rows_device(dA, 2000 * p, 2000 * (p + 1), dB, hsmu.default_opts(rowptr64=1))
            nnz.append(int(c.nnz))
            ctx.free_result(c)
            held.append(hsmu.lib.hsmu_ctx_pool_bytes(ctx.handle))
        note(f"nnz {nnz}, pool bytes {held}")
        assert 4 * min(nnz) >= 1 << 30 and len(set(nnz)) > 1
        # every panel's value and column arrays round to the first panel's block sizes
        assert all(bucket(8 * n) == bucket(8 * nnz[0]) and bucket(4 * n) == bucket(4 * nnz[0]) for n in nnz)
        assert held[-1] - held[0] < 1 << 30
    finally:
        dA.free()
        dB.free()
        ctx.trim()
