"""The two large workloads at their full sizes, entry by entry on sampled rows.

R-MAT scale 20 (C = A*A, nnz(C) 9.7e9, about 116 GB) and Zipf A*A^T (A 500,000 x 100,000,
C about 2.7 TB, so computed in row panels of 10,000 rows as the streamed workload does)
cannot be brought to the host whole.  Their device results are checked on every row by
the row-sum identity rowsum(C) = A * rowsum(B), and entry by entry on sampled rows, read
slice by slice from the device, against the oracle's rows (structure bit-exact, values
within 1e-12 of each entry's sum of magnitudes).
"""
import numpy as np
import pytest

import hsmu
import oracle

pytestmark = pytest.mark.gpu
TOL = 1e-12


def note(msg):
    print(f"[gpu-case] {msg}", flush=True)


def row_of(ctx, c, rp, r):
    """Row r of a device result as a one-row HostMatrix (only that row is copied)."""
    a, b = int(rp[r]), int(rp[r + 1])
    col = ctx.download(c.col + 4 * a, b - a, np.int32)
    val = ctx.download(c.val + 8 * a, b - a, np.float64)
    return hsmu.HostMatrix(1, int(c.N), np.array([0, b - a], np.int64), col, val)


def check_rows(ctx, c, A, B, rows, r0=0):
    """Rows of the device result c (rows r0.. of A*B) against the oracle, and every row sum
    against A * rowsum(B)."""
    rp = ctx.download(c.rowptr, int(c.M) + 1, np.int64 if c.rowptr64 else np.int32)
    assert int(rp[-1]) == int(c.nnz)
    rb = np.bincount(np.repeat(np.arange(B.M), np.diff(B.rowptr)), weights=B.val, minlength=B.M)
    a0, a1 = int(A.rowptr[r0]), int(A.rowptr[r0 + int(c.M)])
    sub = hsmu.HostMatrix(int(c.M), A.N, (A.rowptr[r0:r0 + int(c.M) + 1] - a0).astype(np.int64),
                          A.col[a0:a1], A.val[a0:a1])
    sums, asums = ctx.row_sums(c)
    worst = np.abs(sums - sub.to_scipy() @ rb) / (1 + asums)
    assert worst.max(initial=0.0) <= TOL
    for r in rows:
        ref = oracle.spgemm_rows(A, B, r0 + r, r0 + r + 1)
        oracle.compare(row_of(ctx, c, rp, r), ref, tol=TOL)
    return rp


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
