"""Cross-check of the HIP path against rocSPARSE's SpGEMM on the same device inputs
(the counterpart of the reference's cuSPARSE check, src/cusparse_in_tile.cu:91-272).

rocSPARSE is an independent implementation: row pointers and columns must agree exactly;
values within 1e-12 * (|A|*|B|), the sum of magnitudes computed by rocSPARSE itself from
|A| and |B|.  The comparison routine (hsmu_csr_compare) is checked on planted mistakes.
"""
import os
import subprocess

import numpy as np
import pytest

import hsmu
import oracle

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "hsmu-spgemm_amd", "hsmu_test")


def note(msg):
    print(f"[gpu-case] {msg}", flush=True)


@pytest.fixture(scope="module")
def chk():
    from hsmu import check
    return check


def absolute(m):
    return hsmu.HostMatrix(m.M, m.N, m.rowptr, m.col, np.abs(m.val))


def cross_check(ctx, chk, A, B, name, opts=None):
    note(f"rocsparse cross-check {name}: A {A.M}x{A.N} nnz {A.nnz}, B {B.M}x{B.N} nnz {B.nnz}")
    dA, dB = ctx.matrix(A), ctx.matrix(B)
    aA, aB = ctx.matrix(absolute(A)), ctx.matrix(absolute(B))
    ours, st = ctx.spgemm_device(dA, dB, opts)
    theirs, ms, _ = chk.rocsparse_spgemm(ctx, dA, dB)
    scale, _, _ = chk.rocsparse_spgemm(ctx, aA, aB)
    try:
        d = chk.compare(ours, theirs, scale, 1e-12)
        assert d["rowptr_mismatch"] == 0, d
        assert d["col_mismatch"] == 0, d
        assert d["val_mismatch"] == 0, d
        assert int(ours.nnz) == int(theirs.nnz) == st["nnz_c"]
        return d, ms
    finally:
        ctx.free_result(ours)
        chk.free(theirs)
        chk.free(scale)
        for m in (dA, dB, aA, aB):
            m.free()


def test_laplace2d(ctx, chk):
    A = hsmu.gen_laplace2d(128)
    cross_check(ctx, chk, A, A, "laplace2d 128")


def test_banded_full(ctx, chk):
    A = hsmu.gen_banded()
    d, ms = cross_check(ctx, chk, A, A, "banded 62451")
    note(f"banded: rocSPARSE stages {ms:.2f} ms, max rel err {d['max_rel_err']:.2e}")


@pytest.mark.parametrize("shape", [(300, 200, 5000), (1000, 3000, 70000), (2000, 1500, 1000)])
def test_random_rectangular(ctx, chk, shape):
    m, k, n = shape
    A = hsmu.gen_random(m, k, 12, 21)
    B = hsmu.gen_random(k, n, 40, 22)
    cross_check(ctx, chk, A, B, f"random {shape}")


def test_zipf_aat(ctx, chk):
    A = hsmu.gen_zipf(20000, 8000, 8, 1.2, seed=4)
    cross_check(ctx, chk, A, hsmu.transpose(A), "zipf A*A^T 20000x8000")


def test_rmat_wide_rows(ctx, chk):
    A = hsmu.gen_rmat(13, 16, 0.57, 0.19, 0.19, seed=2)
    cross_check(ctx, chk, A, A, "rmat 13", hsmu.default_opts(rowptr64=0))


def test_empty_A(ctx, chk):
    A = hsmu.HostMatrix(10, 20, np.zeros(11, np.int32), np.zeros(0, np.int32), np.zeros(0))
    B = hsmu.gen_random(20, 30, 5, 1)
    cross_check(ctx, chk, A, B, "empty A")


def test_fp32_against_rocsparse_f32(ctx, chk):
    """Compensated FP32 against rocSPARSE's plain FP32 and the FP64 oracle: same structure;
    the compensated result is at least as close to the exact sum as rocSPARSE's."""
    A = hsmu.gen_banded(6000, 20, 60, 300, seed=9)
    A = hsmu.HostMatrix(A.M, A.N, A.rowptr, A.col, A.val.astype(np.float32).astype(np.float64))
    note("fp32: banded 6000 compensated vs rocSPARSE f32")
    dA = ctx.matrix(A, val32=True)
    ours, _ = ctx.spgemm_device(dA, dA, hsmu.default_opts(precision=hsmu.HSMU_FP32_COMP))
    theirs, _, _ = chk.rocsparse_spgemm(ctx, dA, dA)
    try:
        d = chk.compare(ours, theirs, None, 1e-4)
        assert d["rowptr_mismatch"] == 0 and d["col_mismatch"] == 0, d
        ref = oracle.spgemm(A, A)
        mine = ctx.fetch(ours).val.astype(np.float64)
        rs = ctx.fetch(theirs).val.astype(np.float64)
        e_mine = np.abs(mine - ref.val) / ref.absval
        e_rs = np.abs(rs - ref.val) / ref.absval
        note(f"fp32: max err / sum|ab|: compensated {e_mine.max():.2e}, rocSPARSE f32 {e_rs.max():.2e}")
        assert e_mine.max() <= e_rs.max() + 1e-12
        assert e_mine.max() <= 2.0 ** -23
    finally:
        ctx.free_result(ours)
        chk.free(theirs)
        dA.free()


def test_compare_finds_planted_mistakes(ctx, chk):
    """hsmu_csr_compare reports a changed value, a changed column and a changed row
    pointer (mistakes planted in host copies of a correct result)."""
    A = hsmu.gen_random(200, 300, 10, 5)
    B = hsmu.gen_random(300, 500, 10, 6)
    Cm, _ = ctx.spgemm(A, B)
    good = ctx.matrix(Cm)
    k = Cm.nnz // 2
    v = Cm.val.copy()
    v[k] *= 1 + 1e-9
    bad_val = ctx.matrix(hsmu.HostMatrix(Cm.M, Cm.N, Cm.rowptr, Cm.col, v))
    c = Cm.col.copy()
    row = int(np.searchsorted(Cm.rowptr, k, side="right") - 1)
    last = int(Cm.rowptr[row + 1]) - 1
    c[last] = Cm.N - 1 if c[last] != Cm.N - 1 else Cm.N - 2
    bad_col = ctx.matrix(hsmu.HostMatrix(Cm.M, Cm.N, Cm.rowptr, c, Cm.val))
    rp = Cm.rowptr.astype(np.int32).copy()
    i = next(j for j in range(1, Cm.M) if rp[j - 1] < rp[j] < rp[j + 1])
    rp[i] += 1
    bad_rp = ctx.matrix(hsmu.HostMatrix(Cm.M, Cm.N, rp, Cm.col, Cm.val))
    try:
        assert chk.compare(good.csr, good.csr)["val_mismatch"] == 0
        d = chk.compare(good.csr, bad_val.csr, None, 1e-12)
        assert d["val_mismatch"] == 1 and d["first_entry"] == k and d["col_mismatch"] == 0
        d = chk.compare(good.csr, bad_col.csr)
        assert d["col_mismatch"] == 1 and d["first_entry"] == last
        d = chk.compare(good.csr, bad_rp.csr)
        assert d["rowptr_mismatch"] == 1 and d["first_row"] == i
    finally:
        for m in (good, bad_val, bad_col, bad_rp):
            m.free()


@pytest.mark.parametrize("spec", ["gen:laplace2d:64", "gen:banded:5000", "gen:zipf:3000:2000"])
def test_driver_check_flag(spec):
    """hsmu_test --check: the reference's CHECK_RESULT report, every part right."""
    note(f"driver --check {spec}")
    args = [DRIVER, spec, "--check"] + (["--aat"] if "zipf" in spec else [])
    out = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    for line in ("compute the Cptr is right!", "compute the Cval is right!", "compute the Ccol is right!",
                 "the result is all right!"):
        assert line in out.stdout, out.stdout
