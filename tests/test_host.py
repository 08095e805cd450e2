"""Host-side parts of libhsmu.so that need no GPU: the C ABI surface, the synthetic
generators, the Matrix Market reader and the transpose."""
import ctypes
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import hsmu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("hsmu.h", "NHC_spgemm.h")]


def declared_functions(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while", "for", "return", "sizeof")))


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(hsmu.LIB_PATH)
    names = []
    for h in HEADERS:
        names += declared_functions(h)
    assert "hsmu_spgemm" in names and "NHC_spgemm" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_check_library_exports_its_header():
    """libhsmu_check.so (the rocSPARSE cross-check) loads without a GPU and exports every
    function include/hsmu_check.h declares."""
    from hsmu import check
    names = declared_functions(os.path.join(ROOT, "include", "hsmu_check.h"))
    assert "hsmu_rocsparse_spgemm" in names and "tile_rocsparse_spgemm" in names
    missing = [n for n in names if not hasattr(check.lib, n)]
    assert not missing, missing
    assert ctypes.sizeof(check.Diff) == 56


def test_struct_layouts_match_c():
    """ctypes mirrors of the ABI structs have the C sizes (compiled with gcc here)."""
    src = r"""
    #include <stdio.h>
    #include "hsmu.h"
    #include "NHC_spgemm.h"
    int main(void) {
      printf("%zu %zu %zu %zu %zu\n", sizeof(hsmu_csr), sizeof(hsmu_tiles), sizeof(hsmu_host_csr),
             sizeof(hsmu_opts), sizeof(hsmu_stats));
      printf("%zu %zu %zu %zu\n", sizeof(NHC_CSR), sizeof(NHC_mask_matrix), sizeof(NHC_bin),
             sizeof(compressed_bin));
      return 0;
    }
    """
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "s.c")
        exe = os.path.join(d, "s")
        open(c, "w").write(src)
        subprocess.check_call(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), c, "-o", exe])
        out = subprocess.check_output([exe]).decode().split()
    sizes = list(map(int, out))
    assert sizes[:5] == [ctypes.sizeof(hsmu.Csr), ctypes.sizeof(hsmu.Tiles), ctypes.sizeof(hsmu.HostCsr),
                         ctypes.sizeof(hsmu.Opts), ctypes.sizeof(hsmu.Stats)]
    # the reference's structs on x86-64 (inc/NHC_spgemm.h:66-130): same field order
    assert sizes[5:] == [128, 88, 32, 136]


def test_laplace_sizes():
    for n in (1, 5, 32):
        a = hsmu.gen_laplace2d(n)
        assert a.M == n * n and a.nnz == 5 * n * n - 4 * n
        b = hsmu.gen_laplace3d(n)
        assert b.M == n ** 3 and b.nnz == 7 * n ** 3 - 6 * n * n


def rows_sorted_distinct(m):
    for i in range(m.M):
        c = m.col[m.rowptr[i]:m.rowptr[i + 1]]
        if np.any(np.diff(c) <= 0):
            return False
    return True


def test_banded_spec():
    a = hsmu.gen_banded(5000, 40, 88, 400, seed=1)
    b = hsmu.gen_banded(5000, 40, 88, 400, seed=1)
    assert np.array_equal(a.col, b.col) and np.array_equal(a.val, b.val)
    c = hsmu.gen_banded(5000, 40, 88, 400, seed=2)
    assert not np.array_equal(a.col[:1000], c.col[:1000])
    lens = np.diff(a.rowptr)
    assert lens.min() >= 40 and lens.max() <= 88
    rows = np.repeat(np.arange(a.M), lens)
    assert np.all(np.abs(a.col - rows) <= 400)
    assert rows_sorted_distinct(a)
    assert abs(a.val.mean()) < 0.02 and abs(a.val.std() - 1.0) < 0.02


def test_banded_full_size():
    a = hsmu.gen_banded()
    assert a.M == 62451 and abs(a.nnz / a.M - 64) < 0.5


def test_rmat_small():
    a = hsmu.gen_rmat(10, 16, seed=2)
    assert a.M == a.N == 1024
    assert rows_sorted_distinct(a)
    assert a.nnz < 16 * 1024
    assert a.val.min() >= 0 and a.val.max() < 1
    lens = np.diff(a.rowptr)
    assert lens.max() > 8 * lens.mean()  # skewed rows
    b = hsmu.gen_rmat(10, 16, seed=2)
    assert np.array_equal(a.col, b.col)


def test_zipf_small():
    a = hsmu.gen_zipf(2000, 500, 8, 1.2, seed=4)
    assert a.nnz == 16000 and rows_sorted_distinct(a)
    counts = np.bincount(a.col, minlength=500)
    assert counts[0] == counts.max() and counts[0] > 0.6 * a.M


def test_transpose_roundtrip():
    a = hsmu.gen_random(37, 53, 9, seed=7)
    t = hsmu.transpose(a)
    assert t.M == 53 and t.N == 37 and rows_sorted_distinct(t)
    tt = hsmu.transpose(t)
    assert np.array_equal(tt.rowptr, a.rowptr) and np.array_equal(tt.col, a.col)
    assert np.array_equal(tt.val, a.val)
    assert np.allclose(t.to_scipy().toarray(), a.to_scipy().toarray().T)


def test_read_mtx_semantics(tmp_path):
    p = tmp_path / "g.mtx"
    p.write_text("%%MatrixMarket matrix coordinate integer general\n% c\n3 4 5\n"
                 "1 4 7\n1 2 3\n3 1 -2\n1 2 5\n2 3 1\n")
    m, sym = hsmu.read_mtx(str(p))
    assert not sym and m.M == 3 and m.N == 4
    # duplicates kept, rows stable-sorted by column
    assert m.rowptr.tolist() == [0, 3, 4, 5]
    assert m.col.tolist() == [1, 1, 3, 2, 0]
    assert m.val.tolist() == [3.0, 5.0, 7.0, 1.0, -2.0]
    q = tmp_path / "p.mtx"
    q.write_text("%%MatrixMarket matrix coordinate pattern symmetric\n3 3 3\n1 1\n3 1\n2 2\n")
    m, sym = hsmu.read_mtx(str(q))
    assert sym and m.nnz == 4 and m.val.tolist() == [1.0] * 4
    assert m.rowptr.tolist() == [0, 2, 3, 4] and m.col.tolist() == [0, 2, 1, 0]


def test_read_mtx_rejects_bad_index(tmp_path):
    p = tmp_path / "b.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n")
    with pytest.raises(hsmu.HsmuError):
        hsmu.read_mtx(str(p))


def test_read_mtx_rejects_malformed_headers(tmp_path):
    """A symmetric file must be square (its mirrored entries land in rows up to N), and an
    entry count larger than the file can hold is a corrupt file: both are errors, never an
    out-of-bounds write or an allocation sized by the header."""
    p = tmp_path / "ns.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real symmetric\n2 40000 2\n1 1 1.0\n2 39999 2.0\n")
    with pytest.raises(hsmu.HsmuError) as e:
        hsmu.read_mtx(str(p))
    assert e.value.status == hsmu.HSMU_ERR_INVALID
    q = tmp_path / "big.mtx"
    q.write_text("%%MatrixMarket matrix coordinate real general\n4 4 9000000000000000\n1 1 1.0\n")
    with pytest.raises(hsmu.HsmuError) as e:
        hsmu.read_mtx(str(q))
    assert e.value.status == hsmu.HSMU_ERR_INVALID
    r = tmp_path / "sq.mtx"
    r.write_text("%%MatrixMarket matrix coordinate real symmetric\n3 3 2\n1 1 1.0\n3 2 2.0\n")
    m, sym = hsmu.read_mtx(str(r))
    assert sym and m.nnz == 3 and m.rowptr.tolist() == [0, 1, 2, 3] and m.col.tolist() == [0, 2, 1]
