"""Pins the CPU oracle before it is trusted as the checker of the GPU path.

Every oracle function is checked against something it does not compute itself:
  oracle.spgemm     - the reference's only matrix (nsparse test.mtx) with A*A derived by
                      hand (tests/golden/test_mtx_AxA.json);
                    - closed forms of the 2D/3D Laplacian squares, built from grid
                      geometry (A^2 = 16 I - 8 Adj + Adj^2), not from a matrix product;
                    - brute-force dense products (numpy) on random rectangular inputs,
                      including repeated columns and explicitly stored zeros;
  oracle.products   - closed forms (sum of squared row lengths) and brute force;
  oracle.btiles     - the definition tile = col >> 6, bit = col & 63 evaluated in numpy;
  oracle.ctile_rows - nnz per row from the oracle's own product and from brute force;
  oracle.compare    - must reject planted mistakes (dropped entry, shifted column,
                      transposed operand, perturbed value).
"""
import json
import os

import numpy as np
import pytest

import hsmu
import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def dense(m, indicator=False):
    D = np.zeros((m.M, m.N), dtype=np.int64 if indicator else np.float64)
    for i in range(m.M):
        for k in range(m.rowptr[i], m.rowptr[i + 1]):
            D[i, m.col[k]] += 1 if indicator else m.val[k]
    return D


def brute(A, B):
    """C = A*B by dense products: values, and structure from stored-entry counts."""
    V = dense(A) @ dense(B)
    S = dense(A, True) @ dense(B, True)
    rowptr, col, val = [0], [], []
    for i in range(A.M):
        nz = np.nonzero(S[i])[0]
        col.extend(nz.tolist())
        val.extend(V[i, nz].tolist())
        rowptr.append(len(col))
    products = int((dense(A, True) @ dense(B, True).sum(axis=1)).sum())
    return np.array(rowptr), np.array(col, dtype=np.int32), np.array(val), products


def random_csr(M, N, max_per_row, seed, dup=False, zeros=False):
    rng = np.random.default_rng(seed)
    rowptr, col, val = [0], [], []
    for _ in range(M):
        k = int(rng.integers(0, max_per_row + 1))
        c = np.sort(rng.choice(N, size=min(k, N), replace=False)) if N else np.zeros(0, int)
        if dup and len(c):
            c = np.sort(np.concatenate([c, c[: len(c) // 2]]))
        v = rng.uniform(-1, 1, size=len(c))
        if zeros and len(v):
            v[0] = 0.0
        col.extend(c.tolist())
        val.extend(v.tolist())
        rowptr.append(len(col))
    return hsmu.HostMatrix(M, N, np.array(rowptr, dtype=np.int32), np.array(col, dtype=np.int32),
                           np.array(val))


def load_golden():
    with open(os.path.join(GOLDEN, "test_mtx_AxA.json")) as f:
        return json.load(f)


def test_golden_test_mtx():
    A, sym = hsmu.read_mtx(os.path.join(GOLDEN, "test.mtx"))
    assert sym and A.M == 5 and A.nnz == 9
    g = load_golden()
    P = oracle.spgemm(A, A)
    assert P.rowptr.tolist() == g["rowptr"]
    assert P.col.tolist() == g["col"]
    assert P.val.tolist() == g["val"]
    assert oracle.products(A, A) == g["products"] == 19


def laplace_square_geometry(n, dims):
    """A^2 for the Laplacian from grid geometry: offsets with L1 norm <= 2."""
    M = n ** dims
    diag = 2.0 * dims
    rowptr, col, val = [0], [], []
    strides = [n ** d for d in range(dims)]
    for p in range(M):
        coords = [(p // strides[d]) % n for d in range(dims)]
        nb = sum((coords[d] > 0) + (coords[d] < n - 1) for d in range(dims))
        entries = {}
        offs = [()]
        for _ in range(dims):
            offs = [o + (d,) for o in offs for d in (-2, -1, 0, 1, 2)]
        for o in offs:
            l1 = sum(abs(x) for x in o)
            if l1 > 2:
                continue
            q = [coords[d] + o[d] for d in range(dims)]
            if any(x < 0 or x >= n for x in q):
                continue
            qi = sum(q[d] * strides[d] for d in range(dims))
            if l1 == 0:
                v = diag * diag + nb
            elif l1 == 1:
                v = -2.0 * diag
            elif max(abs(x) for x in o) == 1:
                v = 2.0  # diagonal: two common neighbours
            else:
                v = 1.0  # two steps along one axis: the midpoint
            entries[qi] = v
        for qi in sorted(entries):
            col.append(qi)
            val.append(entries[qi])
        rowptr.append(len(col))
    return np.array(rowptr), np.array(col), np.array(val)


@pytest.mark.parametrize("n", [1, 2, 3, 7, 16])
def test_laplace2d_closed_form(n):
    A = hsmu.gen_laplace2d(n)
    rp, col, val = laplace_square_geometry(n, 2)
    P = oracle.spgemm(A, A)
    assert np.array_equal(P.rowptr, rp) and np.array_equal(P.col, col)
    assert np.array_equal(P.val, val)


@pytest.mark.parametrize("n", [2, 3, 6])
def test_laplace3d_closed_form(n):
    A = hsmu.gen_laplace3d(n)
    rp, col, val = laplace_square_geometry(n, 3)
    P = oracle.spgemm(A, A)
    assert np.array_equal(P.rowptr, rp) and np.array_equal(P.col, col)
    assert np.array_equal(P.val, val)


def test_laplace2d_128_counts():
    """The config's closed forms: nnz(C) = n^2 + 4n(n-1) + 4n(n-2) + 4(n-1)^2,
    products = 4*3^2 + 4(n-2)*4^2 + (n-2)^2*5^2."""
    n = 128
    A = hsmu.gen_laplace2d(n)
    assert A.nnz == 5 * n * n - 4 * n == 81408
    P = oracle.spgemm(A, A, want_abs=False)
    assert int(P.rowptr[-1]) == n * n + 4 * n * (n - 1) + 4 * n * (n - 2) + 4 * (n - 1) ** 2 == 210436
    assert oracle.products(A, A) == 4 * 9 + 4 * (n - 2) * 16 + (n - 2) ** 2 * 25 == 405000


@pytest.mark.parametrize("shape", [(1, 1, 1), (7, 5, 9), (20, 33, 17), (40, 130, 70), (3, 0, 4), (0, 4, 4)])
@pytest.mark.parametrize("flags", [(False, False), (True, True)])
def test_oracle_vs_dense(shape, flags):
    m, k, n = shape
    dup, zeros = flags
    A = random_csr(m, k, 6, seed=m * 131 + k, dup=dup, zeros=zeros)
    B = random_csr(k, n, 9, seed=k * 17 + n, dup=dup, zeros=zeros)
    rp, col, val, products = brute(A, B)
    P = oracle.spgemm(A, B)
    assert np.array_equal(P.rowptr, rp)
    assert np.array_equal(P.col, col)
    np.testing.assert_allclose(P.val, val, rtol=1e-13, atol=1e-13)
    assert oracle.products(A, B) == products


def test_btiles_definition():
    B = random_csr(50, 700, 40, seed=5, dup=True)
    tptr, tcol, tmask = oracle.btiles(B)
    k = 0
    for j in range(B.M):
        cols = B.col[B.rowptr[j]:B.rowptr[j + 1]]
        tiles = {}
        for c in cols:
            tiles[c >> 6] = tiles.get(c >> 6, 0) | (1 << int(c & 63))
        keys = sorted(tiles)
        assert tptr[j + 1] - tptr[j] == len(keys)
        for t in keys:
            assert tcol[k] == t and int(tmask[k]) == tiles[t]
            k += 1
    assert k == len(tcol)


def test_ctile_rows_matches_product():
    A = random_csr(60, 300, 12, seed=9)
    B = random_csr(300, 900, 25, seed=10)
    ct, cn = oracle.ctile_rows(A, B)
    rp, col, _, _ = brute(A, B)
    assert np.array_equal(cn, np.diff(rp))
    for i in range(A.M):
        assert ct[i] == len(set((col[rp[i]:rp[i + 1]] >> 6).tolist()))


def test_compare_rejects_planted_mistakes():
    A = random_csr(30, 40, 8, seed=1)
    B = random_csr(40, 50, 8, seed=2)
    P = oracle.spgemm(A, B)
    good = hsmu.HostMatrix(A.M, B.N, P.rowptr.copy(), P.col.copy(), P.val.copy())
    oracle.compare(good, P)
    # dropped entry
    rp = P.rowptr.copy()
    rp[5:] -= 1
    keep = np.ones(len(P.col), bool)
    keep[P.rowptr[4]] = False
    with pytest.raises(AssertionError):
        oracle.compare(hsmu.HostMatrix(A.M, B.N, rp, P.col[keep], P.val[keep]), P)
    # shifted column
    col = P.col.copy()
    col[3] += 1
    with pytest.raises(AssertionError):
        oracle.compare(hsmu.HostMatrix(A.M, B.N, P.rowptr, col, P.val), P)
    # value off by more than the tolerance
    val = P.val.copy()
    val[7] += 1e-9 * max(1.0, P.absval[7])
    with pytest.raises(AssertionError):
        oracle.compare(hsmu.HostMatrix(A.M, B.N, P.rowptr, P.col, val), P)
    # transposed operand: A * B^T on square inputs differs
    S = random_csr(30, 30, 6, seed=3)
    T = random_csr(30, 30, 6, seed=4)
    PT = oracle.spgemm(S, hsmu.transpose(T))
    with pytest.raises(AssertionError):
        PST = oracle.spgemm(S, T)
        oracle.compare(hsmu.HostMatrix(30, 30, PT.rowptr, PT.col, PT.val), PST)


def test_oracle_row_ranges_compose():
    B = hsmu.gen_banded(3000, 10, 30, 100, seed=3)
    full = oracle.spgemm(B, B)
    parts = [oracle.spgemm_rows(B, B, r0, min(r0 + 700, B.M)) for r0 in range(0, B.M, 700)]
    col = np.concatenate([p.col for p in parts])
    val = np.concatenate([p.val for p in parts])
    assert np.array_equal(col, full.col) and np.array_equal(val, full.val)
