"""Writes the golden fixtures under tests/golden/ (run once, in the build container).

test.mtx is the reference's only in-tree matrix, the 5x5 symmetric
other_spgemm_code/nsparse-master/data/test.mtx:1-21 of wuminqaq/HSMU-SpGEMM; it is
copied verbatim as data.  test_mtx_AxA.json holds A*A for it, computed here with a dense
numpy product of the symmetric expansion (independent of the oracle and of the GPU
path) and checked against the hand-derived values below.
"""
import json
import os
import shutil

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/other_spgemm_code/nsparse-master/data/test.mtx"

# A*A of the expanded matrix, derived by hand from the 7 stored entries
EXPECTED = {
    "rowptr": [0, 3, 4, 7, 8, 11],
    "col": [0, 2, 4, 1, 0, 2, 4, 3, 0, 2, 4],
    "val": [101.0, 40.0, 2.0, 400.0, 40.0, 905.0, 160.0, 1600.0, 2.0, 160.0, 2504.0],
    "products": 19,
}


def main():
    dst = os.path.join(HERE, "test.mtx")
    shutil.copyfile(SRC, dst)
    rows, cols, vals = [], [], []
    with open(dst) as f:
        lines = [ln for ln in f if not ln.startswith("%")]
    m, n, nz = map(int, lines[0].split())
    for ln in lines[1:1 + nz]:
        i, j, v = ln.split()
        rows.append(int(i) - 1)
        cols.append(int(j) - 1)
        vals.append(float(v))
    A = np.zeros((m, n))
    S = np.zeros((m, n), dtype=np.int64)  # stored-entry indicator (structure)
    for i, j, v in zip(rows, cols, vals):
        A[i, j] += v
        S[i, j] += 1
        if i != j:  # symmetric file: mirror
            A[j, i] += v
            S[j, i] += 1
    C = A @ A
    P = S @ S
    rowptr, col, val = [0], [], []
    for i in range(m):
        for j in range(n):
            if P[i, j] > 0:
                col.append(j)
                val.append(float(C[i, j]))
        rowptr.append(len(col))
    products = int(S.sum(axis=0) @ S.sum(axis=1))
    out = {"rowptr": rowptr, "col": col, "val": val, "products": products,
           "input": {"M": m, "N": n, "stored": [[r, c, v] for r, c, v in zip(rows, cols, vals)]}}
    for k in ("rowptr", "col", "val", "products"):
        assert out[k] == EXPECTED[k], (k, out[k], EXPECTED[k])
    with open(os.path.join(HERE, "test_mtx_AxA.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
