"""The large workloads beside the banded benchmark, measured and checked on one GPU.

    python hsmu-spgemm_amd/workloads.py rmat [--scale 20] [--steps 3] [--fp32c]
    python hsmu-spgemm_amd/workloads.py zipf [--rows 500000] [--cols 100000] [--panel-rows 10000] [--fp32c]
    python hsmu-spgemm_amd/workloads.py batch [--replays 10] [--subset all|small|large] [--no-sequential]

rmat: C = A*A, A from R-MAT (a, b, c = .57, .19, .19), edge factor 16, duplicates merged,
values U(0,1), seed 2; int64 row pointers (nnz(C) passes 2^31 at scale 20).
zipf: C = A*A^T, A rows x cols with 8 distinct columns per row drawn Zipf(1.2), values
U(-1,1), seed 4, B = A^T built explicitly.  C (about 2 TB at the full size) cannot be held
in device memory, so it is computed in row panels of A (hsmu_spgemm_rows), each checked and
released before the next: the run covers every product of the full A*A^T.
batch: the 512 uneven Laplacian products (256 2D with side floor(e^U(ln16,ln256)), 256
3D with side floor(e^U(ln8,ln48)), seed 3) as a plan with the autotuner table's options,
replayed as one HIP graph, beside the same products as sequential calls (timed in C);
--subset small keeps the instances with N < 4096, --subset large the others.

Check (rmat, zipf): every row sum of C equals A times the row sums of B (hsmu_csr_row_sums
on the device, the right side with scipy on the host), within 1e-12 of the row's sum of
magnitudes (FP64) or 1e-6 (compensated FP32).  (Entry-by-entry parity with the CPU oracle
is the GPU tests' job.)  One JSON line per workload.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402

import hsmu  # noqa: E402


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def row_sums_host(m):
    return np.bincount(np.repeat(np.arange(m.M), np.diff(m.rowptr)), weights=m.val, minlength=m.M)


def check_sums(ctx, c, A_rows, rb, tol):
    """max over rows of |sum - A*rowsum(B)| / (1 + sum of magnitudes)."""
    sums, asums = ctx.row_sums(c)
    want = A_rows.to_scipy() @ rb if A_rows.M else np.zeros(0)
    err = np.abs(sums - want) / (1.0 + asums)
    worst = float(err.max(initial=0.0))
    if worst > tol:
        raise SystemExit(f"row-sum check failed: {worst:.3e} > {tol:.1e}")
    return worst


def panel_of(A, r0, r1):
    a, b = int(A.rowptr[r0]), int(A.rowptr[r1])
    return hsmu.HostMatrix(r1 - r0, A.N, (A.rowptr[r0:r1 + 1] - a).astype(np.int32), A.col[a:b], A.val[a:b])


def as32(m):
    return hsmu.HostMatrix(m.M, m.N, m.rowptr, m.col, m.val.astype(np.float32).astype(np.float64))


def run_rmat(args):
    ctx = hsmu.Context(0)
    t0 = time.time()
    A = hsmu.gen_rmat(args.scale, 16, 0.57, 0.19, 0.19, seed=2)
    if args.fp32c:
        A = as32(A)
    log(f"rmat scale {args.scale}: {A.M} rows, nnz {A.nnz}, generated in {time.time() - t0:.1f}s")
    dA = ctx.matrix(A, val32=args.fp32c)
    opts = hsmu.default_opts(rowptr64=1, timing=1,
                             precision=hsmu.HSMU_FP32_COMP if args.fp32c else hsmu.HSMU_FP64)
    if args.tuned:
        table = os.path.join(HERE, "tuning", "gfx950.tsv")
        opts, _ = ctx.autotune([dA], [dA], base=opts, table=table, search=False)
    ms = []
    for it in range(args.steps + 1):
        t1 = time.perf_counter()
        c, st = ctx.spgemm_device(dA, dA, opts)
        wall = (time.perf_counter() - t1) * 1e3
        if it > 0:
            ms.append(wall)
        if it == args.steps:
            rb = row_sums_host(A)
            worst = check_sums(ctx, c, A, rb, 1e-6 if args.fp32c else 1e-12)
        ctx.free_result(c)
    med = float(np.median(ms))
    out = {"workload": f"rmat scale {args.scale}, edge factor 16, a/b/c .57/.19/.19, seed 2",
           "precision": "fp32_compensated" if args.fp32c else "fp64",
           "rows": A.M, "nnz_a": A.nnz, "products": st["products"], "nnz_c": st["nnz_c"],
           "ms": round(med, 2), "gflops": round(2.0 * st["products"] / (med * 1e6), 2),
           "alg_GBps": round(st["alg_bytes"] / (med * 1e-3) / 1e9, 1),
           "stage_ms": {k: round(st[k], 2) for k in ("ms_conversion", "ms_form_tileCptr", "ms_tileOR",
                                                      "ms_form_cval")},
           "row_sum_check_max": worst}
    print(json.dumps(out), flush=True)
    dA.free()
    ctx.close()


def run_zipf(args):
    ctx = hsmu.Context(0)
    t0 = time.time()
    A = hsmu.gen_zipf(args.rows, args.cols, 8, 1.2, seed=4)
    if args.fp32c:
        A = as32(A)
    B = hsmu.transpose(A)
    log(f"zipf {A.M}x{A.N}: nnz {A.nnz}, A^T built, {time.time() - t0:.1f}s")
    dA, dB = ctx.matrix(A, val32=args.fp32c), ctx.matrix(B, val32=args.fp32c)
    rb = row_sums_host(B)
    opts = hsmu.default_opts(rowptr64=1, timing=1,
                             precision=hsmu.HSMU_FP32_COMP if args.fp32c else hsmu.HSMU_FP64)
    # one untimed panel first, as the R-MAT driver's warm-up call: the context's pool makes
    # its device allocations there (the first allocations of tens of GB can take seconds
    # while memory another process has just freed is still being released)
    c, _ = ctx.spgemm_rows_device(dA, 0, min(A.M, args.panel_rows), dB, opts)
    ctx.free_result(c)
    products = nnz_c = 0
    gpu_ms = wall_ms = 0.0
    worst = 0.0
    npan = 0
    for r0 in range(0, A.M, args.panel_rows):
        r1 = min(A.M, r0 + args.panel_rows)
        t1 = time.perf_counter()
        c, st = ctx.spgemm_rows_device(dA, r0, r1, dB, opts)
        wall_ms += (time.perf_counter() - t1) * 1e3
        gpu_ms += st["ms_total"]
        products += st["products"]
        nnz_c += st["nnz_c"]
        worst = max(worst, check_sums(ctx, c, panel_of(A, r0, r1), rb, 1e-6 if args.fp32c else 1e-12))
        ctx.free_result(c)
        npan += 1
        log(f"panel {npan}: rows [{r0}, {r1}) nnz {st['nnz_c']} products {st['products']} {st['ms_total']:.1f} ms")
    out = {"workload": f"zipf A*A^T, A {A.M}x{A.N}, 8 columns per row Zipf(1.2), seed 4, "
                       f"row panels of {args.panel_rows} rows streamed (C is not held whole)",
           "precision": "fp32_compensated" if args.fp32c else "fp64",
           "panels": npan, "products": products, "nnz_c": nnz_c,
           "ms_gpu": round(gpu_ms, 1), "ms_wall": round(wall_ms, 1),
           "gflops_gpu": round(2.0 * products / (gpu_ms * 1e6), 2),
           "gflops_wall": round(2.0 * products / (wall_ms * 1e6), 2),
           "row_sum_check_max": worst}
    print(json.dumps(out), flush=True)
    dA.free()
    dB.free()
    ctx.close()


def run_batch(args):
    if args.replays < 1:
        raise SystemExit("--replays must be at least 1")
    ctx = hsmu.Context(0)
    mats = hsmu.gen_laplace_batch(512, seed=3)
    # subsets: the instances below the reference's small-path threshold (N < 4096, which
    # src/NHC_spgemm.cuh:7-43 sends to the dense-bitmap path), or the others
    if args.subset == "small":
        mats = [m for m in mats if m.N < 4096]
    elif args.subset == "large":
        mats = [m for m in mats if m.N >= 4096]
    dms = [ctx.matrix(m) for m in mats]
    ms_seq = None
    if not args.no_sequential:
        ctx.spgemm_each_ms(dms, dms, hsmu.default_opts(timing=0), 1)  # warm-up: the pool's first allocations
        ms_seq = round(ctx.spgemm_each_ms(dms, dms, hsmu.default_opts(timing=0), 3), 3)
    table = os.path.join(HERE, "tuning", "gfx950.tsv")
    opts, found = ctx.autotune(dms, dms, base=hsmu.default_opts(timing=0), table=table if os.path.exists(table) else None,
                               search=False)
    plan = ctx.plan(dms, dms, opts)
    dev = []
    for _ in range(args.replays):
        st = plan.execute()
        dev.append(st["ms_total"])
    out = {"workload": "512 uneven Laplacian products (256 2D, 256 3D, seed 3) as one plan, replayed as a HIP graph",
           "subset": args.subset, "instances": len(mats), "rows": sum(m.M for m in mats), "replays": args.replays,
           "ms_graph_device_median": round(float(np.median(dev)), 3), "products": st["products"],
           "nnz_c": st["nnz_c"], "tuned_from_table": bool(found),
           "ms_sequential_calls": ms_seq}
    print(json.dumps(out), flush=True)
    plan.destroy()
    for d in dms:
        d.free()
    ctx.close()


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="what", required=True)
    r = sub.add_parser("rmat")
    r.add_argument("--scale", type=int, default=20)
    r.add_argument("--steps", type=int, default=3)
    r.add_argument("--fp32c", action="store_true")
    r.add_argument("--tuned", action="store_true", help="options from the autotuner table for the class")
    z = sub.add_parser("zipf")
    z.add_argument("--rows", type=int, default=500000)
    z.add_argument("--cols", type=int, default=100000)
    z.add_argument("--panel-rows", type=int, default=10000)
    z.add_argument("--fp32c", action="store_true")
    b = sub.add_parser("batch")
    b.add_argument("--replays", type=int, default=10, help="graph replays after the plan's recording run (at least 1)")
    b.add_argument("--subset", choices=("all", "small", "large"), default="all",
                   help="all 512 instances, only those with N < 4096, or only the others")
    b.add_argument("--no-sequential", action="store_true",
                   help="skip the sequential calls (profiles of the plan's kernels alone)")
    args = ap.parse_args()
    if args.what == "rmat":
        run_rmat(args)
    elif args.what == "zipf":
        run_zipf(args)
    else:
        run_batch(args)


if __name__ == "__main__":
    main()
