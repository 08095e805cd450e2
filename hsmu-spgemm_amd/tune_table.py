"""Fill the autotuner table for this device: run hsmu_autotune's search on the standard
workloads and append the winners (one line per shape class) to the table.

    python hsmu-spgemm_amd/tune_table.py [--table hsmu-spgemm_amd/tuning/gfx950.tsv] [--only banded,batch]

Workloads: the 2D 5-point Laplacian on a 128x128 grid, the banded 62,451-row matrix,
the 512 uneven Laplacians of the batched form, an R-MAT matrix of the given scale
(C = A*A, int64 row pointers), and one 10,000-row panel of the Zipf A*A^T (A 500,000 x
100,000, the panel's rows as a matrix of their own times A^T, int64 row pointers).
A class already in the table is skipped (delete its line to tune it again).
"""
import argparse
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import hsmu  # noqa: E402


def workloads(names, rmat_scale):
    """(name, (As, Bs)) per workload; Bs is As for self-products, which share arrays."""
    for name in names:
        if name == "laplace":
            ms = [hsmu.gen_laplace2d(128)]
            yield name, (ms, ms)
        elif name == "banded":
            ms = [hsmu.gen_banded(62451, 40, 88, 400, seed=1)]
            yield name, (ms, ms)
        elif name == "batch":
            ms = hsmu.gen_laplace_batch(512, seed=3)
            yield name, (ms, ms)
        elif name == "rmat":
            ms = [hsmu.gen_rmat(rmat_scale, 16, 0.57, 0.19, 0.19, seed=2)]
            yield name, (ms, ms)
        elif name == "zipf":
            A = hsmu.gen_zipf(500000, 100000, 8, 1.2, seed=4)
            r0, r1 = 250000, 260000
            a, b = int(A.rowptr[r0]), int(A.rowptr[r1])
            P = hsmu.HostMatrix(r1 - r0, A.N, (A.rowptr[r0:r1 + 1] - a).astype("int32"), A.col[a:b], A.val[a:b])
            yield name, ([P], [hsmu.transpose(A)])
        else:
            raise SystemExit(f"unknown workload {name}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--table", default=os.path.join(HERE, "tuning", "gfx950.tsv"))
    ap.add_argument("--only", default="laplace,banded,batch")
    ap.add_argument("--rmat-scale", type=int, default=20)
    args = ap.parse_args()
    os.makedirs(os.path.dirname(os.path.abspath(args.table)), exist_ok=True)
    if not os.path.exists(args.table):
        with open(args.table, "w") as f:
            f.write("# class products_per_lane tiles_per_lane b_lanes t_lanes lds_block_budget persistent rebuild_dirs ms\n")
    ctx = hsmu.Context(0)
    for name, (As, Bs) in workloads(args.only.split(","), args.rmat_scale):
        dA = [ctx.matrix(m) for m in As]
        dB = dA if Bs is As else [ctx.matrix(m) for m in Bs]
        cls = hsmu.shape_class(dA, dB)
        t0 = time.time()
        base = hsmu.default_opts(timing=0, rowptr64=1 if name in ("rmat", "zipf") else 0)
        opts, found = ctx.autotune(dA, dB, base=base, table=args.table, search=True)
        state = "in table" if found else f"tuned in {time.time() - t0:.1f}s"
        print(f"{name}: class {cls} {state}: products_per_lane {opts.products_per_lane} tiles_per_lane "
              f"{opts.tiles_per_lane} b_lanes {opts.b_lanes} t_lanes {opts.t_lanes} lds_block_budget "
              f"{opts.lds_block_budget} persistent {opts.persistent} rebuild_dirs {opts.rebuild_dirs}", flush=True)
        for d in dA + ([] if dB is dA else dB):
            d.free()
        ctx.trim()
    ctx.close()


if __name__ == "__main__":
    main()
