"""Fill the autotuner table for this device: run hsmu_autotune's search on the standard
workloads and append the winners (one line per shape class) to the table.

    python hsmu-spgemm_amd/tune_table.py [--table hsmu-spgemm_amd/tuning/gfx950.tsv] [--only banded,batch]

Workloads: the 2D 5-point Laplacian on a 128x128 grid, the banded 62,451-row matrix,
the 512 uneven Laplacians of the batched form, and an R-MAT matrix of the given scale.
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
    for name in names:
        if name == "laplace":
            m = hsmu.gen_laplace2d(128)
            yield name, [m]
        elif name == "banded":
            yield name, [hsmu.gen_banded(62451, 40, 88, 400, seed=1)]
        elif name == "batch":
            yield name, hsmu.gen_laplace_batch(512, seed=3)
        elif name == "rmat":
            yield name, [hsmu.gen_rmat(rmat_scale, 16, 0.57, 0.19, 0.19, seed=2)]
        else:
            raise SystemExit(f"unknown workload {name}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--table", default=os.path.join(HERE, "tuning", "gfx950.tsv"))
    ap.add_argument("--only", default="laplace,banded,batch")
    ap.add_argument("--rmat-scale", type=int, default=16)
    args = ap.parse_args()
    os.makedirs(os.path.dirname(os.path.abspath(args.table)), exist_ok=True)
    if not os.path.exists(args.table):
        with open(args.table, "w") as f:
            f.write("# class products_per_lane tiles_per_lane b_lanes t_lanes lds_block_budget persistent rebuild_dirs ms\n")
    ctx = hsmu.Context(0)
    for name, mats in workloads(args.only.split(","), args.rmat_scale):
        dms = [ctx.matrix(m) for m in mats]
        cls = hsmu.shape_class(dms, dms)
        t0 = time.time()
        opts, found = ctx.autotune(dms, dms, base=hsmu.default_opts(timing=0), table=args.table, search=True)
        state = "in table" if found else f"tuned in {time.time() - t0:.1f}s"
        print(f"{name}: class {cls} {state}: products_per_lane {opts.products_per_lane} tiles_per_lane "
              f"{opts.tiles_per_lane} b_lanes {opts.b_lanes} t_lanes {opts.t_lanes} lds_block_budget "
              f"{opts.lds_block_budget} persistent {opts.persistent} rebuild_dirs {opts.rebuild_dirs}", flush=True)
        for d in dms:
            d.free()
    ctx.close()


if __name__ == "__main__":
    main()
