"""Benchmark: HSMU SpGEMM C = A*A on one MI355X, banded FEM-like workload.

Workload: A 62,451 x 62,451, per-row nnz uniform in [40, 88], distinct columns uniform
within +-400 of the diagonal, values N(0,1), seed 1 (the cant-like configuration the
metric is quoted on).  A step is one full SpGEMM through the C ABI (B -> tiles, row
analysis, symbolic, allocation of C from the context pool, numeric), with A already
resident in HBM; C is released back to the pool after each step.

Prints one JSON line.  value = 2 * products / step time (GFLOP/s, whole job).
roofline: the dominant kernel (the numeric kernel of the bin with the most time),
achieved = its algorithmic bytes (gather model: A row extent + A col/val + B row extent
per a_ij + one B col/val per product + C row pointer + C col/val, over its rows) / its
mean duration from HIP events on the library stream.  traffic: HBM bytes per launch from
the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary when present.
cpu_baseline: the CPU oracle (row-wise Gustavson, C + OpenMP) on the same workload,
repeated for about 10 s, median.  Options come from the autotuner table for the
workload's shape class when the table holds it (hsmu-spgemm_amd/tuning/gfx950.tsv).
Also reported: the compensated-FP32 form, the 512-instance batched form, and rocSPARSE's
SpGEMM on the same inputs.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "hsmu-spgemm_amd"))

PEAK_HBM_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_numeric_banded.json")
TUNE_TABLE = os.path.join(ROOT, "hsmu-spgemm_amd", "tuning", "gfx950.tsv")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def bench_batched(hsmu, ctx, steps, warmup, table):
    """512 uneven Laplacians (256 2D with side floor(e^U(ln16,ln256)), 256 3D with side
    floor(e^U(ln8,ln48)), seed 3), C_i = A_i * A_i: 512 sequential plain calls vs one
    stacked batched call vs the replay of its captured HIP graph (autotuned options from
    the committed table for the batch's shape class).  Results of the graph replay are
    compared entry for entry with the sequential calls' (integer-valued: identical)."""
    import numpy as np
    t0 = time.time()
    mats = hsmu.gen_laplace_batch(512, seed=3)
    dms = [ctx.matrix(m) for m in mats]
    log(f"batched: {len(mats)} instances, rows {sum(m.M for m in mats)}, nnz {sum(m.nnz for m in mats)}, "
        f"uploaded in {time.time() - t0:.1f}s")
    base = hsmu.default_opts(timing=0)
    tuned, found = ctx.autotune(dms, dms, base=base, table=table if os.path.exists(table) else None, search=False)
    cls = hsmu.shape_class(dms, dms)
    plain = hsmu.default_opts(timing=0)

    def seq_pass(keep=False):
        out = []
        for d in dms:
            c, _ = ctx.spgemm_device(d, d, plain)
            if keep:
                out.append(ctx.fetch(c))
            ctx.free_result(c)
        return out

    seq_pass()
    t1 = time.perf_counter()
    for _ in range(steps):
        seq_pass()
    ms_seq = (time.perf_counter() - t1) * 1e3 / steps

    for _ in range(warmup):
        cs, st = ctx.spgemm_batched_device(dms, dms, tuned)
        ctx.free_batch(cs)
    t1 = time.perf_counter()
    for _ in range(steps):
        cs, st = ctx.spgemm_batched_device(dms, dms, tuned)
        ctx.free_batch(cs)
    ms_batch = (time.perf_counter() - t1) * 1e3 / steps

    plan = ctx.plan(dms, dms, tuned)
    for _ in range(warmup):
        plan.execute()
    t1 = time.perf_counter()
    dev = []
    for _ in range(steps):
        dev.append(plan.execute()["ms_total"])
    ms_graph = (time.perf_counter() - t1) * 1e3 / steps

    ref = seq_pass(keep=True)
    identical = True
    for i, r in enumerate(ref):
        g = ctx.fetch(plan.results[i])
        identical = identical and np.array_equal(g.rowptr, r.rowptr) and np.array_equal(g.col, r.col) \
            and np.array_equal(g.val, r.val)
    plan.destroy()
    for d in dms:
        d.free()
    products = st["products"]
    out = {"instances": len(mats), "products": products, "nnz_c": st["nnz_c"],
           "ms_sequential_512_calls": round(ms_seq, 3), "ms_batched_call": round(ms_batch, 3),
           "ms_graph_replay": round(ms_graph, 3), "ms_graph_device": round(statistics.median(dev), 3),
           "speedup_graph_vs_sequential": round(ms_seq / ms_graph, 2),
           "speedup_batched_vs_sequential": round(ms_seq / ms_batch, 2),
           "gflops_graph": round(2.0 * products / (ms_graph * 1e6), 2),
           "alg_GBps_graph": round(st["alg_bytes"] / (ms_graph * 1e-3) / 1e9, 1),
           "identical_to_sequential": bool(identical),
           "shape_class": cls, "tuned_from_table": bool(found),
           "opts": {k: getattr(tuned, k) for k in ("products_per_lane", "tiles_per_lane", "b_lanes", "t_lanes",
                                                   "lds_block_budget", "persistent", "rebuild_dirs")}}
    log(f"batched: {out}")
    return out


def fp32c_errors(hsmu, ctx, A32, dA32, o32):
    """Errors of the compensated-FP32 result against FP64 on the same (FP32-rounded)
    inputs: the FP64 path of this library, whose results the tests hold to 1e-12 of the
    CPU oracle.  Relative error |c32 - c64| / |c64| (max, median over entries with
    c64 != 0) and the error over sum |a*b| (the FP64 product of |A| and |A|)."""
    import numpy as np
    c, _ = ctx.spgemm_device(dA32, dA32, o32)
    got = ctx.fetch(c)
    ctx.free_result(c)
    A64 = hsmu.HostMatrix(A32.M, A32.N, A32.rowptr, A32.col, A32.val.astype(np.float64))
    Aabs = hsmu.HostMatrix(A32.M, A32.N, A32.rowptr, A32.col, np.abs(A64.val))
    ref, _ = ctx.spgemm(A64, A64)
    mag, _ = ctx.spgemm(Aabs, Aabs)
    if not (np.array_equal(np.asarray(got.rowptr, np.int64), np.asarray(ref.rowptr, np.int64))
            and np.array_equal(got.col, ref.col)):
        raise SystemExit("compensated FP32: structure differs from FP64")
    err = np.abs(got.val.astype(np.float64) - ref.val)
    nz = ref.val != 0
    rel = err[nz] / np.abs(ref.val[nz])
    return {"rel_err_max": float(rel.max(initial=0.0)), "rel_err_median": float(np.median(rel)) if rel.size else 0.0,
            "frac_rel_err_above_1e-6": float(np.mean(rel > 1e-6)) if rel.size else 0.0,
            "err_over_sum_abs_max": float((err / np.maximum(mag.val, 1e-300)).max(initial=0.0)),
            "reference": "FP64 path on the same FP32-rounded inputs"}


def bench_rocsparse(hsmu, ctx, dA, products, ms_step, steps, warmup):
    """rocSPARSE's SpGEMM (the vendor library) on the same device inputs: per call, the
    device time of its nnz + compute stages and the host time of the whole call (which
    adds its buffer and C allocations).  The speedup is quoted against the stages only,
    the figure most favourable to rocSPARSE."""
    from hsmu import check
    for _ in range(warmup):
        c, _, _ = check.rocsparse_spgemm(ctx, dA, dA)
        check.free(c)
    stages, walls = [], []
    for _ in range(steps):
        c, ms_s, ms_w = check.rocsparse_spgemm(ctx, dA, dA)
        check.free(c)
        stages.append(ms_s)
        walls.append(ms_w)
    ms_s, ms_w = statistics.median(stages), statistics.median(walls)
    out = {"ms_stages": round(ms_s, 4), "ms_call": round(ms_w, 4),
           "gflops_stages": round(2.0 * products / (ms_s * 1e6), 2),
           "speedup_hsmu_vs_rocsparse_stages": round(ms_s / ms_step, 2),
           "speedup_hsmu_vs_rocsparse_call": round(ms_w / ms_step, 2)}
    log(f"rocsparse: {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32c", action="store_true", help="skip the compensated-FP32 form")
    ap.add_argument("--no-batched", action="store_true", help="skip the 512-instance batched form")
    ap.add_argument("--no-rocsparse", action="store_true", help="skip the rocSPARSE comparison")
    ap.add_argument("--table", default=TUNE_TABLE, help="autotuner table (shape class -> options)")
    ap.add_argument("--opts", default="", help="hsmu_opts overrides, e.g. products_per_lane=16,b_lanes=32")
    ap.add_argument("--copy-mib", type=int, default=1024,
                    help="size of the device-to-device copy probe (achievable HBM bandwidth; 0 = skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported by this build "
              f"(the path has no multi-GPU mode)", file=sys.stderr)
        sys.exit(2)

    import hsmu

    t0 = time.time()
    A = hsmu.gen_banded(62451, 40, 88, 400, seed=1)
    log(f"generated banded A: {A.M} rows, nnz {A.nnz} in {time.time() - t0:.1f}s")
    ctx = hsmu.Context(0)
    dA = ctx.matrix(A)
    copy_gbs = None
    if args.copy_mib > 0:
        nbytes = args.copy_mib << 20
        ms = ctx.copy_kernel_ms(nbytes)
        copy_gbs = 2.0 * nbytes / (ms * 1e-3) / 1e9  # read + write
        log(f"copy probe: {args.copy_mib} MiB in {ms:.3f} ms = {copy_gbs:.0f} GB/s (read+write)")
    overrides = dict(kv.split("=") for kv in args.opts.split(",") if kv)
    tuned_from_table = False
    # timing 2: HIP events around the numeric bins only (the dominant kernel's duration in
    # the timed steps, for the roofline), no per-stage events
    kv = {k: int(v) for k, v in overrides.items()}
    kv.setdefault("timing", 2)
    if overrides or not os.path.exists(args.table):
        opts = hsmu.default_opts(**kv)
    else:
        opts, tuned_from_table = ctx.autotune([dA], [dA], base=hsmu.default_opts(timing=kv["timing"]),
                                              table=args.table, search=False)
    opt_keys = ("products_per_lane", "tiles_per_lane", "b_lanes", "t_lanes", "lds_block_budget", "persistent",
                "rebuild_dirs")
    log(f"options ({'autotuner table' if tuned_from_table else 'defaults/overrides'}): "
        f"{ {k: getattr(opts, k) for k in opt_keys} }")

    for _ in range(args.warmup):
        c, st = ctx.spgemm_device(dA, dA, opts)
        ctx.free_result(c)
    hsmu.lib.hsmu_sync(ctx.handle)

    stats = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        c, st = ctx.spgemm_device(dA, dA, opts)
        ctx.free_result(c)
        stats.append(st)
    hsmu.lib.hsmu_sync(ctx.handle)
    elapsed = time.perf_counter() - t_start
    ms_step = elapsed * 1e3 / args.steps

    st0 = stats[0]
    products = st0["products"]
    nnz_c = st0["nnz_c"]
    gflops = 2.0 * products / (ms_step * 1e6)
    # dominant kernel: numeric bin with the largest mean time
    nb = len(st0["ms_numeric_bin"])
    mean_bin = [statistics.mean(s["ms_numeric_bin"][b] for s in stats) for b in range(nb)]
    dom = max(range(nb), key=lambda b: mean_bin[b])
    dom_ms = mean_bin[dom]
    dom_bytes = st0["num_bin_bytes"][dom]
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9 if dom_ms > 0 else 0.0
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    e2e_gbs = st0["alg_bytes"] / (ms_step * 1e-3) / 1e9
    log(f"step {ms_step:.3f} ms, {gflops:.1f} GFLOP/s, end-to-end {e2e_gbs:.0f} GB/s algorithmic; "
        f"numeric bins {st0['num_bin_rows']}; dominant bin {dom}: {dom_ms:.3f} ms, {achieved:.0f} GB/s")
    # per-stage breakdown: one more call, after the timed region, with every stage event
    o_stage = hsmu.Opts.from_buffer_copy(opts)
    o_stage.timing = 1
    c, st_s = ctx.spgemm_device(dA, dA, o_stage)
    ctx.free_result(c)
    stage = {k: round(st_s[k], 4) for k in ("ms_conversion", "ms_form_tileCptr", "ms_tileOR", "ms_malloc",
                                            "ms_form_cval", "ms_time2", "ms_total")}
    log(f"stages (one call after the timed region, all stage events): {stage}")

    # compensated-FP32 form on the same workload (values rounded to FP32)
    fp32c = None
    if not args.no_fp32c:
        import numpy as np
        A32 = hsmu.HostMatrix(A.M, A.N, A.rowptr, A.col, A.val.astype(np.float32))
        dA32 = ctx.matrix(A32, val32=True)
        o32 = hsmu.Opts.from_buffer_copy(opts)  # the same options and timing level as the FP64 steps
        o32.precision = hsmu.HSMU_FP32_COMP
        for _ in range(args.warmup):
            c, _ = ctx.spgemm_device(dA32, dA32, o32)
            ctx.free_result(c)
        hsmu.lib.hsmu_sync(ctx.handle)
        t1 = time.perf_counter()
        for _ in range(args.steps):
            c, st32 = ctx.spgemm_device(dA32, dA32, o32)
            ctx.free_result(c)
        hsmu.lib.hsmu_sync(ctx.handle)
        ms32 = (time.perf_counter() - t1) * 1e3 / args.steps
        fp32c = {"value": round(2.0 * products / (ms32 * 1e6), 3), "unit": "GFLOP/s", "ms_per_step": round(ms32, 4),
                 "speedup_vs_fp64": round(ms_step / ms32, 3), "numeric_ms": round(sum(st32["ms_numeric_bin"]), 4)}
        fp32c.update(fp32c_errors(hsmu, ctx, A32, dA32, o32))
        log(f"fp32 compensated: {fp32c}")
        dA32.free()

    rocsparse = None
    if not args.no_rocsparse:
        rocsparse = bench_rocsparse(hsmu, ctx, dA, products, ms_step, args.steps, args.warmup)

    batched = None
    if not args.no_batched:
        batched = bench_batched(hsmu, ctx, max(3, min(args.steps, 5)), 1, args.table)

    cpu = None
    if not args.no_cpu:
        import oracle
        runs = []
        t_cpu = time.perf_counter()
        while True:
            t1 = time.perf_counter()
            oracle.spgemm(A, A, want_abs=False)
            runs.append(time.perf_counter() - t1)
            if len(runs) >= 3 and time.perf_counter() - t_cpu > args.cpu_seconds:
                break
        med = statistics.median(runs)
        cpu = {"value": round(2.0 * products / (med * 1e9), 3), "unit": "GFLOP/s",
               "cores": oracle.max_threads(), "kind": "port",
               "sample": f"full workload (C=A*A, all {A.M} rows) repeated {len(runs)} times, "
                         f"median {med * 1e3:.1f} ms, row-wise Gustavson oracle (C, OpenMP)"}
        log(f"cpu baseline: {cpu}")

    out = {
        "metric": "SpGEMM GFLOP/s (2*products/time)",
        "value": round(gflops, 3),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "banded FEM-like C=A*A: 62451 rows, nnz/row U[40,88], columns within +-400, "
                               "N(0,1) values, seed 1",
                   "rows": A.M, "nnz_a": A.nnz, "products": products, "nnz_c": nnz_c,
                   "end_to_end_alg_GBps": round(e2e_gbs, 1), "stage_ms_untimed_call": stage,
                   "opts": {k: getattr(opts, k) for k in opt_keys}, "tuned_from_table": bool(tuned_from_table)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                     "kernel": f"k_numeric_lds bin {dom}" if dom < 8 else "k_numeric_global",
                     "kernel_ms": round(dom_ms, 4), "alg_bytes_per_launch": dom_bytes,
                     "copy_kernel_GBps": round(copy_gbs, 1) if copy_gbs else None},
        "cpu_baseline": cpu,
        "forms": {"fp32_compensated": fp32c, "batched": batched},
        "rocsparse": rocsparse,
    }
    print(json.dumps(out), flush=True)
    dA.free()
    ctx.close()


if __name__ == "__main__":
    main()
