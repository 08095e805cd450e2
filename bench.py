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
the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary when present.  frac_end_to_end: the
whole call's algorithmic bytes / step time / peak; frac_compulsory: the same for the
compulsory bytes (every input read once, C written once).
cpu_baseline: the CPU oracle (row-wise Gustavson, C + OpenMP) on the same workload,
repeated for about 10 s, median.  Options come from the autotuner table for the
workload's shape class when the table holds it (hsmu-spgemm_amd/tuning/gfx950.tsv).
Also reported (forms): the compensated-FP32 form with its errors against the FP64 oracle
on the same FP32-rounded inputs; the 512-instance batched form against 512 sequential
calls timed in C; R-MAT scale 20 and the streamed Zipf A*A^T in FP64 and compensated FP32
(errors on sampled rows against the oracle); and rocSPARSE's SpGEMM on the banded inputs.
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
OPT_KEYS = ("products_per_lane", "tiles_per_lane", "b_lanes", "t_lanes", "lds_block_budget", "persistent",
            "rebuild_dirs")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def as_fp32(hsmu, m):
    import numpy as np
    return hsmu.HostMatrix(m.M, m.N, m.rowptr, m.col, m.val.astype(np.float32).astype(np.float64))


def error_summary(got_val, ref, fp32):
    """Errors of result values against the oracle (same structure): relative error of the
    nonzeros (max, median, share above 1e-6) and the largest share of the tolerance used
    (FP64: 1e-12 sum|a*b|; compensated FP32: 2^-23 |c| + 1e-10 sum|a*b|)."""
    import numpy as np
    err = np.abs(np.asarray(got_val, np.float64) - ref.val)
    bound = (2.0 ** -23 * np.abs(ref.val) + 1e-10 * ref.absval) if fp32 else 1e-12 * ref.absval
    nz = ref.val != 0
    rel = err[nz] / np.abs(ref.val[nz])
    with np.errstate(divide="ignore", invalid="ignore"):
        used = np.where(bound > 0, err / bound, np.where(err > 0, np.inf, 0.0))
    return {"rel_err_max": float(rel.max(initial=0.0)), "rel_err_median": float(np.median(rel)) if rel.size else 0.0,
            "frac_rel_err_above_1e-6": float(np.mean(rel > 1e-6)) if rel.size else 0.0,
            "tolerance_used_max": float(used.max(initial=0.0))}


def row_errors(per_row):
    """Distribution of the per-row maximum and median relative errors."""
    import numpy as np
    mx = np.array([p[0] for p in per_row])
    md = np.array([p[1] for p in per_row])
    return {"rows": len(per_row), "row_max_rel_err": {"median": float(np.median(mx)), "p90": float(np.quantile(mx, 0.9)),
                                                      "max": float(mx.max())},
            "row_median_rel_err": {"median": float(np.median(md)), "max": float(md.max())}}


def full_errors(hsmu, ctx, A, dA, opts, fp32):
    """The whole banded product against the oracle: structure bit-exact (else exit), value
    errors, and the per-row error distribution."""
    import numpy as np
    import oracle
    c, _ = ctx.spgemm_device(dA, dA, opts)
    got = ctx.fetch(c)
    ctx.free_result(c)
    ref = oracle.spgemm(A, A)
    if not (np.array_equal(np.asarray(got.rowptr, np.int64), ref.rowptr) and np.array_equal(got.col, ref.col)):
        raise SystemExit("structure differs from the oracle")
    out = error_summary(got.val, ref, fp32)
    val = np.asarray(got.val, np.float64)
    err = np.abs(val - ref.val)
    per_row = []
    for i in range(0, A.M, max(1, A.M // 2000)):
        a, b = int(ref.rowptr[i]), int(ref.rowptr[i + 1])
        nz = ref.val[a:b] != 0
        r = err[a:b][nz] / np.abs(ref.val[a:b][nz])
        if r.size:
            per_row.append((float(r.max()), float(np.median(r))))
    out["per_row"] = row_errors(per_row)
    out["reference"] = "oracle (FP64 Gustavson) on the same inputs" + (", values rounded to FP32" if fp32 else "")
    return out


def sampled_row_errors(hsmu, ctx, c, A, B, rows, r0, fp32):
    """Sampled rows of a device result (rows r0.. of A*B) against the oracle's rows, read
    slice by slice; structure must be bit-exact.  Returns error summary + per-row errors."""
    import numpy as np
    import oracle
    rp = ctx.download(c.rowptr, int(c.M) + 1, np.int64 if c.rowptr64 else np.int32)
    vsz, vt = (4, np.float32) if c.val32 else (8, np.float64)
    per_row, worst = [], 0.0
    for r in rows:
        a, b = int(rp[r]), int(rp[r + 1])
        ref = oracle.spgemm_rows(A, B, r0 + r, r0 + r + 1)
        col = ctx.download(c.col + 4 * a, b - a, np.int32)
        val = ctx.download(c.val + vsz * a, b - a, vt)
        if b - a != int(ref.rowptr[-1]) or not np.array_equal(col, ref.col):
            raise SystemExit(f"row {r0 + r}: structure differs from the oracle")
        s = error_summary(val, ref, fp32)
        per_row.append((s["rel_err_max"], s["rel_err_median"]))
        worst = max(worst, s["tolerance_used_max"])
    out = row_errors(per_row)
    out["tolerance_used_max"] = worst
    return out


def sample_rows(ctx, c, n_wide, n_rand, seed):
    import numpy as np
    rp = ctx.download(c.rowptr, int(c.M) + 1, np.int64 if c.rowptr64 else np.int32)
    lens = np.diff(rp)
    return [int(r) for r in np.concatenate([np.argsort(-lens, kind="stable")[:n_wide],
                                            np.random.default_rng(seed).integers(0, int(c.M), size=n_rand)])]


def bench_batched(hsmu, ctx, steps, warmup, table):
    """512 uneven Laplacians (256 2D with side floor(e^U(ln16,ln256)), 256 3D with side
    floor(e^U(ln8,ln48)), seed 3), C_i = A_i * A_i: 512 sequential plain calls (timed
    inside the library around the whole loop, hsmu_spgemm_each) vs one stacked batched
    call vs the replay of its captured HIP graph, with the autotuned options of the
    batch's shape class and in the persistent form (work queues).  Results of the graph
    replay are compared entry for entry with the sequential calls' (integer-valued:
    identical); the tests compare every instance with the oracle."""
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

    ctx.spgemm_each_ms(dms, dms, plain, 1)
    ms_seq = ctx.spgemm_each_ms(dms, dms, plain, steps)

    for _ in range(warmup):
        cs, st = ctx.spgemm_batched_device(dms, dms, tuned)
        ctx.free_batch(cs)
    t1 = time.perf_counter()
    for _ in range(steps):
        cs, st = ctx.spgemm_batched_device(dms, dms, tuned)
        ctx.free_batch(cs)
    ms_batch = (time.perf_counter() - t1) * 1e3 / steps

    def replay(opts):
        plan = ctx.plan(dms, dms, opts)
        for _ in range(warmup):
            plan.execute()
        t1 = time.perf_counter()
        dev = []
        for _ in range(steps):
            dev.append(plan.execute()["ms_total"])
        return plan, (time.perf_counter() - t1) * 1e3 / steps, statistics.median(dev)

    pers = hsmu.Opts.from_buffer_copy(tuned)
    pers.persistent = 1
    plan_p, ms_graph_p, dev_p = replay(pers)
    plan_p.destroy()
    plan, ms_graph, dev = replay(tuned)

    ref = []
    for d in dms:
        c, _ = ctx.spgemm_device(d, d, plain)
        ref.append(ctx.fetch(c))
        ctx.free_result(c)
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
           "ms_sequential_512_calls": round(ms_seq, 3),
           "sequential_timing": "hsmu_spgemm_each: host clock around the 512-call loop inside the library",
           "ms_batched_call": round(ms_batch, 3),
           "ms_graph_replay": round(ms_graph, 3), "ms_graph_device": round(dev, 3),
           "ms_graph_replay_persistent": round(ms_graph_p, 3), "ms_graph_device_persistent": round(dev_p, 3),
           "speedup_graph_vs_sequential": round(ms_seq / ms_graph, 2),
           "speedup_graph_persistent_vs_sequential": round(ms_seq / ms_graph_p, 2),
           "speedup_batched_vs_sequential": round(ms_seq / ms_batch, 2),
           "gflops_graph": round(2.0 * products / (ms_graph * 1e6), 2),
           "alg_GBps_graph": round(st["alg_bytes"] / (ms_graph * 1e-3) / 1e9, 1),
           "frac_end_to_end_graph": round(st["alg_bytes"] / (ms_graph * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
           "identical_to_sequential": bool(identical),
           "shape_class": cls, "tuned_from_table": bool(found),
           "opts": {k: getattr(tuned, k) for k in OPT_KEYS}}
    log(f"batched: {out}")
    return out


def timed_calls(ctx, dA, dB, opts, steps, warmup, r0=None, r1=None):
    """Median host time of `steps` calls after `warmup` (results released), with the last
    call's stats."""
    def call():
        if r0 is None:
            return ctx.spgemm_device(dA, dB, opts)
        return ctx.spgemm_rows_device(dA, r0, r1, dB, opts)
    for _ in range(warmup):
        c, st = call()
        ctx.free_result(c)
    ms = []
    for _ in range(steps):
        t1 = time.perf_counter()
        c, st = call()
        ms.append((time.perf_counter() - t1) * 1e3)
        ctx.free_result(c)
    return statistics.median(ms), st


def bench_rmat(hsmu, ctx, steps, table):
    """R-MAT scale 20 (a, b, c = .57, .19, .19, edge factor 16, duplicates merged, values
    U(0,1), seed 2), C = A*A with int64 row pointers, FP64 and compensated FP32 (values
    rounded to FP32): median of `steps` calls after one warm-up, the wide-row bin's time
    and algorithmic bytes, and errors of 16 widest + 32 random rows against the oracle."""
    t0 = time.time()
    A = hsmu.gen_rmat(20, 16, 0.57, 0.19, 0.19, seed=2)
    log(f"rmat scale 20: {A.M} rows, nnz {A.nnz}, generated in {time.time() - t0:.1f}s")
    out = {"workload": "R-MAT scale 20, edge factor 16, a/b/c .57/.19/.19, seed 2, C = A*A"}
    for name, fp32 in (("fp64", False), ("fp32_compensated", True)):
        Am = as_fp32(hsmu, A) if fp32 else A
        dA = ctx.matrix(Am, val32=fp32)
        base = hsmu.default_opts(rowptr64=1, timing=2, precision=hsmu.HSMU_FP32_COMP if fp32 else hsmu.HSMU_FP64)
        opts, found = ctx.autotune([dA], [dA], base=base, table=table if os.path.exists(table) else None,
                                   search=False)
        ms, st = timed_calls(ctx, dA, dA, opts, steps, 1)
        wide = 8
        r = {"ms": round(ms, 2), "gflops": round(2.0 * st["products"] / (ms * 1e6), 2),
             "alg_GBps": round(st["alg_bytes"] / (ms * 1e-3) / 1e9, 1),
             "frac_end_to_end": round(st["alg_bytes"] / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
             "frac_compulsory": round(st["compulsory_bytes"] / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
             "products": st["products"], "nnz_c": st["nnz_c"],
             "wide_bin": {"kernels": "k_wide_count/order/fill, k_wide_cols, k_numeric_wsplit"
                                     + (", k_wide_final" if fp32 else ""),
                          "rows": st["num_bin_rows"][wide], "ms": round(st["ms_numeric_bin"][wide], 2),
                          "alg_bytes": st["num_bin_bytes"][wide],
                          "alg_GBps": round(st["num_bin_bytes"][wide] / (st["ms_numeric_bin"][wide] * 1e-3) / 1e9, 1)
                          if st["ms_numeric_bin"][wide] > 0 else None},
             "ms_numeric_bins": [round(x, 2) for x in st["ms_numeric_bin"]], "lds_use": round(st["lds_use"], 4),
             "shape_class": hsmu.shape_class([dA], [dA]), "tuned_from_table": bool(found),
             "opts": {k: getattr(opts, k) for k in OPT_KEYS}}
        c, _ = ctx.spgemm_device(dA, dA, opts)
        try:
            r["errors"] = sampled_row_errors(hsmu, ctx, c, Am, Am, sample_rows(ctx, c, 16, 32, 20), 0, fp32)
        finally:
            ctx.free_result(c)
        dA.free()
        ctx.trim()
        out[name] = r
        log(f"rmat {name}: {r}")
    out["speedup_fp32c_vs_fp64"] = round(out["fp64"]["ms"] / out["fp32_compensated"]["ms"], 3)
    return out


class _Shape:
    """Sizes of a row panel as the autotuner's shape class sees them (no arrays)."""

    def __init__(self, hsmu, M, N, nnz):
        self.csr = hsmu.Csr(M, N, nnz, 0, 0, 0, 0, 0)


def bench_zipf(hsmu, ctx, table, panel_rows=10000):
    """Zipf A*A^T, A 500,000 x 100,000, 8 distinct columns per row Zipf(1.2), values
    U(-1,1), seed 4: C (about 2.7 TB) streamed in row panels (hsmu_spgemm_rows), every
    product of the full A*A^T computed, in FP64 and compensated FP32; device time summed
    over panels and host time; errors of 64 sampled rows of the panel at row 250,000."""
    t0 = time.time()
    A = hsmu.gen_zipf(500000, 100000, 8, 1.2, seed=4)
    log(f"zipf: {A.M}x{A.N} nnz {A.nnz}, generated in {time.time() - t0:.1f}s")
    out = {"workload": f"Zipf A*A^T, A 500000x100000, 8 columns per row Zipf(1.2), seed 4, "
                       f"streamed in panels of {panel_rows} rows (C is never held whole)"}
    for name, fp32 in (("fp64", False), ("fp32_compensated", True)):
        Am = as_fp32(hsmu, A) if fp32 else A
        Bm = hsmu.transpose(Am)
        dA, dB = ctx.matrix(Am, val32=fp32), ctx.matrix(Bm, val32=fp32)
        base = hsmu.default_opts(rowptr64=1, timing=1, precision=hsmu.HSMU_FP32_COMP if fp32 else hsmu.HSMU_FP64)
        e0 = 250000  # the panel whose rows are checked; its class is every full panel's
        pa = [_Shape(hsmu, panel_rows, A.N, int(A.rowptr[e0 + panel_rows] - A.rowptr[e0]))]
        opts, found = ctx.autotune(pa, [dB], base=base, table=table if os.path.exists(table) else None,
                                   search=False)
        c, _ = ctx.spgemm_rows_device(dA, 0, panel_rows, dB, opts)  # untimed: the pool's first allocations
        ctx.free_result(c)
        products = nnz_c = 0
        dev = wall = alg = 0.0
        for r0 in range(0, A.M, panel_rows):
            r1 = min(A.M, r0 + panel_rows)
            t1 = time.perf_counter()
            c, st = ctx.spgemm_rows_device(dA, r0, r1, dB, opts)
            wall += (time.perf_counter() - t1) * 1e3
            ctx.free_result(c)
            dev += st["ms_total"]
            products += st["products"]
            nnz_c += st["nnz_c"]
            alg += st["alg_bytes"]
        r = {"panels": (A.M + panel_rows - 1) // panel_rows, "products": products, "nnz_c": nnz_c,
             "ms_device": round(dev, 1), "ms_wall": round(wall, 1),
             "gflops_device": round(2.0 * products / (dev * 1e6), 2),
             "gflops_wall": round(2.0 * products / (wall * 1e6), 2),
             "frac_end_to_end_device": round(alg / (dev * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
             "shape_class": hsmu.shape_class(pa, [dB]), "tuned_from_table": bool(found),
             "opts": {k: getattr(opts, k) for k in OPT_KEYS}}
        c, _ = ctx.spgemm_rows_device(dA, e0, e0 + panel_rows, dB, opts)
        try:
            r["errors"] = sampled_row_errors(hsmu, ctx, c, Am, Bm, sample_rows(ctx, c, 8, 56, e0), e0, fp32)
            r["errors"]["rows_sampled_from"] = f"panel [{e0}, {e0 + panel_rows}): 8 widest + 56 random rows"
        finally:
            ctx.free_result(c)
        dA.free()
        dB.free()
        ctx.trim()
        out[name] = r
        log(f"zipf {name}: {r}")
    out["speedup_fp32c_vs_fp64_device"] = round(out["fp64"]["ms_device"] / out["fp32_compensated"]["ms_device"], 3)
    return out


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
    ap.add_argument("--no-rmat", action="store_true", help="skip R-MAT scale 20")
    ap.add_argument("--no-zipf", action="store_true", help="skip the streamed Zipf A*A^T")
    ap.add_argument("--no-errors", action="store_true", help="skip the banded error checks against the oracle")
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

    prov = hsmu.provenance()
    log(f"libhsmu.so build: {prov}")
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
    log(f"options ({'autotuner table' if tuned_from_table else 'defaults/overrides'}): "
        f"{ {k: getattr(opts, k) for k in OPT_KEYS} }")

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
    comp_gbs = st0["compulsory_bytes"] / (ms_step * 1e-3) / 1e9
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
    fp64_err = None
    if not args.no_errors:
        fp64_err = full_errors(hsmu, ctx, A, dA, opts, False)
        log(f"fp64 errors: {fp64_err}")

    # compensated-FP32 form on the same workload (values rounded to FP32)
    fp32c = None
    if not args.no_fp32c:
        import numpy as np
        A32 = as_fp32(hsmu, A)
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
                 "speedup_vs_fp64": round(ms_step / ms32, 3), "numeric_ms": round(sum(st32["ms_numeric_bin"]), 4),
                 "numeric_ms_fp64": round(sum(mean_bin), 4)}
        if not args.no_errors:
            fp32c["errors"] = full_errors(hsmu, ctx, A32, dA32, o32, True)
        log(f"fp32 compensated: {fp32c}")
        dA32.free()

    rocsparse = None
    if not args.no_rocsparse:
        rocsparse = bench_rocsparse(hsmu, ctx, dA, products, ms_step, args.steps, args.warmup)

    batched = None
    if not args.no_batched:
        batched = bench_batched(hsmu, ctx, max(3, min(args.steps, 5)), 1, args.table)

    dA.free()
    ctx.trim()
    rmat = None
    if not args.no_rmat:
        rmat = bench_rmat(hsmu, ctx, 3, args.table)
    zipf = None
    if not args.no_zipf:
        zipf = bench_zipf(hsmu, ctx, args.table)

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
                   "lds_use": round(st0["lds_use"], 4), "lds_launched_bytes": st0["lds_launched"],
                   "opts": {k: getattr(opts, k) for k in OPT_KEYS}, "tuned_from_table": bool(tuned_from_table),
                   "fp64_errors": fp64_err},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                     "kernel": f"k_numeric_lds bin {dom}" if dom < 8 else "wide-row bin",
                     "kernel_ms": round(dom_ms, 4), "alg_bytes_per_launch": dom_bytes,
                     "frac_end_to_end": round(e2e_gbs / PEAK_HBM_GBS, 4),
                     "alg_bytes_end_to_end": st0["alg_bytes"],
                     "compulsory_bytes": st0["compulsory_bytes"],
                     "frac_compulsory": round(comp_gbs / PEAK_HBM_GBS, 4),
                     "copy_kernel_GBps": round(copy_gbs, 1) if copy_gbs else None},
        "cpu_baseline": cpu,
        "forms": {"fp32_compensated": fp32c, "batched": batched, "rmat20": rmat, "zipf_aat": zipf},
        "rocsparse": rocsparse,
        "build": prov,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
