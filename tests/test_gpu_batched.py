"""Batched form, HIP-graph plans, persistent work-queue kernels and the autotuner,
through the C ABI, against the CPU oracle and against single hsmu_spgemm calls.

Structure must be bit-exact; FP64 values within 1e-12 * sum |a*b| (integer-valued
Laplacians: exact); compensated FP32 within the bound of the FP32 tests.
"""
import ctypes

import numpy as np
import pytest

import hsmu
import oracle

pytestmark = pytest.mark.gpu
TOL = 1e-12


def note(msg):
    print(f"[gpu-case] {msg}", flush=True)


def laplace_batch(n2, n3, seed):
    """Uneven 2D/3D Laplacians (sides log-uniform, as the batched workload draws them)."""
    rng = np.random.default_rng(seed)
    mats = [hsmu.gen_laplace2d(int(np.exp(rng.uniform(np.log(2), np.log(40))))) for _ in range(n2)]
    mats += [hsmu.gen_laplace3d(int(np.exp(rng.uniform(np.log(2), np.log(12))))) for _ in range(n3)]
    order = rng.permutation(len(mats))
    return [mats[i] for i in order]


def mixed_pairs(seed, n):
    """Uneven, rectangular, empty and wide-N instances."""
    rng = np.random.default_rng(seed)
    pairs = []
    for i in range(n):
        kind = i % 6
        if kind == 0:
            A = hsmu.gen_laplace2d(int(rng.integers(2, 30)))
            B = A
        elif kind == 1:
            m, k, n_ = (int(v) for v in rng.integers(1, 300, 3))
            A = hsmu.gen_random(m, k, 8, seed * 100 + i)
            B = hsmu.gen_random(k, n_, 20, seed * 100 + i + 50)
        elif kind == 2:
            A = hsmu.gen_banded(int(rng.integers(100, 1500)), 5, 40, 60, seed=seed * 100 + i)
            B = A
        elif kind == 3:
            A = hsmu.gen_random(int(rng.integers(1, 60)), 40, 5, seed * 100 + i)
            B = hsmu.gen_random(40, 9000, 4, seed * 100 + i + 7)  # wide N, several tiles
        elif kind == 4:
            A = hsmu.gen_random(int(rng.integers(1, 30)), 25, 0, seed * 100 + i)  # no entries
            B = hsmu.gen_random(25, 70, 6, seed * 100 + i + 3)
        else:
            A = hsmu.gen_random(0, 12, 4, seed * 100 + i)  # no rows
            B = hsmu.gen_random(12, 33, 4, seed * 100 + i + 5)
        pairs.append((A, B))
    return pairs


def upload_pairs(ctx, pairs, val32=False):
    dA, dB = [], []
    for A, B in pairs:
        a = ctx.matrix(A, val32)
        dA.append(a)
        dB.append(a if B is A else ctx.matrix(B, val32))
    return dA, dB


def free_all(mats):
    seen = set()
    for m in mats:
        if id(m) not in seen:
            seen.add(id(m))
            m.free()


def check_views(ctx, cs, pairs, tol=TOL, exact=False):
    for i, (A, B) in enumerate(pairs):
        Cm = ctx.fetch(cs[i])
        ref = oracle.spgemm(A, B)
        oracle.compare(Cm, ref, tol=tol)
        if exact:
            assert np.array_equal(Cm.val, ref.val), f"instance {i}"


def test_batched_laplacians_match_oracle_and_single_calls(ctx):
    mats = laplace_batch(12, 12, seed=3)
    pairs = [(m, m) for m in mats]
    note(f"batched laplacians: {len(pairs)} instances, rows {sum(m.M for m in mats)}")
    dA, dB = upload_pairs(ctx, pairs)
    cs, st = ctx.spgemm_batched_device(dA, dB)
    try:
        check_views(ctx, cs, pairs, exact=True)
        assert st["products"] == sum(oracle.products(m, m) for m in mats)
        assert st["nnz_c"] == sum(int(cs[i].nnz) for i in range(len(mats)))
        for i, a in enumerate(dA[:6]):
            c1, _ = ctx.spgemm_device(a, a)
            one = ctx.fetch(c1)
            ctx.free_result(c1)
            many = ctx.fetch(cs[i])
            assert np.array_equal(one.rowptr, many.rowptr) and np.array_equal(one.col, many.col)
            assert np.array_equal(one.val, many.val)
    finally:
        ctx.free_batch(cs)
        free_all(dA)


@pytest.mark.parametrize("seed", [1, 2])
def test_batched_mixed_instances(ctx, seed):
    pairs = mixed_pairs(seed, 18)
    note(f"batched mixed seed {seed}: shapes {[(A.M, A.N, B.N) for A, B in pairs]}")
    dA, dB = upload_pairs(ctx, pairs)
    cs, _ = ctx.spgemm_batched_device(dA, dB)
    try:
        check_views(ctx, cs, pairs)
        for i, (A, B) in enumerate(pairs):
            assert cs[i].M == A.M and cs[i].N == B.N
    finally:
        ctx.free_batch(cs)
        free_all(dA + dB)


@pytest.mark.parametrize("opts", [dict(persistent=1), dict(force_global=1), dict(rowptr64=1),
                                  dict(lds_block_budget=4096), dict(rebuild_dirs=1),
                                  dict(products_per_lane=4, tiles_per_lane=1, b_lanes=2, t_lanes=1)])
def test_batched_options(ctx, opts):
    pairs = mixed_pairs(5, 12)
    note(f"batched options {opts}")
    dA, dB = upload_pairs(ctx, pairs)
    cs, _ = ctx.spgemm_batched_device(dA, dB, hsmu.default_opts(**opts))
    try:
        check_views(ctx, cs, pairs)
    finally:
        ctx.free_batch(cs)
        free_all(dA + dB)


def as_fp32(m):
    """The same matrix with values rounded to FP32 (kept in float64 for the oracle)."""
    return hsmu.HostMatrix(m.M, m.N, m.rowptr, m.col, m.val.astype(np.float32).astype(np.float64))


def test_batched_fp32_compensated(ctx):
    pairs = []
    for A, B in mixed_pairs(7, 12):
        A32 = as_fp32(A)
        pairs.append((A32, A32) if B is A else (A32, as_fp32(B)))
    note(f"batched fp32c: {len(pairs)} instances")
    dA, dB = upload_pairs(ctx, pairs, val32=True)
    cs, _ = ctx.spgemm_batched_device(dA, dB, hsmu.default_opts(precision=hsmu.HSMU_FP32_COMP))
    try:
        for i, (A, B) in enumerate(pairs):
            Cm = ctx.fetch(cs[i])
            assert Cm.val.dtype == np.float32
            ref = oracle.spgemm(A, B)
            assert np.array_equal(np.asarray(Cm.rowptr, np.int64), ref.rowptr) and np.array_equal(Cm.col, ref.col)
            bound = 2.0 ** -23 * np.abs(ref.val) + 1e-10 * ref.absval
            assert np.all(np.abs(Cm.val.astype(np.float64) - ref.val) <= bound), f"instance {i}"
    finally:
        ctx.free_batch(cs)
        free_all(dA + dB)


def wide_pair(seed, heavy, na, K, N, blen):
    """A with `heavy` rows of `na` entries (then 50 rows of 0..20) over B (K x N, `blen`
    entries per row): the heavy rows of C are wide and split into several work items."""
    rng = np.random.default_rng(seed)
    rows_b = [np.sort(rng.choice(N, size=blen, replace=False)).astype(np.int32) for _ in range(K)]
    rp_b = np.zeros(K + 1, np.int32)
    rp_b[1:] = np.cumsum([len(r) for r in rows_b])
    B = hsmu.HostMatrix(K, N, rp_b, np.concatenate(rows_b), rng.uniform(-1, 1, int(rp_b[-1])))
    rows_a = [np.sort(rng.choice(K, size=na if i < heavy else int(rng.integers(0, 21)), replace=False)).astype(np.int32)
              for i in range(heavy + 50)]
    rp_a = np.zeros(len(rows_a) + 1, np.int32)
    rp_a[1:] = np.cumsum([len(r) for r in rows_a])
    A = hsmu.HostMatrix(len(rows_a), K, rp_a, np.concatenate(rows_a), rng.uniform(-1, 1, int(rp_a[-1])))
    return A, B


def check_fp32c_views(ctx, cs, pairs):
    for i, (A, B) in enumerate(pairs):
        Cm = ctx.fetch(cs[i])
        assert Cm.val.dtype == np.float32
        ref = oracle.spgemm(A, B)
        assert np.array_equal(np.asarray(Cm.rowptr, np.int64), ref.rowptr) and np.array_equal(Cm.col, ref.col)
        bound = 2.0 ** -23 * np.abs(ref.val) + 1e-10 * ref.absval
        assert np.all(np.abs(Cm.val.astype(np.float64) - ref.val) <= bound), f"instance {i}"


def test_batched_and_plan_fp32c_split_wide_rows(ctx):
    """Compensated FP32 with wide rows split into work items in several instances, batched
    and as a plan replayed with new values: each split row's (s, c) scratch range is handed
    out across the stacked instances, and the plan's capture sizes the scratch from the
    count its recording run kept."""
    R = as_fp32(hsmu.gen_rmat(11, 16, seed=2))
    pairs = [tuple(as_fp32(m) for m in wide_pair(1, 3, 2000, 3000, 200000, 40)), (R, R),
             tuple(as_fp32(m) for m in wide_pair(2, 2, 1500, 2500, 60000, 120))]
    dA, dB = upload_pairs(ctx, pairs, val32=True)
    o = hsmu.default_opts(precision=hsmu.HSMU_FP32_COMP)
    note("fp32c batched split wide rows")
    cs, _ = ctx.spgemm_batched_device(dA, dB, o)
    try:
        check_fp32c_views(ctx, cs, pairs)
    finally:
        ctx.free_batch(cs)
    note("fp32c plan with split wide rows: record + capture")
    plan = ctx.plan(dA, dB, o)
    try:
        check_fp32c_views(ctx, plan.results, pairs)
        rng = np.random.default_rng(5)
        for rep in range(2):
            new_pairs = []
            for (A, B), a, b in zip(pairs, dA, dB):
                A2 = hsmu.HostMatrix(A.M, A.N, A.rowptr, A.col, rng.uniform(-1, 1, A.nnz).astype(np.float32).astype(np.float64))
                set_values(ctx, a, A2.val.astype(np.float32))
                if b is a:
                    B2 = A2
                else:
                    B2 = hsmu.HostMatrix(B.M, B.N, B.rowptr, B.col,
                                         rng.uniform(-1, 1, B.nnz).astype(np.float32).astype(np.float64))
                    set_values(ctx, b, B2.val.astype(np.float32))
                new_pairs.append((A2, B2))
            note(f"fp32c plan replay {rep}")
            plan.execute()
            check_fp32c_views(ctx, plan.results, new_pairs)
    finally:
        plan.destroy()
        free_all(dA + dB)


def test_batched_all_empty(ctx):
    pairs = [(hsmu.gen_random(5 + i, 9, 0, i), hsmu.gen_random(9, 11, 3, i + 9)) for i in range(4)]
    dA, dB = upload_pairs(ctx, pairs)
    cs, st = ctx.spgemm_batched_device(dA, dB)
    try:
        assert st["nnz_c"] == 0
        for i, (A, B) in enumerate(pairs):
            Cm = ctx.fetch(cs[i])
            assert Cm.nnz == 0 and Cm.rowptr.tolist() == [0] * (A.M + 1)
    finally:
        ctx.free_batch(cs)
        free_all(dA + dB)


def test_batched_count_one_is_a_single_call(ctx):
    A = hsmu.gen_banded(3000, 20, 60, 300, seed=4)
    dA, _ = upload_pairs(ctx, [(A, A)])
    cs, _ = ctx.spgemm_batched_device(dA, dA)
    try:
        check_views(ctx, cs, [(A, A)])
    finally:
        ctx.free_batch(cs)
        free_all(dA)


def test_batched_errors_fail_loudly(ctx):
    good = hsmu.gen_laplace2d(6)
    bad = hsmu.HostMatrix(good.M, good.N, good.rowptr, good.col.copy(), good.val)
    bad.col[3] = good.N + 5  # column beyond the rows of B
    dA, _ = upload_pairs(ctx, [(good, good)])
    dbad = ctx.matrix(bad)
    try:
        with pytest.raises(hsmu.HsmuError) as e:
            ctx.spgemm_batched_device([dA[0], dbad], [dA[0], dA[0]])
        assert e.value.status == hsmu.HSMU_ERR_INDEX
        other = ctx.matrix(hsmu.gen_random(4, 7, 3, 1))
        with pytest.raises(hsmu.HsmuError) as e:  # A.N != B.M in instance 1
            ctx.spgemm_batched_device([dA[0], other], [dA[0], dA[0]])
        assert e.value.status == hsmu.HSMU_ERR_INVALID
        other.free()
        cs, _ = ctx.spgemm_batched_device(dA, dA)  # the context still works
        check_views(ctx, cs, [(good, good)])
        ctx.free_batch(cs)
    finally:
        dbad.free()
        free_all(dA)


def set_values(ctx, dm, vals):
    vals = np.ascontiguousarray(vals)
    hsmu._check(hsmu.lib.hsmu_copy_h2d(ctx.handle, dm.val, vals.ctypes.data, vals.nbytes), "h2d")


def set_cols(ctx, dm, cols):
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    hsmu._check(hsmu.lib.hsmu_copy_h2d(ctx.handle, dm.col, cols.ctypes.data, cols.nbytes), "h2d")


def test_plan_replays_match_eager(ctx):
    pairs = mixed_pairs(11, 14)
    dA, dB = upload_pairs(ctx, pairs)
    note("plan: record + capture")
    plan = ctx.plan(dA, dB, hsmu.default_opts(persistent=1))
    try:
        check_views(ctx, plan.results, pairs)
        rng = np.random.default_rng(0)
        for rep in range(3):
            # new values, same structure; eager work between the replays
            new_pairs = []
            for (A, B), a, b in zip(pairs, dA, dB):
                A2 = hsmu.HostMatrix(A.M, A.N, A.rowptr, A.col, rng.standard_normal(A.nnz))
                set_values(ctx, a, A2.val)
                if b is a:
                    B2 = A2
                else:
                    B2 = hsmu.HostMatrix(B.M, B.N, B.rowptr, B.col, rng.standard_normal(B.nnz))
                    set_values(ctx, b, B2.val)
                new_pairs.append((A2, B2))
            note(f"plan replay {rep}")
            st = plan.execute()
            assert st["ms_total"] > 0
            check_views(ctx, plan.results, new_pairs)
            cs, _ = ctx.spgemm_batched_device(dA, dB)
            for i in range(len(pairs)):
                e, r = ctx.fetch(cs[i]), ctx.fetch(plan.results[i])
                assert np.array_equal(e.rowptr, r.rowptr) and np.array_equal(e.col, r.col)
            ctx.free_batch(cs)
    finally:
        plan.destroy()
        free_all(dA + dB)


def test_plan_single_product(ctx):
    A = hsmu.gen_banded(8000, 30, 70, 500, seed=8)
    dA, _ = upload_pairs(ctx, [(A, A)])
    plan = ctx.plan(dA, dA)
    try:
        for rep in range(3):
            note(f"single plan replay {rep}")
            plan.execute()
            check_views(ctx, plan.results, [(A, A)])
    finally:
        plan.destroy()
        free_all(dA)


def test_plan_detects_structure_change(ctx):
    """A replay either reports HSMU_ERR_PLAN or computes the product of the changed
    inputs exactly; after the inputs are restored it is correct again."""
    mats = laplace_batch(4, 4, seed=9)
    pairs = [(m, m) for m in mats]
    dA, dB = upload_pairs(ctx, pairs)
    plan = ctx.plan(dA, dB)
    try:
        A = mats[0]
        # move every entry of the first row to column 0 (duplicates allowed): nnz per
        # row and tile spans change
        col = A.col.copy()
        col[A.rowptr[0]:A.rowptr[1]] = 0
        changed = hsmu.HostMatrix(A.M, A.N, A.rowptr, col, A.val)
        set_cols(ctx, dA[0], col)
        note("plan replay on changed structure")
        try:
            plan.execute()
            check_views(ctx, plan.results[:1], [(changed, changed)])
        except hsmu.HsmuError as e:
            assert e.status == hsmu.HSMU_ERR_PLAN
        set_cols(ctx, dA[0], A.col)
        plan.execute()
        check_views(ctx, plan.results, pairs, exact=True)
    finally:
        plan.destroy()
        free_all(dA)


@pytest.mark.parametrize("case", ["banded", "rmat", "laplace"])
def test_persistent_matches_static(ctx, case):
    if case == "banded":
        A = hsmu.gen_banded(20000, 20, 80, 900, seed=2)
    elif case == "rmat":
        A = hsmu.gen_rmat(12, 16, 0.57, 0.19, 0.19, seed=2)
    else:
        A = hsmu.gen_laplace3d(20)
    note(f"persistent vs static: {case}")
    C1, _ = ctx.spgemm(A, A, persistent=1)
    ref = oracle.spgemm(A, A)
    oracle.compare(C1, ref, tol=TOL)


def test_autotune_table_roundtrip(ctx, tmp_path):
    mats = laplace_batch(3, 3, seed=12)
    pairs = [(m, m) for m in mats]
    dA, dB = upload_pairs(ctx, pairs)
    table = str(tmp_path / "tune.tsv")
    try:
        cls = hsmu.shape_class(dA, dB)
        note(f"autotune class {cls}")
        o1, found1 = ctx.autotune(dA, dB, table=table, search=True)
        assert not found1
        lines = [l for l in open(table) if not l.startswith("#")]
        assert len(lines) == 1 and lines[0].split()[0] == cls
        o2, found2 = ctx.autotune(dA, dB, table=table, search=False)
        assert found2
        for f in ("products_per_lane", "tiles_per_lane", "b_lanes", "t_lanes", "lds_block_budget", "persistent",
                  "rebuild_dirs"):
            assert getattr(o1, f) == getattr(o2, f)
        cs, _ = ctx.spgemm_batched_device(dA, dB, o2)
        check_views(ctx, cs, pairs, exact=True)
        ctx.free_batch(cs)
    finally:
        free_all(dA)
