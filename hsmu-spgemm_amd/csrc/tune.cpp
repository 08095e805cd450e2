// tune.cpp -- autotuner of the launch parameters, stored per shape class.
//
// The reference fixes its bin ranges and workgroup sizes in the source for one GPU
// (src/bining.cuh:12,121,228; SM counts in inc/NHC_spgemm.h).  Here the parameters that
// set a row's bin and workgroup shape are options (hsmu_opts), and this file picks them
// per shape class of the input by timing candidates on the device, keeping the winners
// in a plain-text table:
//
//   # class products_per_lane tiles_per_lane b_lanes t_lanes lds_block_budget persistent rebuild_dirs ms
//   c10_a3_b3_m22 32 4 0 0 65536 1 0 3.512
//
// Lines are appended; the last line of a class wins.  Tuning changes only how rows are
// grouped and walked, never the result's structure (FP64 values may differ in the last
// bits, because the order of the atomic additions changes with the grouping).

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hsmu.h"

namespace {

struct Knobs {
    int ppl, tpl, bl, tl, budget, persistent, rebuild;
};

Knobs from_opts(const hsmu_opts &o)
{
    return {o.products_per_lane, o.tiles_per_lane, o.b_lanes, o.t_lanes, o.lds_block_budget, o.persistent,
            o.rebuild_dirs};
}

void to_opts(const Knobs &k, hsmu_opts &o)
{
    o.products_per_lane = k.ppl;
    o.tiles_per_lane = k.tpl;
    o.b_lanes = k.bl;
    o.t_lanes = k.tl;
    o.lds_block_budget = k.budget;
    o.persistent = k.persistent;
    o.rebuild_dirs = k.rebuild;
}

bool lookup(const char *path, const std::string &cls, Knobs &k)
{
    if (!path)
        return false;
    FILE *f = fopen(path, "r");
    if (!f)
        return false;
    char line[512];
    bool found = false;
    while (fgets(line, sizeof(line), f)) {
        if (line[0] == '#')
            continue;
        char name[128];
        Knobs t;
        double ms;
        if (sscanf(line, "%127s %d %d %d %d %d %d %d %lf", name, &t.ppl, &t.tpl, &t.bl, &t.tl, &t.budget,
                   &t.persistent, &t.rebuild, &ms) == 9 &&
            cls == name) {
            k = t;
            found = true;
        }
    }
    fclose(f);
    return found;
}

// One timed call: the whole SpGEMM (or batch) on the device, results released.
int run_once(hsmu_ctx *ctx, int64_t count, const hsmu_csr *A, const hsmu_csr *B, const hsmu_opts &o, double *ms)
{
    std::vector<hsmu_csr> C((size_t)count);
    const auto t0 = std::chrono::steady_clock::now();
    int rc = count == 1 ? hsmu_spgemm(ctx, A, B, C.data(), &o, nullptr)
                        : hsmu_spgemm_batched(ctx, count, A, B, C.data(), &o, nullptr);
    *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (rc)
        return rc;
    return count == 1 ? hsmu_csr_free(ctx, C.data()) : hsmu_batch_free(ctx, count, C.data());
}

// median of three timed calls after one warm-up
int measure(hsmu_ctx *ctx, int64_t count, const hsmu_csr *A, const hsmu_csr *B, const hsmu_opts &o, double *ms)
{
    double t[3], w;
    int rc = run_once(ctx, count, A, B, o, &w);
    for (int i = 0; i < 3 && !rc; ++i)
        rc = run_once(ctx, count, A, B, o, &t[i]);
    if (rc)
        return rc;
    std::sort(t, t + 3);
    *ms = t[1];
    return HSMU_OK;
}

} // namespace

extern "C" int hsmu_shape_class(int64_t count, const hsmu_csr *A, const hsmu_csr *B, char *cls, size_t cls_len)
{
    if (count < 1 || !A || !B || !cls || cls_len == 0)
        return HSMU_ERR_INVALID;
    double M = 0, nnzA = 0, K = 0, nnzB = 0;
    for (int64_t i = 0; i < count; ++i) {
        M += (double)A[i].M;
        nnzA += (double)A[i].nnz;
        K += (double)B[i].M;
        nnzB += (double)B[i].nnz;
    }
    auto rlog2 = [](double v) { return v > 0 ? (int)std::lround(std::log2(v)) : -1; };
    const int c = count == 1 ? 0 : 1 + (int)std::floor(std::log2((double)count));
    const int ea = rlog2(M > 0 ? nnzA / M : 0), eb = rlog2(K > 0 ? nnzB / K : 0);
    const int em = M >= 1 ? (int)std::floor(std::log2(M)) : -1;
    const int n = snprintf(cls, cls_len, "c%d_a%d_b%d_m%d", c, ea, eb, em);
    return (n < 0 || (size_t)n >= cls_len) ? HSMU_ERR_INVALID : HSMU_OK;
}

extern "C" int hsmu_autotune(hsmu_ctx *ctx, int64_t count, const hsmu_csr *A, const hsmu_csr *B,
                             const hsmu_opts *base, const char *table_path, int search, hsmu_opts *out,
                             int32_t *found)
{
    if (!ctx || !out || count < 1)
        return HSMU_ERR_INVALID;
    hsmu_opts o;
    if (base)
        o = *base;
    else
        hsmu_opts_default(&o);
    if (found)
        *found = 0;
    char cls[128];
    int rc = hsmu_shape_class(count, A, B, cls, sizeof(cls));
    if (rc)
        return rc;
    Knobs k = from_opts(o);
    if (lookup(table_path, cls, k)) {
        to_opts(k, o);
        *out = o;
        if (found)
            *found = 1;
        return HSMU_OK;
    }
    if (!search) {
        *out = o;
        return HSMU_OK;
    }
    o.timing = 0;
    // coordinate descent: one knob at a time, in order of measured effect, keeping the
    // best value of each before moving to the next
    double best;
    rc = measure(ctx, count, A, B, o, &best);
    if (rc)
        return rc;
    struct Axis {
        int Knobs::*field;
        std::vector<int> values;
    };
    const std::vector<Axis> axes = {
        {&Knobs::persistent, {0, 1}},
        {&Knobs::ppl, {8, 16, 32, 64, 128}},
        {&Knobs::bl, {0, 2, 4, 8, 16, 32}},
        {&Knobs::tpl, {2, 4, 8, 16}},
        {&Knobs::tl, {0, 1, 2, 4, 8}},
        {&Knobs::budget, {2048, 4096, 6144, 8192, 12288, 16384, 32768}},
        {&Knobs::rebuild, {0, 1}},
    };
    for (const Axis &ax : axes) {
        const int keep = k.*(ax.field);
        int pick = keep;
        for (int v : ax.values) {
            if (v == keep)
                continue;
            Knobs t = k;
            t.*(ax.field) = v;
            hsmu_opts ot = o;
            to_opts(t, ot);
            double ms;
            rc = measure(ctx, count, A, B, ot, &ms);
            if (rc)
                return rc;
            if (ms < best) {
                best = ms;
                pick = v;
            }
        }
        k.*(ax.field) = pick;
    }
    hsmu_opts res = base ? *base : o;
    if (!base)
        hsmu_opts_default(&res);
    to_opts(k, res);
    *out = res;
    if (table_path) {
        FILE *f = fopen(table_path, "a");
        if (f) {
            fprintf(f, "%s %d %d %d %d %d %d %d %.4f\n", cls, k.ppl, k.tpl, k.bl, k.tl, k.budget, k.persistent,
                    k.rebuild, best);
            fclose(f);
        }
    }
    return HSMU_OK;
}
