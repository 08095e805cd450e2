// api.cpp -- host orchestrator and C ABI of libhsmu.so.
//
// hsmu_spgemm replaces the reference's NHC_spgemm -> compressed_NHC_spgemm /
// small_scale_NHC_spgemm (src/NHC_spgemm.cuh:7-43, src/compressed/compressed_NHC_spgemm.cu:22-248,
// src/small/small_scale_NHC_spgemm.cu:4-121).  One HIP stream per context instead of the
// reference's 14 (src/loading.cuh:55-59); two host synchronisations per call instead of
// about ten (the reference copies bin sizes and whole row-pointer arrays back at every
// stage): one after the row analysis (validation and symbolic launch sizes) and one
// after the symbolic phase (nnz(C) and numeric launch sizes).

#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/hsmu.h"
#include "hsmu_internal.h"

using namespace hsmu;

namespace {

thread_local std::string g_last_error;

int fail(int status, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return status;
}

#define HIP_TRY(expr)                                                                                                  \
    do {                                                                                                               \
        hipError_t e_ = (expr);                                                                                        \
        if (e_ != hipSuccess)                                                                                          \
            return fail(HSMU_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__);      \
    } while (0)

// Caching device allocator: freed blocks stay mapped and are handed out again to
// requests of up to twice their size, so repeated calls do not hipMalloc.  All work
// runs on the context's one stream, so stream order makes reuse safe.
class Pool {
  public:
    ~Pool() { trim(); }

    void *alloc(size_t bytes)
    {
        if (bytes == 0)
            bytes = 1;
        const size_t gran = bytes < (size_t(1) << 20) ? 256 : (size_t(2) << 20);
        const size_t sz = (bytes + gran - 1) / gran * gran;
        auto it = free_.lower_bound(sz);
        if (it != free_.end() && it->first <= 2 * sz) {
            void *p = it->second;
            used_[p] = it->first;
            free_.erase(it);
            return p;
        }
        void *p = nullptr;
        if (hipMalloc(&p, sz) != hipSuccess) {
            (void)hipGetLastError();
            trim();
            if (hipMalloc(&p, sz) != hipSuccess) {
                (void)hipGetLastError();
                return nullptr;
            }
        }
        held_ += sz;
        used_[p] = sz;
        return p;
    }

    bool owns(void *p) const { return used_.count(p) != 0; }

    void release(void *p)
    {
        auto it = used_.find(p);
        if (it == used_.end())
            return;
        free_.emplace(it->second, p);
        used_.erase(it);
    }

    void trim()
    {
        for (auto &kv : free_) {
            (void)hipFree(kv.second);
            held_ -= kv.first;
        }
        free_.clear();
    }

    size_t held() const { return held_; }

  private:
    std::multimap<size_t, void *> free_;
    std::unordered_map<void *, size_t> used_;
    size_t held_ = 0;
};

// Scoped workspace: everything taken from the pool inside one call goes back at exit.
class Scratch {
  public:
    explicit Scratch(Pool &p) : pool_(p) {}
    ~Scratch()
    {
        for (void *q : taken_)
            pool_.release(q);
    }
    template <typename T> T *get(size_t count)
    {
        void *q = pool_.alloc(count * sizeof(T));
        if (q)
            taken_.push_back(q);
        return (T *)q;
    }

  private:
    Pool &pool_;
    std::vector<void *> taken_;
};

int pow2_floor(long long v)
{
    int r = 1;
    while ((long long)r * 2 <= v)
        r *= 2;
    return r;
}

int ilog2(int v)
{
    int r = 0;
    while ((1 << (r + 1)) <= v)
        ++r;
    return r;
}

} // namespace

enum { EV_START, EV_CONV, EV_ANALYSIS, EV_SYMBOLIC, EV_ALLOC, EV_NUMERIC, EV_COUNT };

struct hsmu_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    Pool pool;
    DevStats *host_stats = nullptr; // pinned
    hipEvent_t ev[EV_COUNT] = {};
    hipEvent_t evb[NBINS][2] = {};
    int num_cus = 0;
    int lds_max = 0;
};

extern "C" {

const char *hsmu_last_error(void) { return g_last_error.c_str(); }

const char *hsmu_status_string(int status)
{
    switch (status) {
    case HSMU_OK: return "ok";
    case HSMU_ERR_INVALID: return "invalid argument";
    case HSMU_ERR_UNSORTED: return "row of B not sorted by column";
    case HSMU_ERR_INDEX: return "index out of range";
    case HSMU_ERR_OVERFLOW: return "nnz(C) overflows the row-pointer type";
    case HSMU_ERR_NOMEM: return "out of memory";
    case HSMU_ERR_HIP: return "HIP runtime error";
    case HSMU_ERR_UNSUPPORTED: return "unsupported";
    default: return "unknown status";
    }
}

void hsmu_opts_default(hsmu_opts *o)
{
    if (!o)
        return;
    memset(o, 0, sizeof(*o));
    o->precision = HSMU_FP64;
    o->timing = 1;
}

int hsmu_ctx_create(int device, hsmu_ctx **out)
{
    if (!out)
        return fail(HSMU_ERR_INVALID, "hsmu_ctx_create: out is NULL");
    *out = nullptr;
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev)
        return fail(HSMU_ERR_INVALID, "hsmu_ctx_create: device %d of %d", device, ndev);
    HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    hsmu_ctx *c = new hsmu_ctx();
    c->device = device;
    c->num_cus = prop.multiProcessorCount;
    c->lds_max = (int)prop.sharedMemPerBlock;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipHostMalloc((void **)&c->host_stats, sizeof(DevStats), hipHostMallocDefault) != hipSuccess) {
        delete c;
        return fail(HSMU_ERR_HIP, "hsmu_ctx_create: stream or pinned memory");
    }
    for (int i = 0; i < EV_COUNT; ++i)
        HIP_TRY(hipEventCreate(&c->ev[i]));
    for (int b = 0; b < NBINS; ++b)
        for (int k = 0; k < 2; ++k)
            HIP_TRY(hipEventCreate(&c->evb[b][k]));
    static std::mutex mu;
    static std::vector<int> configured;
    {
        std::lock_guard<std::mutex> lk(mu);
        bool done = false;
        for (int d : configured)
            done |= d == device;
        if (!done) {
            // Best effort: HIP on AMD accepts dynamic LDS up to the device limit without
            // the attribute; a refusal here is not an error, a launch that needs more
            // LDS than allowed reports its own.
            if (configure_kernels(c->lds_max - LDS_STATIC_RESERVE) != hipSuccess)
                (void)hipGetLastError();
            configured.push_back(device);
        }
    }
    *out = c;
    return HSMU_OK;
}

void hsmu_ctx_destroy(hsmu_ctx *c)
{
    if (!c)
        return;
    (void)hipSetDevice(c->device);
    if (c->stream)
        (void)hipStreamSynchronize(c->stream);
    c->pool.trim();
    for (int i = 0; i < EV_COUNT; ++i)
        if (c->ev[i])
            (void)hipEventDestroy(c->ev[i]);
    for (int b = 0; b < NBINS; ++b)
        for (int k = 0; k < 2; ++k)
            if (c->evb[b][k])
                (void)hipEventDestroy(c->evb[b][k]);
    if (c->host_stats)
        (void)hipHostFree(c->host_stats);
    if (c->stream)
        (void)hipStreamDestroy(c->stream);
    delete c;
}

void *hsmu_ctx_stream(hsmu_ctx *c) { return c ? (void *)c->stream : nullptr; }
int hsmu_ctx_trim(hsmu_ctx *c)
{
    if (!c)
        return fail(HSMU_ERR_INVALID, "ctx is NULL");
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->pool.trim();
    return HSMU_OK;
}
int64_t hsmu_ctx_pool_bytes(hsmu_ctx *c) { return c ? (int64_t)c->pool.held() : 0; }
int hsmu_ctx_lds_max(hsmu_ctx *c) { return c ? c->lds_max : 0; }
int hsmu_ctx_num_cus(hsmu_ctx *c) { return c ? c->num_cus : 0; }

int hsmu_device_alloc(hsmu_ctx *c, size_t bytes, void **ptr)
{
    if (!c || !ptr)
        return fail(HSMU_ERR_INVALID, "hsmu_device_alloc: NULL argument");
    HIP_TRY(hipSetDevice(c->device));
    *ptr = nullptr;
    if (hipMalloc(ptr, bytes ? bytes : 1) != hipSuccess) {
        (void)hipGetLastError();
        return fail(HSMU_ERR_NOMEM, "hipMalloc(%zu) failed", bytes);
    }
    return HSMU_OK;
}

int hsmu_device_free(hsmu_ctx *c, void *ptr)
{
    if (!c)
        return fail(HSMU_ERR_INVALID, "ctx is NULL");
    if (ptr)
        HIP_TRY(hipFree(ptr));
    return HSMU_OK;
}

int hsmu_copy_h2d(hsmu_ctx *c, void *dst, const void *src, size_t bytes)
{
    if (!c || (bytes && (!dst || !src)))
        return fail(HSMU_ERR_INVALID, "hsmu_copy_h2d: NULL argument");
    if (bytes) {
        HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return HSMU_OK;
}

int hsmu_copy_d2h(hsmu_ctx *c, void *dst, const void *src, size_t bytes)
{
    if (!c || (bytes && (!dst || !src)))
        return fail(HSMU_ERR_INVALID, "hsmu_copy_d2h: NULL argument");
    if (bytes) {
        HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return HSMU_OK;
}

int hsmu_sync(hsmu_ctx *c)
{
    if (!c)
        return fail(HSMU_ERR_INVALID, "ctx is NULL");
    HIP_TRY(hipStreamSynchronize(c->stream));
    return HSMU_OK;
}

int hsmu_copy_kernel(hsmu_ctx *c, void *dst, const void *src, size_t bytes, double *ms)
{
    if (!c || !dst || !src)
        return fail(HSMU_ERR_INVALID, "hsmu_copy_kernel: NULL argument");
    if (bytes % 16)
        return fail(HSMU_ERR_INVALID, "hsmu_copy_kernel: bytes must be a multiple of 16");
    HIP_TRY(hipEventRecord(c->ev[EV_START], c->stream));
    HIP_TRY(launch_copy(src, dst, bytes, c->num_cus, c->stream));
    HIP_TRY(hipEventRecord(c->ev[EV_CONV], c->stream));
    HIP_TRY(hipEventSynchronize(c->ev[EV_CONV]));
    float t = 0;
    HIP_TRY(hipEventElapsedTime(&t, c->ev[EV_START], c->ev[EV_CONV]));
    if (ms)
        *ms = t;
    return HSMU_OK;
}

} // extern "C"

namespace {

int check_input(const hsmu_csr *X, const char *name, bool want_val32)
{
    if (!X)
        return fail(HSMU_ERR_INVALID, "%s is NULL", name);
    if (X->M < 0 || X->N < 0 || X->nnz < 0)
        return fail(HSMU_ERR_INVALID, "%s has a negative size", name);
    if (X->M >= INT32_MAX || X->N > INT32_MAX || X->nnz > INT32_MAX)
        return fail(HSMU_ERR_UNSUPPORTED, "%s: sizes beyond int32 indices", name);
    if (X->rowptr64)
        return fail(HSMU_ERR_UNSUPPORTED, "%s: input row pointers must be int32", name);
    if ((bool)X->val32 != want_val32)
        return fail(HSMU_ERR_INVALID, "%s: value type does not match the precision", name);
    if (!X->rowptr)
        return fail(HSMU_ERR_INVALID, "%s: rowptr is NULL", name);
    if (X->nnz > 0 && (!X->col || !X->val))
        return fail(HSMU_ERR_INVALID, "%s: col/val NULL with nnz > 0", name);
    return HSMU_OK;
}

int device_error_status(int err)
{
    if (err & ERR_ROWPTR)
        return fail(HSMU_ERR_INDEX, "row pointer decreasing or beyond nnz");
    if (err & ERR_ACOL)
        return fail(HSMU_ERR_INDEX, "column index of A outside [0, rows of B)");
    if (err & ERR_BCOL)
        return fail(HSMU_ERR_INDEX, "column index of B outside [0, N)");
    if (err & ERR_UNSORTED)
        return fail(HSMU_ERR_UNSORTED, "a row of B is not sorted by column");
    if (err & ERR_INTERNAL)
        return fail(HSMU_ERR_HIP, "internal: symbolic and numeric nnz disagree");
    return HSMU_OK;
}

// Allocate one output array: pooled or plain hipMalloc.
void *out_alloc(hsmu_ctx *c, size_t bytes, bool plain)
{
    if (plain) {
        void *p = nullptr;
        if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        return p;
    }
    return c->pool.alloc(bytes);
}

void out_free(hsmu_ctx *c, void *p)
{
    if (!p)
        return;
    if (c->pool.owns(p))
        c->pool.release(p);
    else
        (void)hipFree(p);
}

float elapsed(hipEvent_t a, hipEvent_t b)
{
    float t = 0;
    if (hipEventElapsedTime(&t, a, b) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return t;
}

// C with M rows and no entries.
int empty_product(hsmu_ctx *c, int64_t M, int64_t N, hsmu_csr *C, bool rp64, bool plain, bool val32, hsmu_stats *st)
{
    const size_t rpb = (size_t)(M + 1) * (rp64 ? 8 : 4);
    void *rp = out_alloc(c, rpb, plain);
    if (!rp)
        return fail(HSMU_ERR_NOMEM, "allocating C row pointers");
    HIP_TRY(launch_fill_zero(rp, rpb, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    C->M = M;
    C->N = N;
    C->nnz = 0;
    C->rowptr = rp;
    C->col = nullptr;
    C->val = nullptr;
    C->rowptr64 = rp64;
    C->val32 = val32;
    if (st)
        memset(st, 0, sizeof(*st));
    return HSMU_OK;
}

} // namespace

extern "C" int hsmu_btiles(hsmu_ctx *c, const hsmu_csr *B, hsmu_tiles *T)
{
    if (!c || !T)
        return fail(HSMU_ERR_INVALID, "hsmu_btiles: NULL argument");
    int rc = check_input(B, "B", false);
    if (rc)
        return rc;
    HIP_TRY(hipSetDevice(c->device));
    memset(T, 0, sizeof(*T));
    const int64_t K = B->M;
    Scratch ws(c->pool);
    int *tcnt = (int *)ws.get<char>(btiles_ws_bytes(B->nnz));
    void *scan_ws = ws.get<char>(scan_ws_bytes(B->nnz + 1));
    DevStats *dst = ws.get<DevStats>(1);
    int *tptr = (int *)c->pool.alloc(((size_t)K + 1) * sizeof(int));
    int *tcol = (int *)c->pool.alloc((size_t)(B->nnz > 0 ? B->nnz : 1) * sizeof(int));
    unsigned long long *tmask =
        (unsigned long long *)c->pool.alloc((size_t)(B->nnz > 0 ? B->nnz : 1) * sizeof(unsigned long long));
    if (!tcnt || !scan_ws || !dst || !tptr || !tcol || !tmask) {
        c->pool.release(tptr);
        c->pool.release(tcol);
        c->pool.release(tmask);
        return fail(HSMU_ERR_NOMEM, "hsmu_btiles: workspace");
    }
    HIP_TRY(launch_fill_zero(dst, sizeof(DevStats), c->stream));
    HIP_TRY(launch_btiles(K, B->N, B->nnz, (const int *)B->rowptr, B->col, tcnt, tptr, tcol, tmask, scan_ws, dst,
                          c->num_cus, c->stream));
    HIP_TRY(hipMemcpyAsync(c->host_stats, dst, sizeof(DevStats), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    rc = device_error_status(c->host_stats->err);
    if (rc) {
        c->pool.release(tptr);
        c->pool.release(tcol);
        c->pool.release(tmask);
        return rc;
    }
    T->M = K;
    T->ntiles = c->host_stats->ntiles_b;
    T->tptr = tptr;
    T->tcol = tcol;
    T->tmask = (uint64_t *)tmask;
    return HSMU_OK;
}

extern "C" int hsmu_tiles_free(hsmu_ctx *c, hsmu_tiles *T)
{
    if (!c || !T)
        return fail(HSMU_ERR_INVALID, "hsmu_tiles_free: NULL argument");
    HIP_TRY(hipStreamSynchronize(c->stream));
    out_free(c, T->tptr);
    out_free(c, T->tcol);
    out_free(c, T->tmask);
    memset(T, 0, sizeof(*T));
    return HSMU_OK;
}

extern "C" int hsmu_csr_free(hsmu_ctx *c, hsmu_csr *C)
{
    if (!c || !C)
        return fail(HSMU_ERR_INVALID, "hsmu_csr_free: NULL argument");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    out_free(c, C->rowptr);
    out_free(c, C->col);
    out_free(c, C->val);
    memset(C, 0, sizeof(*C));
    return HSMU_OK;
}

extern "C" int hsmu_spgemm(hsmu_ctx *c, const hsmu_csr *A, const hsmu_csr *B, hsmu_csr *C, const hsmu_opts *opts_in,
                           hsmu_stats *stats)
{
    if (!c || !C)
        return fail(HSMU_ERR_INVALID, "hsmu_spgemm: ctx or C is NULL");
    hsmu_opts opts;
    if (opts_in)
        opts = *opts_in;
    else
        hsmu_opts_default(&opts);
    if (opts.precision != HSMU_FP64 && opts.precision != HSMU_FP32_COMP)
        return fail(HSMU_ERR_INVALID, "unknown precision %d", opts.precision);
    const bool fp32 = opts.precision == HSMU_FP32_COMP;
    const double vsz = fp32 ? 4.0 : 8.0; // bytes per value
    int rc = check_input(A, "A", fp32);
    if (!rc)
        rc = check_input(B, "B", fp32);
    if (rc)
        return rc;
    if (A->N != B->M)
        return fail(HSMU_ERR_INVALID, "A is %lld x %lld but B has %lld rows", (long long)A->M, (long long)A->N,
                    (long long)B->M);
    HIP_TRY(hipSetDevice(c->device));
    memset(C, 0, sizeof(*C));
    const bool rp64 = opts.rowptr64 != 0;
    const bool plain = opts.plain_alloc != 0;
    const int64_t M = A->M, K = B->M, N = B->N;
    if (stats)
        memset(stats, 0, sizeof(*stats));
    if (M == 0 || A->nnz == 0 || B->nnz == 0)
        return empty_product(c, M, N, C, rp64, plain, fp32, stats);

    hipStream_t s = c->stream;
    const bool timing = opts.timing != 0;
    Scratch ws(c->pool);
    DevStats *dst = ws.get<DevStats>(1);
    int *tcnt = (int *)ws.get<char>(btiles_ws_bytes(B->nnz));
    int *tptr = ws.get<int>((size_t)K + 1);
    int *tcol = ws.get<int>((size_t)B->nnz);
    unsigned long long *tmask = ws.get<unsigned long long>((size_t)B->nnz);
    int64_t nmax = M > K ? M : K;
    nmax = nmax > B->nnz ? nmax : B->nnz;
    void *scan_ws = ws.get<char>(scan_ws_bytes(nmax + 1));
    long long *rprod = ws.get<long long>((size_t)M);
    int *rtp = ws.get<int>((size_t)M);
    int *rmin = ws.get<int>((size_t)M);
    int *rmax = ws.get<int>((size_t)M);
    int *rnnz = ws.get<int>((size_t)M);
    int *binrows = ws.get<int>((size_t)NBINS * (size_t)M);
    long long *rdir = ws.get<long long>((size_t)M);
    if (!dst || !tcnt || !tptr || !tcol || !tmask || !scan_ws || !rprod || !rtp || !rmin || !rmax || !rnnz || !binrows ||
        !rdir)
        return fail(HSMU_ERR_NOMEM, "workspace allocation");

    // ---- 1. B -> tiles ----
    HIP_TRY(hipEventRecord(c->ev[EV_START], s));
    HIP_TRY(launch_fill_zero(dst, sizeof(DevStats), s));
    HIP_TRY(launch_btiles(K, N, B->nnz, (const int *)B->rowptr, B->col, tcnt, tptr, tcol, tmask, scan_ws, dst,
                          c->num_cus, s));
    HIP_TRY(hipEventRecord(c->ev[EV_CONV], s));

    // ---- 2. row analysis + symbolic binning ----
    const int avg_a = (int)(A->nnz / (M > 0 ? M : 1));
    // the spare bin region (bin 9) holds the list of long rows until binning runs
    HIP_TRY(launch_row_info(M, K, A->nnz, (const int *)A->rowptr, A->col, (const int *)B->rowptr, tptr, tcol, rprod,
                            rtp, rmin, rmax, rnnz, binrows + (size_t)(NBINS - 1) * M, dst, avg_a, c->num_cus, s));
    BinParams bp;
    const int lds_dyn = c->lds_max - LDS_STATIC_RESERVE;
    bp.lds_row_max = opts.lds_row_max > 0 ? opts.lds_row_max : lds_dyn;
    if (bp.lds_row_max > lds_dyn)
        bp.lds_row_max = lds_dyn;
    bp.lds_block_budget = opts.lds_block_budget > 0 ? opts.lds_block_budget : 64 * 1024;
    if (bp.lds_block_budget > lds_dyn)
        bp.lds_block_budget = lds_dyn;
    bp.per_lane = opts.tiles_per_lane > 0 ? opts.tiles_per_lane : 4;
    bp.force_global = opts.force_global;
    bp.pad = 0;
    HIP_TRY(launch_bin(0, M, (const int *)A->rowptr, rprod, rtp, rmin, rmax, rnnz, bp, binrows, rdir, dst, s));
    HIP_TRY(hipEventRecord(c->ev[EV_ANALYSIS], s));
    HIP_TRY(hipMemcpyAsync(c->host_stats, dst, sizeof(DevStats), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const DevStats h1 = *c->host_stats;
    rc = device_error_status(h1.err);
    if (rc)
        return rc;

    // lanes per A entry: walk one B row (numeric) / one B tile row (tile OR)
    const long long avg_b = B->nnz / (K > 0 ? K : 1);
    const long long avg_t = (long long)h1.ntiles_b / (K > 0 ? K : 1);
    // about four products per lane per B row: shorter sub-groups keep more A entries in
    // flight per wave and idle fewer lanes on ragged B rows (measured on the banded
    // workload: 16 lanes for 64-entry rows beats 64 lanes by 1.2x in the numeric phase)
    int blanes = opts.b_lanes > 0 ? opts.b_lanes : pow2_floor(avg_b / 4 > 0 ? avg_b / 4 : 1);
    int tlanes = opts.t_lanes > 0 ? opts.t_lanes : pow2_floor(avg_t > 0 ? avg_t : 1);
    if (blanes > 64)
        blanes = 64;
    if (tlanes > 64)
        tlanes = 64;
    const int bshift = ilog2(blanes), tshift = ilog2(tlanes);

    // global workspace for rows that do not fit LDS
    int max_span = 1;
    for (int b = 0; b < NBINS; ++b)
        max_span = h1.sym.max_span[b] > max_span ? h1.sym.max_span[b] : max_span;
    const long long gws_tiles = max_span;
    const long long gws_stride = (gws_tiles + (gws_tiles + 1) / 2 + 31) & ~31LL; // u64 units, 256 B aligned
    int gws_blocks = c->num_cus * 2;
    unsigned long long *gws = nullptr;
    auto need_gws = [&](int count) -> int {
        if (gws)
            return HSMU_OK;
        int blocks = count < gws_blocks ? count : gws_blocks;
        if (blocks < 1)
            blocks = 1;
        gws_blocks = blocks;
        gws = ws.get<unsigned long long>((size_t)gws_stride * (size_t)blocks);
        return gws ? HSMU_OK : fail(HSMU_ERR_NOMEM, "global accumulator workspace");
    };

    Operands op;
    op.Arp = (const int *)A->rowptr;
    op.Acol = A->col;
    op.Aval = A->val;
    op.Brp = (const int *)B->rowptr;
    op.Bcol = B->col;
    op.Bval = B->val;
    op.val32 = fp32 ? 1 : 0;
    op.tptr = tptr;
    op.tcol = tcol;
    op.tmask = tmask;
    op.rmin = rmin;
    op.rmax = rmax;
    // Row directories (tile masks + slot offsets) of LDS-bin rows are kept from the
    // symbolic to the numeric phase when they fit the budget (12 bytes per spanned tile:
    // 15 MB on the banded workload); otherwise the numeric kernel rebuilds them.
    op.rdir = rdir;
    op.dir_mask = nullptr;
    op.dir_off = nullptr;
    const unsigned long long dir_total = h1.dir_total;
    const unsigned long long dir_budget = 1ull << 31;
    if (!opts.rebuild_dirs && dir_total > 0 && dir_total * 12ull <= dir_budget) {
        op.dir_mask = ws.get<unsigned long long>((size_t)dir_total);
        op.dir_off = ws.get<unsigned>((size_t)dir_total);
        if (!op.dir_mask || !op.dir_off) {
            op.dir_mask = nullptr;
            op.dir_off = nullptr;
        }
    }

    // ---- 3. symbolic ----
    for (int b = 0; b < NBINS; ++b) {
        const int cnt = h1.sym.count[b];
        if (cnt <= 0)
            continue;
        if (b == BIN_GLOBAL) {
            rc = need_gws(cnt);
            if (rc)
                return rc;
        }
        HIP_TRY(launch_symbolic(b, binrows + (size_t)b * M, &dst->sym.count[b], cnt, h1.sym.max_need[b], op, rnnz,
                                tshift, gws, gws_stride, c->num_cus, gws_blocks, s));
    }
    // C row pointers (int64); they become C's rowptr when int64 output is asked for
    long long *crp = (long long *)out_alloc(c, ((size_t)M + 1) * sizeof(long long), plain && rp64);
    if (!crp)
        return fail(HSMU_ERR_NOMEM, "C row pointers");
    struct CrpGuard {
        hsmu_ctx *c;
        void *p;
        ~CrpGuard()
        {
            if (p)
                out_free(c, p);
        }
    } crp_guard{c, crp};
    HIP_TRY(launch_scan_i32_i64(rnnz, M, crp, scan_ws, &dst->nnz_c, s));
    bp.per_lane = opts.products_per_lane > 0 ? opts.products_per_lane : 32;
    HIP_TRY(launch_bin(1, M, (const int *)A->rowptr, rprod, rtp, rmin, rmax, rnnz, bp, binrows, rdir, dst, s));
    HIP_TRY(hipEventRecord(c->ev[EV_SYMBOLIC], s));
    HIP_TRY(hipMemcpyAsync(c->host_stats, dst, sizeof(DevStats), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const DevStats h2 = *c->host_stats;
    rc = device_error_status(h2.err);
    if (rc)
        return rc;
    const long long nnzc = h2.nnz_c;
    if (!rp64 && nnzc > INT32_MAX)
        return fail(HSMU_ERR_OVERFLOW, "nnz(C) = %lld needs int64 row pointers (opts.rowptr64)", nnzc);

    // ---- 4. allocate C ----
    const auto t_alloc0 = std::chrono::steady_clock::now();
    int *Ccol = (int *)out_alloc(c, (size_t)nnzc * sizeof(int), plain);
    void *Cval = out_alloc(c, (size_t)nnzc * (size_t)vsz, plain);
    void *Crp = rp64 ? (void *)crp : out_alloc(c, ((size_t)M + 1) * sizeof(int), plain);
    if (!Ccol || !Cval || !Crp) {
        out_free(c, Ccol);
        out_free(c, Cval);
        if (!rp64)
            out_free(c, Crp);
        return fail(HSMU_ERR_NOMEM, "allocating C (nnz %lld)", nnzc);
    }
    const double ms_alloc_host =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_alloc0).count();
    HIP_TRY(hipEventRecord(c->ev[EV_ALLOC], s));

    // ---- 5. numeric ----
    int *errp = &dst->err;
    for (int b = 0; b < NBINS; ++b) {
        const int cnt = h2.num.count[b];
        if (cnt <= 0)
            continue;
        if (b == BIN_GLOBAL) {
            rc = need_gws(cnt);
            if (rc)
                return rc;
        }
        // compensation terms of FP32C rows accumulated in global memory
        float *comp = nullptr;
        if (fp32 && b == BIN_GLOBAL) {
            comp = ws.get<float>((size_t)nnzc);
            if (!comp)
                return fail(HSMU_ERR_NOMEM, "compensation scratch (nnz %lld)", nnzc);
        }
        if (timing)
            HIP_TRY(hipEventRecord(c->evb[b][0], s));
        HIP_TRY(launch_numeric(b, fp32 ? 1 : 0, binrows + (size_t)b * M, &dst->num.count[b], cnt, h2.num.max_need[b], op,
                               crp, Ccol, Cval, comp, bshift, tshift, gws, gws_stride, gws_tiles, gws_blocks, errp, s));
        if (timing)
            HIP_TRY(hipEventRecord(c->evb[b][1], s));
    }
    if (!rp64)
        HIP_TRY(launch_rowptr_i64_to_i32(crp, M + 1, (int *)Crp, s));
    HIP_TRY(hipEventRecord(c->ev[EV_NUMERIC], s));
    HIP_TRY(hipMemcpyAsync(&c->host_stats->err, errp, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    rc = device_error_status(c->host_stats->err);
    if (rc) {
        out_free(c, Ccol);
        out_free(c, Cval);
        if (!rp64)
            out_free(c, Crp);
        return rc;
    }
    if (rp64)
        crp_guard.p = nullptr; // crp is now C's row pointer array

    C->M = M;
    C->N = N;
    C->nnz = nnzc;
    C->rowptr = Crp;
    C->col = Ccol;
    C->val = Cval;
    C->rowptr64 = rp64;
    C->val32 = fp32 ? 1 : 0;

    if (stats) {
        hsmu_stats &S = *stats;
        S.ms_conversion = elapsed(c->ev[EV_START], c->ev[EV_CONV]);
        S.ms_form_tileCptr = elapsed(c->ev[EV_CONV], c->ev[EV_ANALYSIS]);
        S.ms_tileOR = elapsed(c->ev[EV_ANALYSIS], c->ev[EV_SYMBOLIC]);
        S.ms_malloc = elapsed(c->ev[EV_SYMBOLIC], c->ev[EV_ALLOC]);
        if (S.ms_malloc < ms_alloc_host)
            S.ms_malloc = ms_alloc_host;
        S.ms_get_Ccol = 0.0;
        S.ms_form_cval = elapsed(c->ev[EV_ALLOC], c->ev[EV_NUMERIC]);
        S.ms_time2 = elapsed(c->ev[EV_CONV], c->ev[EV_NUMERIC]);
        S.ms_total = elapsed(c->ev[EV_START], c->ev[EV_NUMERIC]);
        for (int b = 0; b < NBINS; ++b)
            S.ms_numeric_bin[b] = (timing && h2.num.count[b] > 0) ? elapsed(c->evb[b][0], c->evb[b][1]) : 0.0;
        S.products = 0;
        S.tile_products = 0;
        S.rows_nonempty = 0;
        for (int b = 0; b < NBINS; ++b) {
            S.products += (int64_t)h2.num.work[b];
            S.tile_products += (int64_t)h2.sym.work[b];
            S.rows_nonempty += h2.sym.count[b];
        }
        S.nnz_c = nnzc;
        S.tiles_b = h2.ntiles_b;
        for (int b = 0; b < NBINS; ++b) {
            S.sym_bin_rows[b] = h2.sym.count[b];
            S.num_bin_rows[b] = h2.num.count[b];
            // gather-model bytes of the rows of this numeric bin: A row extent, A col+val,
            // B row extent per a_ij, one B (col, val) per product, C row pointer, C col+val
            S.num_bin_bytes[b] = 4.0 * 2.0 * h2.num.count[b] + (4.0 + vsz) * (double)h2.num.nnz_a[b] +
                                 8.0 * (double)h2.num.nnz_a[b] + (4.0 + vsz) * (double)h2.num.work[b] +
                                 (rp64 ? 8.0 : 4.0) * h2.num.count[b] + (4.0 + vsz) * (double)h2.num.nnz_c[b];
            S.num_bin_products[b] = (int64_t)h2.num.work[b];
        }
        S.compress_rate = nnzc > 0 ? (double)S.products / (double)nnzc : 0.0;
        S.gflops = S.ms_time2 > 0 ? 2.0 * (double)S.products / (S.ms_time2 * 1e6) : 0.0;
        S.space_compressed_mask_mb =
            ((double)(K + 1) * 4.0 + (double)h2.ntiles_b * (4.0 + 8.0)) / (1024.0 * 1024.0);
        S.alg_bytes = 4.0 * (double)(M + 1) + (12.0 + vsz) * (double)A->nnz + (4.0 + vsz) * (double)S.products +
                      (rp64 ? 8.0 : 4.0) * (double)(M + 1) + (4.0 + vsz) * (double)nnzc;
    }
    return HSMU_OK;
}
