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
        size_t gran = bytes < (size_t(1) << 20) ? 256 : (size_t(2) << 20);
        // Blocks of a GiB or more come in eighth-octave sizes: C of a streamed sequence of
        // row panels grows by fractions of a percent from panel to panel, and exact sizes
        // made each slightly larger panel allocate a new set of tens of GB, keeping the
        // old ones cached.  (Such allocations stalled for seconds once device memory had
        // been freed in bulk.)  At most 12.5% of a block is unused.
        if (bytes >= (size_t(1) << 30))
            gran = size_t(1) << (63 - __builtin_clzll((unsigned long long)bytes) - 3);
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

    // Free every block, cached or still handed out (results the caller did not release).
    void release_all()
    {
        trim();
        for (auto &kv : used_) {
            (void)hipFree(kv.first);
            held_ -= kv.second;
        }
        used_.clear();
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
    size_t total_mem = 0; // device memory (bytes)
    int num_xcds = 0;     // XCDs (hipDeviceAttributeNumberOfXccs)
    void *scratch = nullptr; // device address of hsmu_spgemm_noalloc's array on this device
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
    case HSMU_ERR_PLAN: return "inputs changed structure since the plan was made";
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
    c->total_mem = prop.totalGlobalMem;
    if (hipDeviceGetAttribute(&c->num_xcds, hipDeviceAttributeNumberOfXccs, device) != hipSuccess ||
        c->num_xcds < 1) {
        (void)hipGetLastError();
        c->num_xcds = MAX_XCDS;
    }
    // every failure from here on releases what was already created
    bool ok = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess &&
              hipHostMalloc((void **)&c->host_stats, sizeof(DevStats), hipHostMallocDefault) == hipSuccess;
    for (int i = 0; ok && i < EV_COUNT; ++i)
        ok = hipEventCreate(&c->ev[i]) == hipSuccess;
    for (int b = 0; ok && b < NBINS; ++b)
        for (int k = 0; ok && k < 2; ++k)
            ok = hipEventCreate(&c->evb[b][k]) == hipSuccess;
    if (!ok) {
        const hipError_t e = hipGetLastError();
        hsmu_ctx_destroy(c);
        return fail(HSMU_ERR_HIP, "hsmu_ctx_create: stream, pinned memory or events: %s", hipGetErrorString(e));
    }
    // the allocation-free entry point's workspace on this device (looked up once per
    // device); without it only that entry point is unavailable
    if (scratch_setup(device, &c->scratch) != hipSuccess) {
        (void)hipGetLastError();
        c->scratch = nullptr;
    }
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
            if (configure_kernels(c->lds_max - LDS_STATIC_RESERVE, c->num_xcds) != hipSuccess)
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
    c->pool.release_all();
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
    HIP_TRY(launch_btiles(K, B->N, B->nnz, (const int *)B->rowptr, B->col, tcnt, tptr, tcol, tmask, nullptr, scan_ws, dst,
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

// A plan: the recorded run of one batch (or one SpGEMM) and its captured HIP graph.
struct hsmu_plan {
    hsmu_ctx *c = nullptr;
    hsmu_opts opts;
    std::vector<void *> keep;     // every device block the plan uses, in allocation order
    DevStats rec[2];              // the two checkpoints of the recording run
    DevStats *expect = nullptr;   // device copy of rec, compared by the replay
    DevStats *dst = nullptr;      // device counters of the plan's runs
    hsmu_stats stats0;            // statistics of the recording run
    bool empty = false;           // product with no entries: nothing to replay
    bool dirs_kept = false;       // the recording run kept the row directories
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

// Device memory of one pipeline run.  A one-shot call takes its workspace from the pool
// and gives it back on exit; its outputs survive the call.  A plan keeps every block,
// workspace and outputs alike, until it is destroyed, and its capture pass is handed the
// same blocks again in the same order (the sequence of requests depends only on the
// recorded checkpoints, which the capture pass reuses).
// The caller's arrays for C in an allocation-free call: row pointers of M + 1 entries,
// columns and values of cap_nnz entries each (vsz bytes per value).
struct FixedOut {
    void *rowptr;
    int *col;
    void *val;
    int64_t cap_nnz;
    size_t vsz;
};

enum OutRole { OUT_ROWPTR, OUT_COL, OUT_VAL };

class Arena {
  public:
    Arena(hsmu_ctx *c, std::vector<void *> *keep, bool replay) : c_(c), keep_(keep), replay_(replay) {}
    // Allocation-free: the workspace is carved from [base, base + size) (256-byte aligned
    // pieces, nothing is allocated or freed) and C goes into the caller's arrays.
    Arena(hsmu_ctx *c, char *base, size_t size, const FixedOut *fixed)
        : c_(c), keep_(nullptr), replay_(false), base_(base), size_(size), fixed_(fixed)
    {
    }
    ~Arena()
    {
        for (void *p : taken_)
            c_->pool.release(p);
    }
    void *raw(size_t bytes)
    {
        if (base_) {
            const size_t o = (off_ + 255) & ~(size_t)255;
            if (o > size_ || bytes > size_ - o) {
                if (!short_)
                    short_ = o + bytes;
                return nullptr;
            }
            off_ = o + (bytes ? bytes : 1);
            return base_ + o;
        }
        if (replay_)
            return pos_ < keep_->size() ? (*keep_)[pos_++] : nullptr;
        void *p = c_->pool.alloc(bytes);
        if (p)
            (keep_ ? *keep_ : taken_).push_back(p);
        return p;
    }
    template <typename T> T *get(size_t n) { return (T *)raw(n * sizeof(T)); }
    // bytes a further workspace request may take (unbounded unless allocation-free)
    size_t remaining() const { return base_ ? (off_ < size_ ? size_ - off_ : 0) : SIZE_MAX; }
    bool fixed() const { return fixed_ != nullptr; }
    // an output array of a one-shot call (plain hipMalloc or the pool, not given back at
    // exit); a plan's outputs are kept blocks like the rest; an allocation-free call's are
    // the caller's arrays (nullptr when C's entries exceed their capacity)
    void *out(size_t bytes, bool plain, OutRole role)
    {
        if (fixed_) {
            if (role == OUT_ROWPTR)
                return fixed_->rowptr;
            const size_t esz = role == OUT_COL ? sizeof(int) : fixed_->vsz;
            if (bytes > (size_t)fixed_->cap_nnz * esz) {
                want_nnz_ = (int64_t)(bytes / esz);
                return nullptr;
            }
            void *p = role == OUT_COL ? (void *)fixed_->col : fixed_->val;
            return p ? p : base_; // nnz(C) = 0 with no arrays: any valid address, never written
        }
        return keep_ ? raw(bytes) : out_alloc(c_, bytes, plain);
    }
    void drop_out(void *p)
    {
        if (!keep_ && !fixed_)
            out_free(c_, p);
    }
    int64_t want_nnz() const { return want_nnz_; }    // nnz(C) that exceeded the capacity
    size_t short_bytes() const { return short_; }     // workspace end a failed request needed

  private:
    hsmu_ctx *c_;
    std::vector<void *> *keep_;
    bool replay_;
    size_t pos_ = 0;
    std::vector<void *> taken_;
    char *base_ = nullptr;
    size_t size_ = 0, off_ = 0, short_ = 0;
    const FixedOut *fixed_ = nullptr;
    int64_t want_nnz_ = 0;
};

// How a pipeline run meets its two checkpoints: eagerly (copy the device counters to
// the host and synchronise; a plan's recording run keeps them), or, while a plan is
// being captured, with no synchronisation at all: the recorded counters stand in for the
// device's, and a check kernel compares the two when the graph is replayed.
struct RunMode {
    hsmu_plan *plan = nullptr;
    bool capture = false;
};

int checkpoint(hsmu_ctx *c, const RunMode &m, DevStats *dst, int k, DevStats &out)
{
    hipStream_t s = c->stream;
    if (m.capture) {
        HIP_TRY(launch_plan_check(dst, m.plan->expect + k, k, s));
        out = m.plan->rec[k];
        return HSMU_OK;
    }
    HIP_TRY(hipMemcpyAsync(c->host_stats, dst, sizeof(DevStats), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    out = *c->host_stats;
    if (m.plan)
        m.plan->rec[k] = out;
    return device_error_status(out.err);
}

// The operands of one pipeline run (device arrays, int32 row pointers): one SpGEMM, or
// the stacked instances of a batch.
struct Problem {
    int64_t M, K, N, nnzA, nnzB;
    const int *Arp, *Acol;
    const void *Aval;
    const int *Brp, *Bcol;
    const void *Bval;
};

// Set for a batch: the output row pointers are split per instance instead of converted.
struct BatchSplit {
    int64_t count;
    BatchOffsets off; // device
    std::vector<long long> *bases; // host: where each C_i's entries start (eager runs)
    long long *bases_dev = nullptr; // the same on the device (set by the pipeline)
};

struct Output {
    int64_t nnzc = 0;
    void *rowptr = nullptr;
    int *col = nullptr;
    void *val = nullptr;
};

// B -> tiles, row analysis, symbolic, C allocation, numeric.  *dst must be zeroed on the
// stream before the call (the batch stacking kernels report into it first).
int pipeline(hsmu_ctx *c, const Problem &P, const hsmu_opts &opts, Arena &ar, const RunMode &mode, DevStats *dst,
             BatchSplit *batch, Output &res, hsmu_stats *stats)
{
    const bool fp32 = opts.precision == HSMU_FP32_COMP;
    const double vsz = fp32 ? 4.0 : 8.0; // bytes per value
    const bool cap = mode.capture;
    const bool rp64 = opts.rowptr64 != 0;
    const bool plain = opts.plain_alloc != 0;
    const int64_t M = P.M, K = P.K, N = P.N;
    hipStream_t s = c->stream;
    // hipEvents cost the device a few microseconds each: numeric-bin events for timing 1
    // and 2, stage events for timing 1 only
    const bool timing = opts.timing != 0 && !cap;
    const bool stage_timing = opts.timing == 1 && !cap;
    auto mark = [&](hipEvent_t e) -> hipError_t { return stage_timing ? hipEventRecord(e, s) : hipSuccess; };
    int rc;

    int *tcnt = (int *)ar.get<char>(btiles_ws_bytes(P.nnzB));
    int *tptr = ar.get<int>((size_t)K + 1);
    int *tcol = ar.get<int>((size_t)P.nnzB);
    unsigned long long *tmask = ar.get<unsigned long long>((size_t)P.nnzB);
    int64_t nmax = M > K ? M : K;
    nmax = nmax > P.nnzB ? nmax : P.nnzB;
    void *scan_ws = ar.get<char>(scan_ws_bytes(nmax + 1));
    long long *rprod = ar.get<long long>((size_t)M);
    int *rtp = ar.get<int>((size_t)M);
    int *rmin = ar.get<int>((size_t)M);
    int *rmax = ar.get<int>((size_t)M);
    int *rnnz = ar.get<int>((size_t)M);
    int *binrows = ar.get<int>((size_t)NBINS * (size_t)M);
    long long *rdir = ar.get<long long>((size_t)M);
    BRow *brow = ar.get<BRow>((size_t)(K > 0 ? K : 1));
    if (!tcnt || !tptr || !tcol || !tmask || !scan_ws || !rprod || !rtp || !rmin || !rmax || !rnnz || !binrows || !rdir ||
        !brow)
        return fail(HSMU_ERR_NOMEM, "workspace allocation");

    // ---- 1. B -> tiles ----
    HIP_TRY(launch_btiles(K, N, P.nnzB, P.Brp, P.Bcol, tcnt, tptr, tcol, tmask, brow, scan_ws, dst, c->num_cus, s));
    HIP_TRY(mark(c->ev[EV_CONV]));

    // ---- 2. row analysis + symbolic binning ----
    const int avg_a = (int)(P.nnzA / (M > 0 ? M : 1));
    // the spare bin region (bin 9) holds the list of long rows until binning runs
    HIP_TRY(launch_row_info(M, K, P.nnzA, P.Arp, P.Acol, brow, rprod, rtp, rmin, rmax, rnnz,
                            binrows + (size_t)(NBINS - 1) * M, dst, avg_a, c->num_cus, s));
    BinParams bp;
    const int lds_dyn = c->lds_max - LDS_STATIC_RESERVE;
    bp.lds_row_max = opts.lds_row_max > 0 ? opts.lds_row_max : lds_dyn;
    if (bp.lds_row_max > lds_dyn)
        bp.lds_row_max = lds_dyn;
    // 8 KiB per one-wave workgroup that packs several rows: 20 such workgroups fit a CU's
    // 160 KiB; rows needing more move up to a bin with fewer rows per workgroup (measured
    // best on the batched Laplacians)
    bp.lds_block_budget = opts.lds_block_budget > 0 ? opts.lds_block_budget : 8 * 1024;
    if (bp.lds_block_budget > lds_dyn)
        bp.lds_block_budget = lds_dyn;
    bp.per_lane = opts.tiles_per_lane > 0 ? opts.tiles_per_lane : 4;
    const long long bp_sym_row_max = bp.lds_row_max;
    bp.force_global = opts.force_global;
    bp.lds_cu = lds_dyn;
    bp.min_waves = opts.min_waves > 0 ? opts.min_waves : 8;
    HIP_TRY(launch_bin(0, M, P.Arp, rprod, rtp, rmin, rmax, rnnz, bp, binrows, rdir, dst, c->num_cus, s));
    HIP_TRY(mark(c->ev[EV_ANALYSIS]));
    DevStats h1;
    rc = checkpoint(c, mode, dst, 0, h1);
    if (rc)
        return rc;

    // lanes per A entry: walk one B row (numeric) / one B tile row (tile OR)
    const long long avg_b = P.nnzB / (K > 0 ? K : 1);
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
    // B has rows long enough for the whole-wave walks (the binning sized LDS for them)
    const int long_rows = h1.max_brow >= LONG_TILES ? 1 : 0;

    Operands op;
    op.Arp = P.Arp;
    op.Acol = P.Acol;
    op.Aval = P.Aval;
    op.Brp = P.Brp;
    op.Bcol = P.Bcol;
    op.Bval = P.Bval;
    op.val32 = fp32 ? 1 : 0;
    op.tptr = tptr;
    op.tcol = tcol;
    op.tmask = tmask;
    op.rmin = rmin;
    op.rmax = rmax;
    op.rtp = rtp;
    // Row directories (tile masks + slot offsets) are kept from the symbolic to the
    // numeric phase (12 bytes per directory entry) when that is cheaper than rebuilding
    // them -- the directories hold at most half as many entries as the tile-products that
    // rebuild them (banded: 1.3M entries for 10M tile-products; the batched Laplacians:
    // about as many, and rebuilding is 0.6 ms faster there) -- or when wide rows need them
    // for the compact windows, and they fit a quarter of the device memory (72 GB on an
    // MI355X; R-MAT scale 20 needs 43 GB).  Otherwise the numeric LDS kernels rebuild them and wide rows take the dense
    // window kernel.  A plan's capture pass repeats the recorded decision.
    op.rdir = rdir;
    op.dir_mask = nullptr;
    op.dir_off = nullptr;
    const unsigned long long dir_total = h1.dir_total;
    bool keep_dirs;
    if (cap) {
        keep_dirs = mode.plan->dirs_kept;
    } else {
        unsigned long long tile_products = 0;
        for (int b = 0; b < NBINS; ++b)
            tile_products += h1.sym.work[b];
        const bool worth = 2 * dir_total <= tile_products || h1.sym.count[BIN_GLOBAL] > 0 ||
                           h1.num_wide_hint > 0;
        unsigned long long budget = c->total_mem / 4 > (1ull << 31) ? c->total_mem / 4 : (1ull << 31);
        // allocation-free: at most half of the workspace left, the rest serves C's row
        // pointers and the numeric phase
        if (budget > ar.remaining() / 2)
            budget = ar.remaining() / 2;
        keep_dirs = !opts.rebuild_dirs && worth && dir_total > 0 && dir_total * 12ull <= budget;
        if (mode.plan)
            mode.plan->dirs_kept = keep_dirs;
    }
    if (keep_dirs) {
        op.dir_mask = ar.get<unsigned long long>((size_t)dir_total);
        op.dir_off = ar.get<unsigned>((size_t)dir_total);
        if (!op.dir_mask || !op.dir_off)
            return fail(HSMU_ERR_NOMEM, "row directories (%llu entries)", dir_total);
    }
    // tile-mask workspace of the symbolic phase for rows that do not fit LDS when the
    // directories are not kept: one span of masks per workgroup (widest span in the bin).
    // Not needed when that span's directory fits the LDS a row may use (k_symbolic_wide).
    const int gws_tiles = h1.sym.max_span[BIN_GLOBAL] > 0 ? h1.sym.max_span[BIN_GLOBAL] : 1;
    const long long gws_stride = ((long long)gws_tiles + 31) & ~31LL; // u64 units, 256 B aligned
    int gws_blocks = 0;
    unsigned long long *gws = nullptr;
    if (h1.sym.count[BIN_GLOBAL] > 0) {
        gws_blocks = h1.sym.count[BIN_GLOBAL] < 2 * c->num_cus ? h1.sym.count[BIN_GLOBAL] : 2 * c->num_cus;
        if (!keep_dirs && (unsigned long long)gws_stride * 8ull > (unsigned long long)bp.lds_row_max) {
            gws = ar.get<unsigned long long>((size_t)gws_stride * (size_t)gws_blocks);
            if (!gws)
                return fail(HSMU_ERR_NOMEM, "symbolic tile-mask workspace");
        }
    }

    // ---- 3. symbolic ----
    for (int b = 0; b < NBINS; ++b) {
        const int cnt = h1.sym.count[b];
        if (cnt <= 0)
            continue;
        HIP_TRY(launch_symbolic(b, binrows + (size_t)b * M, &dst->sym.count[b], cnt, h1.sym.max_need[b], op, rnnz,
                                tshift, h1.sym.nsparse[b] > 0 ? 1 : 0, long_rows, gws, gws_stride, c->num_cus,
                                gws_blocks, (int)bp.lds_row_max,
                                opts.persistent ? dst->queue[0][b] : nullptr, s));
    }
    // Arrays of C handed to the caller; dropped again (one-shot calls) on any failure.
    struct Outs {
        Arena &ar;
        std::vector<void *> p;
        ~Outs()
        {
            for (void *q : p)
                ar.drop_out(q);
        }
    } outs{ar, {}};
    // C row pointers (int64); they are C's rowptr when a single product asks for int64
    long long *crp;
    if (rp64 && !batch) {
        crp = (long long *)ar.out(((size_t)M + 1) * sizeof(long long), plain, OUT_ROWPTR);
        if (crp)
            outs.p.push_back(crp);
    } else {
        crp = ar.get<long long>((size_t)M + 1);
    }
    if (!crp)
        return fail(HSMU_ERR_NOMEM, "C row pointers");
    HIP_TRY(launch_scan_i32_i64(rnnz, M, crp, scan_ws, &dst->nnz_c, s));
    bp.per_lane = opts.products_per_lane > 0 ? opts.products_per_lane : 32;
    HIP_TRY(launch_bin(1, M, P.Arp, rprod, rtp, rmin, rmax, rnnz, bp, binrows, rdir, dst, c->num_cus, s));
    HIP_TRY(mark(c->ev[EV_SYMBOLIC]));
    DevStats h2;
    rc = checkpoint(c, mode, dst, 1, h2);
    if (rc)
        return rc;
    const long long nnzc = h2.nnz_c;
    if (!rp64 && nnzc > INT32_MAX)
        return fail(HSMU_ERR_OVERFLOW, "nnz(C) = %lld needs int64 row pointers (opts.rowptr64)", nnzc);

    // ---- 4. allocate C ----
    const auto t_alloc0 = std::chrono::steady_clock::now();
    int *Ccol = (int *)ar.out((size_t)nnzc * sizeof(int), plain, OUT_COL);
    outs.p.push_back(Ccol);
    void *Cval = ar.out((size_t)nnzc * (size_t)vsz, plain, OUT_VAL);
    outs.p.push_back(Cval);
    void *Crp;
    long long *bases = nullptr;
    if (batch) {
        Crp = ar.out((size_t)(M + batch->count) * (rp64 ? 8 : 4), plain, OUT_ROWPTR);
        outs.p.push_back(Crp);
        bases = ar.get<long long>((size_t)batch->count + 1);
        batch->bases_dev = bases;
    } else if (rp64) {
        Crp = crp;
    } else {
        Crp = ar.out(((size_t)M + 1) * sizeof(int), plain, OUT_ROWPTR);
        outs.p.push_back(Crp);
    }
    if (!Ccol || !Cval || !Crp || (batch && !bases))
        return fail(HSMU_ERR_NOMEM, "allocating C (nnz %lld)", nnzc);
    const double ms_alloc_host =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_alloc0).count();
    HIP_TRY(mark(c->ev[EV_ALLOC]));

    // ---- 5. numeric ----
    int *errp = &dst->err;
    unsigned long long lds_launched = 0;
    for (int b = 0; b < NBINS; ++b) {
        const int cnt = h2.num.count[b];
        if (cnt <= 0)
            continue;
        // wide rows: a workspace per window workgroup (two 1024-lane workgroups fit a CU's
        // LDS) for the tiles of a two-level directory and, without split items, the cursors
        // of its row's A entries.  Rows with kept directories (FP64, and the compensated form
        // when its scratch fits) run as split work items (at most count + nnz(A) /
        // WS_MIN_SPLIT of them) on a full persistent grid.
        WideRows wide = {};
        if (b == BIN_GLOBAL) {
            // the compensated form's split rows need an FP64 sum per nonzero of theirs
            const unsigned long long scratch_bytes = h2.wide_split_nnz * sizeof(unsigned long long);
            unsigned long long scratch_max =
                opts.wide_scratch_mb > 0 ? (unsigned long long)opts.wide_scratch_mb << 20 : c->total_mem / 4;
            if (scratch_max > ar.remaining() / 2) // allocation-free: what the workspace has left
                scratch_max = ar.remaining() / 2;
            const bool split_items = keep_dirs && (!fp32 || scratch_bytes <= scratch_max);
            // split items: the window size asked for (the default one otherwise), and
            // narrow workgroups for sparse rows (few columns per tile spanned)
            wide.vals = split_items && opts.wide_window ? opts.wide_window : WS_VALS;
            wide.sparse =
                (unsigned long long)h2.num.nnz_c[b] < (unsigned long long)WS_SPARSE_DENSITY * h2.wide_span ? 1 : 0;
            const int per_cu = wide.vals == WS_VALS ? 2 : 1; // workgroups of the window kernel a CU holds
            wide.blocks = split_items ? per_cu * c->num_cus : (cnt > 2 * c->num_cus ? 2 * c->num_cus : cnt);
            wide.cur_stride = split_items ? 0 : ((long long)h2.num.max_na[b] + 63) & ~63LL;
            wide.ft_stride = keep_dirs ? (((long long)h2.num.max_span[b] + 63) & ~63LL) : 0;
            wide.sym_row_max = bp_sym_row_max;
            wide.sym_all_global = opts.force_global ? 1 : 0;
            wide.long_rows = long_rows;
            wide.ws = ar.get<int>(wide_ws_ints(wide.cur_stride, wide.ft_stride, wide.blocks) + 1);
            if (split_items) {
                wide.items_cap = (long long)cnt + P.nnzA / WS_MIN_SPLIT + 1;
                wide.items = ar.get<WideItem>((size_t)wide.items_cap);
                wide.ctr = ar.get<unsigned>(WIDE_CTRS);
                if (fp32) {
                    wide.scratch_n = (long long)h2.wide_split_nnz;
                    wide.scratch = ar.get<unsigned long long>((size_t)(wide.scratch_n > 0 ? wide.scratch_n : 1));
                    wide.sctr = ar.get<unsigned long long>(1);
                }
            }
            if (!wide.ws || (split_items && (!wide.items || !wide.ctr || (fp32 && (!wide.scratch || !wide.sctr)))))
                return fail(HSMU_ERR_NOMEM, "wide-row workspace");
        }
        if (timing)
            HIP_TRY(hipEventRecord(c->evb[b][0], s));
        HIP_TRY(launch_numeric(b, fp32 ? 1 : 0, binrows + (size_t)b * M, &dst->num.count[b], cnt, h2.num.max_need[b], op,
                               crp, Ccol, Cval, bshift, tshift, h2.num.nsparse[b] > 0 ? 1 : 0, long_rows, wide, errp,
                               c->num_cus,
                               opts.persistent ? dst->queue[1][b] : nullptr, &lds_launched, s));
        if (timing)
            HIP_TRY(hipEventRecord(c->evb[b][1], s));
    }
    if (batch)
        HIP_TRY(launch_batch_split(batch->count, batch->off, crp, Crp, rp64 ? 1 : 0, bases, M + batch->count, s));
    else if (!rp64)
        HIP_TRY(launch_rowptr_i64_to_i32(crp, M + 1, (int *)Crp, s));
    HIP_TRY(mark(c->ev[EV_NUMERIC]));
    if (!cap) {
        HIP_TRY(hipMemcpyAsync(&c->host_stats->err, errp, sizeof(int), hipMemcpyDeviceToHost, s));
        if (batch) {
            batch->bases->assign((size_t)batch->count + 1, 0);
            HIP_TRY(hipMemcpyAsync(batch->bases->data(), bases, ((size_t)batch->count + 1) * sizeof(long long),
                                   hipMemcpyDeviceToHost, s));
        }
        HIP_TRY(hipStreamSynchronize(s));
        rc = device_error_status(c->host_stats->err);
        if (rc)
            return rc;
    }
    outs.p.clear(); // the arrays now belong to the result

    res.nnzc = nnzc;
    res.rowptr = Crp;
    res.col = Ccol;
    res.val = Cval;
    if (cap)
        return HSMU_OK;

    if (stats) {
        hsmu_stats &S = *stats;
        auto stage = [&](int from, int to) { return stage_timing ? (double)elapsed(c->ev[from], c->ev[to]) : 0.0; };
        S.ms_conversion = stage(EV_START, EV_CONV);
        S.ms_form_tileCptr = stage(EV_CONV, EV_ANALYSIS);
        S.ms_tileOR = stage(EV_ANALYSIS, EV_SYMBOLIC);
        S.ms_malloc = stage(EV_SYMBOLIC, EV_ALLOC);
        if (S.ms_malloc < ms_alloc_host)
            S.ms_malloc = ms_alloc_host;
        S.ms_get_Ccol = 0.0;
        S.ms_form_cval = stage(EV_ALLOC, EV_NUMERIC);
        S.ms_time2 = stage(EV_CONV, EV_NUMERIC);
        S.ms_total = stage(EV_START, EV_NUMERIC);
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
        S.alg_bytes = 4.0 * (double)(M + 1) + (12.0 + vsz) * (double)P.nnzA + (4.0 + vsz) * (double)S.products +
                      (rp64 ? 8.0 : 4.0) * (double)(M + 1) + (4.0 + vsz) * (double)nnzc;
        // compulsory traffic: every input array read once and C written once (A and B the
        // same arrays for a self-product A*A)
        const bool self = P.Arp == P.Brp && P.Acol == P.Bcol && P.Aval == P.Bval;
        S.compulsory_bytes = (4.0 * (double)(M + 1) + (4.0 + vsz) * (double)P.nnzA) +
                             (self ? 0.0 : 4.0 * (double)(K + 1) + (4.0 + vsz) * (double)P.nnzB) +
                             (rp64 ? 8.0 : 4.0) * (double)(M + 1) + (4.0 + vsz) * (double)nnzc;
        unsigned long long nnz_lds = 0; // nonzeros of the rows in LDS bins
        for (int b = 0; b < NLDS; ++b)
            nnz_lds += h2.num.nnz_c[b];
        S.lds_launched = (double)lds_launched;
        S.lds_use = lds_launched ? 12.0 * (double)nnz_lds / (double)lds_launched : 0.0;
    }
    return HSMU_OK;
}

hsmu_opts resolve_opts(const hsmu_opts *in)
{
    hsmu_opts o;
    if (in)
        o = *in;
    else
        hsmu_opts_default(&o);
    return o;
}

int check_opts(const hsmu_opts &o)
{
    if (o.precision != HSMU_FP64 && o.precision != HSMU_FP32_COMP)
        return fail(HSMU_ERR_INVALID, "unknown precision %d", o.precision);
    if (o.wide_window != 0 && o.wide_window != WS_VALS && o.wide_window != WS_VALS_BIG)
        return fail(HSMU_ERR_INVALID, "wide_window %d: 0, %d or %d", o.wide_window, WS_VALS, WS_VALS_BIG);
    return HSMU_OK;
}

int check_pair(const hsmu_csr *A, const hsmu_csr *B, bool fp32)
{
    int rc = check_input(A, "A", fp32);
    if (!rc)
        rc = check_input(B, "B", fp32);
    if (rc)
        return rc;
    if (A->N != B->M)
        return fail(HSMU_ERR_INVALID, "A is %lld x %lld but B has %lld rows", (long long)A->M, (long long)A->N,
                    (long long)B->M);
    return HSMU_OK;
}

Problem single_problem(const hsmu_csr *A, const hsmu_csr *B)
{
    Problem P;
    P.M = A->M;
    P.K = B->M;
    P.N = B->N;
    P.nnzA = A->nnz;
    P.nnzB = B->nnz;
    P.Arp = (const int *)A->rowptr;
    P.Acol = A->col;
    P.Aval = A->val;
    P.Brp = (const int *)B->rowptr;
    P.Bcol = B->col;
    P.Bval = B->val;
    return P;
}

void fill_result(hsmu_csr *C, const Output &r, int64_t M, int64_t N, const hsmu_opts &o)
{
    C->M = M;
    C->N = N;
    C->nnz = r.nnzc;
    C->rowptr = r.rowptr;
    C->col = r.col;
    C->val = r.val;
    C->rowptr64 = o.rowptr64 != 0;
    C->val32 = o.precision == HSMU_FP32_COMP;
}

// ---- batches -------------------------------------------------------------------------------

// Host side of a batch: the instance table, the offsets and the stacked sizes.
struct BatchPrep {
    int64_t count = 0;
    std::vector<BatchInst> inst;
    std::vector<long long> offs; // mo, ko, ea, eb: count + 1 each
    bool self = false;           // every instance is A_i * A_i on the same arrays
    int64_t M = 0, K = 0, N = 0, nnzA = 0, nnzB = 0;
};

int prepare_batch(int64_t count, const hsmu_csr *A, const hsmu_csr *B, bool fp32, BatchPrep &bp)
{
    if (count < 1)
        return fail(HSMU_ERR_INVALID, "batch count %lld", (long long)count);
    if (count >= INT32_MAX)
        return fail(HSMU_ERR_UNSUPPORTED, "batch count %lld", (long long)count);
    if (!A || !B)
        return fail(HSMU_ERR_INVALID, "batch A or B is NULL");
    bp.count = count;
    bp.inst.resize((size_t)count);
    bp.offs.assign(4 * ((size_t)count + 1), 0);
    long long *mo = bp.offs.data(), *ko = mo + count + 1, *ea = ko + count + 1, *eb = ea + count + 1;
    bool self = true;
    for (int64_t i = 0; i < count; ++i) {
        int rc = check_pair(A + i, B + i, fp32);
        if (rc)
            return fail(rc, "instance %lld: %s", (long long)i, hsmu_last_error());
        if ((A[i].M == 0 && A[i].nnz > 0) || (B[i].M == 0 && B[i].nnz > 0))
            return fail(HSMU_ERR_INVALID, "instance %lld: entries in a matrix without rows", (long long)i);
        BatchInst &I = bp.inst[(size_t)i];
        I.Arp = (const int *)A[i].rowptr;
        I.Acol = A[i].col;
        I.Aval = A[i].val;
        I.Brp = (const int *)B[i].rowptr;
        I.Bcol = B[i].col;
        I.Bval = B[i].val;
        I.M = (int)A[i].M;
        I.K = (int)B[i].M;
        I.N = (int)B[i].N;
        I.pad = 0;
        mo[i + 1] = mo[i] + A[i].M;
        ko[i + 1] = ko[i] + B[i].M;
        ea[i + 1] = ea[i] + A[i].nnz;
        eb[i + 1] = eb[i] + B[i].nnz;
        bp.N = B[i].N > bp.N ? B[i].N : bp.N;
        self = self && A[i].rowptr == B[i].rowptr && A[i].col == B[i].col && A[i].val == B[i].val &&
               A[i].M == B[i].M && A[i].nnz == B[i].nnz;
    }
    bp.M = mo[count];
    bp.K = ko[count];
    bp.nnzA = ea[count];
    bp.nnzB = eb[count];
    bp.self = self;
    if (bp.M + count >= INT32_MAX || bp.K >= INT32_MAX || bp.nnzA > INT32_MAX || bp.nnzB > INT32_MAX)
        return fail(HSMU_ERR_UNSUPPORTED, "stacked batch beyond int32 indices (rows %lld, nnz %lld / %lld)",
                    (long long)bp.M, (long long)bp.nnzA, (long long)bp.nnzB);
    return HSMU_OK;
}

// Stack the instances on the device (the table and offsets are uploaded by eager runs
// only: a capture pass reuses what the recording run uploaded into the same blocks).
int stack_inputs(hsmu_ctx *c, const BatchPrep &bp, const hsmu_opts &opts, Arena &ar, const RunMode &mode,
                 DevStats *dst, Problem &P, BatchOffsets &off)
{
    hipStream_t s = c->stream;
    const bool fp32 = opts.precision == HSMU_FP32_COMP;
    const size_t vsz = fp32 ? 4 : 8;
    const int64_t n1 = bp.count + 1;
    BatchInst *inst = ar.get<BatchInst>((size_t)bp.count);
    long long *offs = ar.get<long long>(4 * (size_t)n1);
    int *Brp = ar.get<int>((size_t)bp.K + 1);
    int *Bcol = ar.get<int>((size_t)bp.nnzB);
    void *Bval = ar.raw((size_t)bp.nnzB * vsz);
    int *Acol = ar.get<int>((size_t)bp.nnzA);
    int *Arp = bp.self ? Brp : ar.get<int>((size_t)bp.M + 1);
    void *Aval = bp.self ? Bval : ar.raw((size_t)bp.nnzA * vsz);
    if (!inst || !offs || !Brp || !Bcol || !Bval || !Acol || !Arp || !Aval)
        return fail(HSMU_ERR_NOMEM, "stacked batch inputs");
    if (!mode.capture) {
        HIP_TRY(hipMemcpyAsync(inst, bp.inst.data(), bp.inst.size() * sizeof(BatchInst), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(offs, bp.offs.data(), bp.offs.size() * sizeof(long long), hipMemcpyHostToDevice, s));
    }
    off.mo = offs;
    off.ko = offs + n1;
    off.ea = offs + 2 * n1;
    off.eb = offs + 3 * n1;
    const int v32 = fp32 ? 1 : 0;
    HIP_TRY(launch_batch_rowptr(1, bp.count, inst, off, bp.K, bp.nnzB, Brp, dst, s));
    HIP_TRY(launch_batch_entries(1, bp.count, inst, off, bp.nnzB, v32, Bcol, Bval, dst, s));
    if (!bp.self)
        HIP_TRY(launch_batch_rowptr(0, bp.count, inst, off, bp.M, bp.nnzA, Arp, dst, s));
    HIP_TRY(launch_batch_entries(0, bp.count, inst, off, bp.nnzA, v32, Acol, bp.self ? nullptr : Aval, dst, s));
    P.M = bp.M;
    P.K = bp.K;
    P.N = bp.N;
    P.nnzA = bp.nnzA;
    P.nnzB = bp.nnzB;
    P.Arp = Arp;
    P.Acol = Acol;
    P.Aval = Aval;
    P.Brp = Brp;
    P.Bcol = Bcol;
    P.Bval = Bval;
    return HSMU_OK;
}

// Views of the batch's result: C_i's row pointers at rowptr + mo_i + i, entries at bases[i].
void fill_views(const BatchPrep &bp, const hsmu_csr *B, const Output &r, const std::vector<long long> &bases,
                const hsmu_opts &o, hsmu_csr *C)
{
    const bool rp64 = o.rowptr64 != 0;
    const size_t vsz = o.precision == HSMU_FP32_COMP ? 4 : 8;
    const long long *mo = bp.offs.data();
    for (int64_t i = 0; i < bp.count; ++i) {
        hsmu_csr &X = C[i];
        X.M = bp.inst[(size_t)i].M;
        X.N = B[i].N;
        X.nnz = r.col ? bases[(size_t)i + 1] - bases[(size_t)i] : 0;
        X.rowptr = (char *)r.rowptr + (size_t)(mo[i] + i) * (rp64 ? 8 : 4);
        X.col = r.col ? r.col + bases[(size_t)i] : nullptr;
        X.val = r.col ? (char *)r.val + (size_t)bases[(size_t)i] * vsz : nullptr;
        X.rowptr64 = rp64;
        X.val32 = o.precision == HSMU_FP32_COMP;
    }
}

// Run a batch (count > 1) or a single product (count == 1) through the pipeline; with a
// plan, in its recording or capture pass.  Fills C (views for a batch).
int run_batch(hsmu_ctx *c, int64_t count, const hsmu_csr *A, const hsmu_csr *B, hsmu_csr *C, const hsmu_opts &opts,
              Arena &ar, const RunMode &mode, hsmu_stats *stats, bool *empty)
{
    const bool fp32 = opts.precision == HSMU_FP32_COMP;
    hipStream_t s = c->stream;
    *empty = false;
    BatchPrep bp;
    int rc;
    if (count == 1) {
        rc = check_pair(A, B, fp32);
        if (rc)
            return rc;
    } else {
        rc = prepare_batch(count, A, B, fp32, bp);
        if (rc)
            return rc;
    }
    const int64_t M = count == 1 ? A->M : bp.M;
    const int64_t nnzA = count == 1 ? A->nnz : bp.nnzA, nnzB = count == 1 ? B->nnz : bp.nnzB;
    if (M == 0 || nnzA == 0 || nnzB == 0) {
        // no products: C has row pointers only, all zero
        *empty = true;
        if (stats)
            memset(stats, 0, sizeof(*stats));
        if (mode.capture)
            return HSMU_OK;
        const bool rp64 = opts.rowptr64 != 0;
        const size_t rpb = (size_t)(M + count) * (rp64 ? 8 : 4);
        void *rp = ar.out(rpb, opts.plain_alloc != 0, OUT_ROWPTR);
        if (!rp)
            return fail(HSMU_ERR_NOMEM, "allocating C row pointers");
        HIP_TRY(launch_fill_zero(rp, rpb, s));
        HIP_TRY(hipStreamSynchronize(s));
        Output r;
        r.rowptr = rp;
        if (count == 1) {
            fill_result(C, r, A->M, B->N, opts);
        } else {
            std::vector<long long> bases((size_t)count + 1, 0);
            fill_views(bp, B, r, bases, opts, C);
        }
        return HSMU_OK;
    }
    DevStats *dst = ar.get<DevStats>(1);
    if (!dst)
        return fail(HSMU_ERR_NOMEM, "workspace allocation");
    if (mode.plan && !mode.capture) {
        mode.plan->dst = dst;
        mode.plan->expect = ar.get<DevStats>(2);
        if (!mode.plan->expect)
            return fail(HSMU_ERR_NOMEM, "workspace allocation");
    } else if (mode.capture) {
        ar.get<DevStats>(2); // the recorded expect block, in allocation order
    }
    if (!mode.capture && opts.timing == 1)
        HIP_TRY(hipEventRecord(c->ev[EV_START], s));
    HIP_TRY(launch_fill_zero(dst, sizeof(DevStats), s));
    Output r;
    if (count == 1) {
        rc = pipeline(c, single_problem(A, B), opts, ar, mode, dst, nullptr, r, stats);
        if (rc)
            return rc;
        if (!mode.capture)
            fill_result(C, r, A->M, B->N, opts);
        return HSMU_OK;
    }
    Problem P;
    BatchSplit split;
    std::vector<long long> bases;
    split.count = count;
    split.bases = &bases;
    rc = stack_inputs(c, bp, opts, ar, mode, dst, P, split.off);
    if (rc)
        return rc;
    rc = pipeline(c, P, opts, ar, mode, dst, &split, r, stats);
    if (rc)
        return rc;
    if (mode.plan) {
        // a replay must leave every C_i where the views point: compare the starts
        long long *want = ar.get<long long>((size_t)count + 1);
        if (!want)
            return fail(HSMU_ERR_NOMEM, "plan workspace");
        if (mode.capture) {
            HIP_TRY(launch_plan_bases(split.bases_dev, want, count + 1, dst, s));
        } else {
            HIP_TRY(hipMemcpyAsync(want, bases.data(), bases.size() * sizeof(long long), hipMemcpyHostToDevice, s));
            HIP_TRY(hipStreamSynchronize(s));
        }
    }
    if (!mode.capture)
        fill_views(bp, B, r, bases, opts, C);
    return HSMU_OK;
}

} // namespace

extern "C" int hsmu_spgemm(hsmu_ctx *c, const hsmu_csr *A, const hsmu_csr *B, hsmu_csr *C, const hsmu_opts *opts_in,
                           hsmu_stats *stats)
{
    if (!c || !C)
        return fail(HSMU_ERR_INVALID, "hsmu_spgemm: ctx or C is NULL");
    const hsmu_opts opts = resolve_opts(opts_in);
    int rc = check_opts(opts);
    if (rc)
        return rc;
    HIP_TRY(hipSetDevice(c->device));
    memset(C, 0, sizeof(*C));
    if (stats)
        memset(stats, 0, sizeof(*stats));
    Arena ar(c, nullptr, false);
    bool empty;
    return run_batch(c, 1, A, B, C, opts, ar, RunMode(), stats, &empty);
}

static_assert(HSMU_NOALLOC_SCRATCH_BYTES == NOALLOC_SCRATCH_BYTES, "scratch size in hsmu.h");

extern "C" int hsmu_spgemm_noalloc(hsmu_ctx *c, const hsmu_csr *A, const hsmu_csr *B, hsmu_csr *C, int64_t cap_nnz,
                                   const hsmu_opts *opts_in, hsmu_stats *stats)
{
    if (!c || !C || cap_nnz < 0)
        return fail(HSMU_ERR_INVALID, "hsmu_spgemm_noalloc: ctx or C is NULL, or cap_nnz < 0");
    const hsmu_opts opts = resolve_opts(opts_in);
    int rc = check_opts(opts);
    if (rc)
        return rc;
    const bool fp32 = opts.precision == HSMU_FP32_COMP;
    rc = check_pair(A, B, fp32);
    if (rc)
        return rc;
    if (!C->rowptr || (cap_nnz > 0 && (!C->col || !C->val)))
        return fail(HSMU_ERR_INVALID, "hsmu_spgemm_noalloc: C's row pointers, or its columns/values with "
                                      "cap_nnz > 0, are NULL");
    if (!c->scratch)
        return fail(HSMU_ERR_UNSUPPORTED, "hsmu_spgemm_noalloc: the scratch array is not available on device %d",
                    c->device);
    HIP_TRY(hipSetDevice(c->device));
    if (stats)
        memset(stats, 0, sizeof(*stats));
    // every call on this device shares the array: one at a time in this process
    std::lock_guard<std::mutex> lk(scratch_lock(c->device));
    const FixedOut fo{C->rowptr, C->col, C->val, cap_nnz, fp32 ? sizeof(float) : sizeof(double)};
    Arena ar(c, (char *)c->scratch, NOALLOC_SCRATCH_BYTES, &fo);
    hsmu_csr R;
    memset(&R, 0, sizeof(R));
    bool empty;
    rc = run_batch(c, 1, A, B, &R, opts, ar, RunMode(), stats, &empty);
    // A call that failed part-way may have left kernels running on this context's stream;
    // they use the shared array, so drain them before the next call (from any context)
    // may take it.  (A successful call has waited for its stream already.)
    if (rc != HSMU_OK && hipStreamSynchronize(c->stream) != hipSuccess)
        (void)hipGetLastError();
    if (rc == HSMU_ERR_NOMEM && ar.want_nnz() > 0) {
        C->nnz = ar.want_nnz();
        return fail(HSMU_ERR_NOMEM, "hsmu_spgemm_noalloc: nnz(C) = %lld exceeds the capacity of C's arrays (%lld)",
                    (long long)ar.want_nnz(), (long long)cap_nnz);
    }
    if (rc == HSMU_ERR_NOMEM && ar.short_bytes() > 0)
        return fail(HSMU_ERR_NOMEM, "hsmu_spgemm_noalloc: the workspace needs more than the %zu MiB scratch array "
                                    "(a request reached byte %zu)",
                    NOALLOC_SCRATCH_BYTES >> 20, ar.short_bytes());
    if (rc)
        return rc;
    C->M = R.M;
    C->N = R.N;
    C->nnz = R.nnz;
    C->rowptr64 = R.rowptr64;
    C->val32 = R.val32;
    return HSMU_OK;
}

extern "C" int hsmu_spgemm_rows(hsmu_ctx *c, const hsmu_csr *A, int64_t r0, int64_t r1, const hsmu_csr *B, hsmu_csr *C,
                                const hsmu_opts *opts_in, hsmu_stats *stats)
{
    if (!c || !C || !A)
        return fail(HSMU_ERR_INVALID, "hsmu_spgemm_rows: NULL argument");
    const hsmu_opts opts = resolve_opts(opts_in);
    int rc = check_opts(opts);
    if (!rc)
        rc = check_pair(A, B, opts.precision == HSMU_FP32_COMP);
    if (rc)
        return rc;
    if (r0 < 0 || r1 < r0 || r1 > A->M)
        return fail(HSMU_ERR_INVALID, "hsmu_spgemm_rows: rows [%lld, %lld) of %lld", (long long)r0, (long long)r1,
                    (long long)A->M);
    HIP_TRY(hipSetDevice(c->device));
    memset(C, 0, sizeof(*C));
    if (stats)
        memset(stats, 0, sizeof(*stats));
    // the panel as a CSR matrix of its own: row pointers rebased to 0, columns and values
    // from the panel's first entry on
    const int *rp = (const int *)A->rowptr;
    int ends[2];
    HIP_TRY(hipMemcpyAsync(&ends[0], rp + r0, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(&ends[1], rp + r1, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (ends[0] < 0 || ends[1] < ends[0] || ends[1] > A->nnz)
        return fail(HSMU_ERR_INDEX, "hsmu_spgemm_rows: row pointers of the panel outside [0, nnz]");
    Arena ar(c, nullptr, false);
    int *prp = ar.get<int>((size_t)(r1 - r0) + 1);
    if (!prp)
        return fail(HSMU_ERR_NOMEM, "panel row pointers");
    HIP_TRY(launch_rowptr_rebase(rp + r0, r1 - r0 + 1, prp, c->stream));
    const size_t vsz = opts.precision == HSMU_FP32_COMP ? 4 : 8;
    hsmu_csr P = *A;
    P.M = r1 - r0;
    P.nnz = ends[1] - ends[0];
    P.rowptr = prp;
    P.col = A->col + ends[0];
    P.val = (char *)A->val + (size_t)ends[0] * vsz;
    bool empty;
    return run_batch(c, 1, &P, B, C, opts, ar, RunMode(), stats, &empty);
}

extern "C" int hsmu_csr_row_sums(hsmu_ctx *c, const hsmu_csr *C, double *sums, double *abs_sums)
{
    if (!c || !C || !sums || !abs_sums)
        return fail(HSMU_ERR_INVALID, "hsmu_csr_row_sums: NULL argument");
    if (C->M < 0 || !C->rowptr || (C->nnz > 0 && !C->val))
        return fail(HSMU_ERR_INVALID, "hsmu_csr_row_sums: incomplete matrix");
    HIP_TRY(hipSetDevice(c->device));
    Arena ar(c, nullptr, false);
    double *d = ar.get<double>(2 * (size_t)(C->M > 0 ? C->M : 1));
    if (!d)
        return fail(HSMU_ERR_NOMEM, "row sums");
    if (C->M > 0) {
        HIP_TRY(launch_row_sums(C->M, C->rowptr, C->rowptr64, C->val, C->val32, d, d + C->M, c->stream));
        HIP_TRY(hipMemcpyAsync(sums, d, (size_t)C->M * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipMemcpyAsync(abs_sums, d + C->M, (size_t)C->M * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    return HSMU_OK;
}

extern "C" int hsmu_spgemm_batched(hsmu_ctx *c, int64_t count, const hsmu_csr *A, const hsmu_csr *B, hsmu_csr *C,
                                   const hsmu_opts *opts_in, hsmu_stats *stats)
{
    if (!c || !C || count < 1)
        return fail(HSMU_ERR_INVALID, "hsmu_spgemm_batched: ctx or C is NULL, or count < 1");
    const hsmu_opts opts = resolve_opts(opts_in);
    int rc = check_opts(opts);
    if (rc)
        return rc;
    HIP_TRY(hipSetDevice(c->device));
    memset(C, 0, (size_t)count * sizeof(*C));
    if (stats)
        memset(stats, 0, sizeof(*stats));
    Arena ar(c, nullptr, false);
    bool empty;
    return run_batch(c, count, A, B, C, opts, ar, RunMode(), stats, &empty);
}

extern "C" int hsmu_spgemm_each(hsmu_ctx *c, int64_t count, const hsmu_csr *A, const hsmu_csr *B,
                                const hsmu_opts *opts, int32_t passes, double *ms_per_pass)
{
    if (!c || !A || !B || count < 1 || passes < 1 || !ms_per_pass)
        return fail(HSMU_ERR_INVALID, "hsmu_spgemm_each: NULL argument, count < 1 or passes < 1");
    HIP_TRY(hipSetDevice(c->device));
    const auto t0 = std::chrono::steady_clock::now();
    for (int32_t p = 0; p < passes; ++p) {
        for (int64_t i = 0; i < count; ++i) {
            hsmu_csr C;
            int rc = hsmu_spgemm(c, A + i, B + i, &C, opts, nullptr);
            if (rc)
                return fail(rc, "hsmu_spgemm_each: instance %lld: %s", (long long)i, hsmu_last_error());
            rc = hsmu_csr_free(c, &C);
            if (rc)
                return rc;
        }
    }
    *ms_per_pass = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / passes;
    return HSMU_OK;
}

extern "C" int hsmu_batch_free(hsmu_ctx *c, int64_t count, hsmu_csr *C)
{
    if (!c || !C || count < 1)
        return fail(HSMU_ERR_INVALID, "hsmu_batch_free: NULL argument or count < 1");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    out_free(c, C[0].rowptr);
    out_free(c, C[0].col);
    out_free(c, C[0].val);
    memset(C, 0, (size_t)count * sizeof(*C));
    return HSMU_OK;
}

// ---- plans -------------------------------------------------------------------------------------

extern "C" void hsmu_plan_destroy(hsmu_plan *p)
{
    if (!p)
        return;
    hsmu_ctx *c = p->c;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (p->exec)
        (void)hipGraphExecDestroy(p->exec);
    if (p->graph)
        (void)hipGraphDestroy(p->graph);
    for (void *q : p->keep)
        c->pool.release(q);
    delete p;
}

extern "C" int hsmu_plan_create(hsmu_ctx *c, int64_t count, const hsmu_csr *A, const hsmu_csr *B, hsmu_csr *C,
                                const hsmu_opts *opts_in, hsmu_plan **out)
{
    if (!c || !C || !out || count < 1)
        return fail(HSMU_ERR_INVALID, "hsmu_plan_create: NULL argument or count < 1");
    *out = nullptr;
    hsmu_opts opts = resolve_opts(opts_in);
    int rc = check_opts(opts);
    if (rc)
        return rc;
    opts.plain_alloc = 0; // the plan owns its outputs
    HIP_TRY(hipSetDevice(c->device));
    memset(C, 0, (size_t)count * sizeof(*C));
    hsmu_plan *p = new hsmu_plan();
    p->c = c;
    p->opts = opts;
    memset(&p->stats0, 0, sizeof(p->stats0));
    hipStream_t s = c->stream;
    // recording run: eager, every block kept
    {
        Arena ar(c, &p->keep, false);
        RunMode m;
        m.plan = p;
        rc = run_batch(c, count, A, B, C, opts, ar, m, &p->stats0, &p->empty);
    }
    if (!rc && !p->empty) {
        if (hipMemcpyAsync(p->expect, p->rec, sizeof(p->rec), hipMemcpyHostToDevice, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            rc = fail(HSMU_ERR_HIP, "hsmu_plan_create: uploading the checkpoints");
    }
    // capture pass: the same launches, the recorded sizes, no synchronisation
    if (!rc && !p->empty) {
        if (hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed) != hipSuccess) {
            rc = fail(HSMU_ERR_HIP, "hsmu_plan_create: hipStreamBeginCapture");
        } else {
            std::vector<hsmu_csr> scratch_c((size_t)count);
            Arena ar(c, &p->keep, true);
            RunMode m;
            m.plan = p;
            m.capture = true;
            bool empty;
            rc = run_batch(c, count, A, B, scratch_c.data(), opts, ar, m, nullptr, &empty);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(s, &g);
            if (!rc && e != hipSuccess)
                rc = fail(HSMU_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
            p->graph = g;
            if (!rc) {
                const hipError_t ei = hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0);
                if (ei != hipSuccess)
                    rc = fail(HSMU_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(ei));
            }
        }
    }
    if (rc) {
        (void)hipGetLastError();
        memset(C, 0, (size_t)count * sizeof(*C));
        hsmu_plan_destroy(p);
        return rc;
    }
    *out = p;
    return HSMU_OK;
}

extern "C" int hsmu_plan_execute(hsmu_plan *p, hsmu_stats *stats)
{
    if (!p)
        return fail(HSMU_ERR_INVALID, "hsmu_plan_execute: plan is NULL");
    hsmu_ctx *c = p->c;
    HIP_TRY(hipSetDevice(c->device));
    if (stats)
        *stats = p->stats0;
    if (p->empty)
        return HSMU_OK;
    hipStream_t s = c->stream;
    HIP_TRY(hipEventRecord(c->ev[EV_START], s));
    HIP_TRY(hipGraphLaunch(p->exec, s));
    HIP_TRY(hipEventRecord(c->ev[EV_NUMERIC], s));
    HIP_TRY(hipMemcpyAsync(&c->host_stats->err, &p->dst->err, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const int err = c->host_stats->err;
    if (err & ERR_PLAN)
        return fail(HSMU_ERR_PLAN, "the plan's inputs changed structure since it was made");
    int rc = device_error_status(err);
    if (rc)
        return rc;
    if (stats) {
        for (int b = 0; b < NBINS; ++b)
            stats->ms_numeric_bin[b] = 0.0;
        stats->ms_conversion = stats->ms_form_tileCptr = stats->ms_tileOR = stats->ms_malloc = 0.0;
        stats->ms_form_cval = 0.0;
        stats->ms_total = elapsed(c->ev[EV_START], c->ev[EV_NUMERIC]);
        stats->ms_time2 = stats->ms_total;
        stats->gflops = stats->ms_total > 0 ? 2.0 * (double)stats->products / (stats->ms_total * 1e6) : 0.0;
    }
    return HSMU_OK;
}
