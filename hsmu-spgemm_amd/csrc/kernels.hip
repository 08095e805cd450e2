// kernels.hip -- HIP kernels of the HSMU SpGEMM path for gfx950 (MI355X).
//
// Pipeline for C = A * B (all on one stream; host synchronises twice, to validate the
// input / size the symbolic launches and to size C):
//   1. B -> 64-column tiles (k_btile_count, scan, k_btile_fill).
//      Reference: src/compressed/compressed_mask64.cuh:61-424.  Tiles of a row come out
//      sorted by tile column (the reference's order is whatever LDS atomics give).
//   2. Row analysis of A (k_row_info): products, tile-products and the tile span
//      [rmin, rmax] of each row of C.  Reference: k_compute_flop, src/small/mask.cuh:14-51.
//   3. Binning (k_bin) by work and LDS need, wave/workgroup-aggregated appends.
//      Reference: src/bining.cuh:1-395 (fixed ranges there; derived from LDS need here).
//   4. Symbolic (k_symbolic_lds / k_symbolic_global): OR of B's tile masks over the A row
//      into a dense mask directory over the row's tile span, nnz(C_i) = sum popcount.
//      Reference: compressed_tilePtr_symbolic.cuh + compressed_tileOR.cuh, which do the
//      same in two hash passes; the dense directory needs one pass and no hashing.
//   5. Scan of nnz(C_i) to int64 row pointers.
//   6. Numeric (k_numeric_lds / k_numeric_global): the HSMU accumulator.  The row's
//      value array is exactly nnz(C_i) long; the slot of column c is
//        off[t] + popcount(mask[t] & ((1 << (c & 63)) - 1)),  t = (c >> 6) - rmin,
//      i.e. the rank of c among C's sorted columns, found in O(1) from the tile
//      directory where the reference binary-searches a sorted column list
//      (src/compressed/compressed_Cval.cuh:5-69, src/small/numeric.cuh:233-253).
//      The same kernel writes C's sorted columns by expanding the tile masks, which
//      replaces the reference's separate O(n^2) rank-sort pass
//      (src/compressed/compressed_Ccol_symbolic.cuh:4-63).
//   Rows whose directory + values do not fit LDS keep them in a global workspace and
//   accumulate with FP64 atomics straight into C (the reference's bin-13 dense spECK
//   path, src/external/spECK/dense_acc.cu:246-267, whose early return leaves rows
//   unwritten; every row is written here).

#include <hip/hip_runtime.h>

#include <climits>

#include "hsmu_internal.h"

namespace hsmu {

typedef unsigned long long u64;

#define DEVI __device__ __forceinline__

// ---- wave64 helpers ---------------------------------------------------------------

// number of set bits of m below this lane (v_mbcnt_lo/hi)
DEVI int mbcnt(u64 m)
{
    return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

DEVI u64 below(int b) { return (1ull << b) - 1ull; }

// Map the hardware block index so that each XCD (blocks are dealt round-robin over the
// 8 XCDs) works on a contiguous range of logical blocks: neighbouring rows, which share
// rows of B, then share one XCD's L2.
DEVI int xcd_block(int b, int n)
{
    const int x = b & 7, q = b >> 3, per = n >> 3, rem = n & 7;
    return x * per + (x < rem ? x : rem) + q;
}

// Sum over a group of G consecutive lanes.  G <= 64: shuffles; G > 64: through LDS
// (every thread of the block must call it).
template <int G, typename T> DEVI T group_sum(T v, T *red)
{
    if constexpr (G <= 64) {
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1)
            v += __shfl_xor(v, o, G);
        return v;
    } else {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1)
            v += __shfl_xor(v, o, 64);
        const int w = threadIdx.x >> 6;
        __syncthreads();
        if ((threadIdx.x & 63) == 0)
            red[w] = v;
        __syncthreads();
        const int w0 = (threadIdx.x / G) * (G / 64);
        T s = 0;
#pragma unroll
        for (int k = 0; k < G / 64; ++k)
            s += red[w0 + k];
        return s;
    }
}

template <int G, typename T> DEVI T group_min(T v)
{
    static_assert(G <= 64, "group_min is a wave operation");
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
        T y = __shfl_xor(v, o, G);
        v = y < v ? y : v;
    }
    return v;
}

template <int G, typename T> DEVI T group_max(T v)
{
    static_assert(G <= 64, "group_max is a wave operation");
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
        T y = __shfl_xor(v, o, G);
        v = y > v ? y : v;
    }
    return v;
}

// Exclusive scan over a group of G lanes; total = group sum.  G > 64 goes through LDS
// (every thread of the block must call it).
template <int G, typename T> DEVI T group_excl_scan(T v, T *red, T &total)
{
    constexpr int W = G < 64 ? G : 64;
    const int l = threadIdx.x & (W - 1);
    T x = v;
#pragma unroll
    for (int o = 1; o < W; o <<= 1) {
        T y = __shfl_up(x, o, W);
        if (l >= o)
            x += y;
    }
    if constexpr (G <= 64) {
        total = __shfl(x, W - 1, W);
        return x - v;
    } else {
        const int w = threadIdx.x >> 6;
        __syncthreads();
        if (l == 63)
            red[w] = x;
        __syncthreads();
        const int w0 = (threadIdx.x / G) * (G / 64);
        T before = 0, tot = 0;
#pragma unroll
        for (int k = 0; k < G / 64; ++k) {
            const T r = red[w0 + k];
            if (w0 + k < w)
                before += r;
            tot += r;
        }
        total = tot;
        return before + x - v;
    }
}

// Workgroup barrier that also orders global memory for the whole workgroup at device
// scope: prior stores are complete in L2 and this CU's L1 is invalidated afterwards,
// so plain loads see values other waves wrote with stores or L2 atomics.
DEVI void block_sync_global()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// ---- fill / copy -------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_fill_zero(uint4 *p, size_t n16, unsigned char *tail, size_t ntail)
{
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += stride)
        p[i] = make_uint4(0, 0, 0, 0);
    if (blockIdx.x == 0 && threadIdx.x < ntail)
        tail[threadIdx.x] = 0;
}

hipError_t launch_fill_zero(void *p, size_t bytes, hipStream_t s)
{
    if (bytes == 0)
        return hipSuccess;
    const size_t n16 = bytes / 16;
    const size_t ntail = bytes - n16 * 16;
    size_t blocks = (n16 + 255) / 256;
    if (blocks > 4096)
        blocks = 4096;
    if (blocks == 0)
        blocks = 1;
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)blocks), dim3(256), 0, s, (uint4 *)p, n16,
                       (unsigned char *)p + n16 * 16, ntail);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n16)
{
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += stride)
        dst[i] = src[i];
}

hipError_t launch_copy(const void *src, void *dst, size_t bytes, int num_cus, hipStream_t s)
{
    const size_t n16 = bytes / 16;
    size_t blocks = (n16 + 255) / 256;
    const size_t cap = (size_t)num_cus * 16;
    if (blocks > cap)
        blocks = cap;
    if (blocks == 0)
        blocks = 1;
    hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, s, (const uint4 *)src, (uint4 *)dst, n16);
    return hipGetLastError();
}

// ---- exclusive scan (three phases: block sums, scan of sums, apply) ---------------

constexpr int SCAN_BS = 256;
constexpr int SCAN_IPT = 16;
constexpr int SCAN_TILE = SCAN_BS * SCAN_IPT;

size_t scan_ws_bytes(int64_t n)
{
    const int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    return (size_t)(nb + 2) * sizeof(long long);
}

template <typename Tin>
__global__ __launch_bounds__(SCAN_BS) void k_scan_reduce(const Tin *__restrict__ in, int64_t n, long long *__restrict__ bsum)
{
    __shared__ long long red[SCAN_BS / 64];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    long long s = 0;
    for (int i = threadIdx.x; i < SCAN_TILE; i += SCAN_BS) {
        const int64_t k = base + i;
        if (k < n)
            s += (long long)in[k];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long t = 0;
        for (int w = 0; w < SCAN_BS / 64; ++w)
            t += red[w];
        bsum[blockIdx.x] = t;
    }
}

// One workgroup: exclusive scan of nb block sums in place; bsum[nb] = total.
__global__ __launch_bounds__(1024) void k_scan_blocks(long long *bsum, int64_t nb)
{
    __shared__ long long red[16];
    __shared__ long long carry_s;
    if (threadIdx.x == 0)
        carry_s = 0;
    __syncthreads();
    for (int64_t base = 0; base < nb; base += 1024) {
        const int64_t k = base + threadIdx.x;
        const long long v = k < nb ? bsum[k] : 0;
        long long total;
        const long long ex = group_excl_scan<1024>(v, red, total);
        const long long carry = carry_s;
        if (k < nb)
            bsum[k] = carry + ex;
        __syncthreads();
        if (threadIdx.x == 0)
            carry_s = carry + total;
        __syncthreads();
    }
    if (threadIdx.x == 0)
        bsum[nb] = carry_s;
}

template <typename Tin, typename Tout>
__global__ __launch_bounds__(SCAN_BS) void k_scan_apply(const Tin *__restrict__ in, int64_t n, const long long *__restrict__ bsum,
                                                         int64_t nb, Tout *__restrict__ out, long long *total_copy)
{
    __shared__ long long red[SCAN_BS / 64];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_IPT;
    long long v[SCAN_IPT];
    long long loc = 0;
#pragma unroll
    for (int i = 0; i < SCAN_IPT; ++i) {
        const int64_t k = base + i;
        v[i] = k < n ? (long long)in[k] : 0;
        loc += v[i];
    }
    long long total;
    long long pre = group_excl_scan<SCAN_BS>(loc, red, total) + bsum[blockIdx.x];
#pragma unroll
    for (int i = 0; i < SCAN_IPT; ++i) {
        const int64_t k = base + i;
        if (k < n)
            out[k] = (Tout)pre;
        pre += v[i];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        out[n] = (Tout)bsum[nb];
        if (total_copy)
            *total_copy = bsum[nb];
    }
}

template <typename Tin, typename Tout>
static hipError_t scan_excl(const Tin *in, int64_t n, Tout *out, void *ws, long long *total_copy, hipStream_t s)
{
    const int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    long long *bsum = (long long *)ws;
    if (nb > 0)
        hipLaunchKernelGGL((k_scan_reduce<Tin>), dim3((unsigned)nb), dim3(SCAN_BS), 0, s, in, n, bsum);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, s, bsum, nb);
    hipLaunchKernelGGL((k_scan_apply<Tin, Tout>), dim3((unsigned)(nb > 0 ? nb : 1)), dim3(SCAN_BS), 0, s, in, n, bsum, nb,
                       out, total_copy);
    return hipGetLastError();
}

hipError_t launch_scan_i32_i64(const int *in, int64_t n, long long *out, void *scan_ws, long long *total_copy,
                               hipStream_t s)
{
    return scan_excl<int, long long>(in, n, out, scan_ws, total_copy, s);
}

// ---- B -> tiles ------------------------------------------------------------------------

// Tiles per row of B: a stored column starts a tile when its tile column differs from
// its left neighbour's.  Also validates row pointers, column range and sortedness.
template <int G>
__global__ __launch_bounds__(256) void k_btile_count(int64_t K, int64_t N, int64_t nnz, const int *__restrict__ rp,
                                                     const int *__restrict__ col, int *__restrict__ tcnt, DevStats *st)
{
    constexpr int RPB = 256 / G;
    const int g = threadIdx.x / G, l = threadIdx.x & (G - 1);
    int bad = 0;
    for (int64_t j = (int64_t)blockIdx.x * RPB + g; j < K; j += (int64_t)gridDim.x * RPB) {
        const int s = rp[j], e = rp[j + 1];
        int c = 0;
        if (s < 0 || e < s || (int64_t)e > nnz) {
            bad |= ERR_ROWPTR;
        } else {
            for (int k = s + l; k < e; k += G) {
                const int cc = col[k];
                if ((unsigned)cc >= (unsigned long long)N)
                    bad |= ERR_BCOL;
                if (k == s) {
                    ++c;
                } else {
                    const int pc = col[k - 1];
                    if (pc > cc)
                        bad |= ERR_UNSORTED;
                    if ((pc >> 6) != (cc >> 6))
                        ++c;
                }
            }
        }
        c = group_sum<G>(c, (int *)nullptr);
        if (l == 0)
            tcnt[j] = c;
    }
    if (bad)
        atomicOr(&st->err, bad);
}

// Write tile columns and masks.  Heads (first element of each tile) get their tile
// index from a ballot prefix count; each head ORs its tile's bits.
template <int G>
__global__ __launch_bounds__(256) void k_btile_fill(int64_t K, int64_t nnz, const int *__restrict__ rp,
                                                    const int *__restrict__ col, const int *__restrict__ tptr,
                                                    int *__restrict__ tcol, u64 *__restrict__ tmask)
{
    constexpr int RPB = 256 / G;
    const int g = threadIdx.x / G, l = threadIdx.x & (G - 1);
    const int lane = threadIdx.x & 63;
    u64 gmask = ~0ull;
    if constexpr (G < 64)
        gmask = ((1ull << G) - 1ull) << (lane & ~(G - 1));
    for (int64_t j = (int64_t)blockIdx.x * RPB + g; j < K; j += (int64_t)gridDim.x * RPB) {
        const int s = rp[j], e = rp[j + 1];
        if (s < 0 || e < s || (int64_t)e > nnz)
            continue;
        int run = tptr[j];
        for (int k0 = s; k0 < e; k0 += G) {
            const int k = k0 + l;
            bool head = false;
            int cc = 0;
            if (k < e) {
                cc = col[k];
                head = (k == s) || ((col[k - 1] >> 6) != (cc >> 6));
            }
            const u64 bal = __ballot(head) & gmask;
            if (head) {
                const int t = run + mbcnt(bal);
                const int tc = cc >> 6;
                u64 m = 0;
                for (int q = k; q < e; ++q) {
                    const int c2 = col[q];
                    if ((c2 >> 6) != tc)
                        break;
                    m |= 1ull << (c2 & 63);
                }
                tcol[t] = tc;
                tmask[t] = m;
            }
            run += __popcll(bal);
        }
    }
}

template <int G>
static void btiles_g(int64_t K, int64_t N, int64_t nnz, const int *rp, const int *col, int *tcnt, int *tptr, int *tcol,
                     u64 *tmask, void *scan_ws, DevStats *st, int num_cus, hipStream_t s)
{
    constexpr int RPB = 256 / G;
    int64_t blocks = (K + RPB - 1) / RPB;
    const int64_t cap = (int64_t)num_cus * 32;
    if (blocks > cap)
        blocks = cap;
    if (blocks < 1)
        blocks = 1;
    hipLaunchKernelGGL((k_btile_count<G>), dim3((unsigned)blocks), dim3(256), 0, s, K, N, nnz, rp, col, tcnt, st);
    (void)scan_excl<int, int>(tcnt, K, tptr, scan_ws, nullptr, s);
    hipLaunchKernelGGL((k_btile_fill<G>), dim3((unsigned)blocks), dim3(256), 0, s, K, nnz, rp, col, tptr, tcol, tmask);
}

__global__ void k_copy_ntiles(const int *tptr, int64_t K, DevStats *st) { st->ntiles_b = tptr[K]; }

hipError_t launch_btiles(int64_t K, int64_t N, int64_t nnz, const int *rp, const int *col, int *tcnt, int *tptr,
                         int *tcol, u64 *tmask, void *scan_ws, DevStats *st, int num_cus, hipStream_t s)
{
    const int64_t avg = K > 0 ? nnz / K : 0;
    if (avg <= 12)
        btiles_g<8>(K, N, nnz, rp, col, tcnt, tptr, tcol, tmask, scan_ws, st, num_cus, s);
    else if (avg <= 24)
        btiles_g<16>(K, N, nnz, rp, col, tcnt, tptr, tcol, tmask, scan_ws, st, num_cus, s);
    else if (avg <= 48)
        btiles_g<32>(K, N, nnz, rp, col, tcnt, tptr, tcol, tmask, scan_ws, st, num_cus, s);
    else
        btiles_g<64>(K, N, nnz, rp, col, tcnt, tptr, tcol, tmask, scan_ws, st, num_cus, s);
    hipLaunchKernelGGL(k_copy_ntiles, dim3(1), dim3(1), 0, s, tptr, K, st);
    return hipGetLastError();
}

// ---- row analysis of A ---------------------------------------------------------------

// Per row i of A: products sum_j nnz(B_j), tile-products sum_j tiles(B_j), and the span
// [rmin, rmax] of tile columns of row i of C (rows of B are sorted, so each B row's
// first and last tile bound it).  Reference: k_compute_flop, src/small/mask.cuh:14-51.
template <int G>
__global__ __launch_bounds__(256) void k_row_info(int64_t M, int64_t K, int64_t nnzA, const int *__restrict__ Arp,
                                                  const int *__restrict__ Acol, const int *__restrict__ Brp,
                                                  const int *__restrict__ tptr, const int *__restrict__ tcol,
                                                  long long *__restrict__ rprod, int *__restrict__ rtp,
                                                  int *__restrict__ rmin, int *__restrict__ rmax, int *__restrict__ rnnz,
                                                  DevStats *st)
{
    constexpr int RPB = 256 / G;
    const int g = threadIdx.x / G, l = threadIdx.x & (G - 1);
    int bad = 0;
    unsigned long long tot_p = 0, tot_tp = 0, nonempty = 0;
    int span_max = 0;
    for (int64_t i = (int64_t)blockIdx.x * RPB + g; i < M; i += (int64_t)gridDim.x * RPB) {
        const int s = Arp[i], e = Arp[i + 1];
        long long p = 0;
        int tp = 0, mn = INT_MAX, mx = -1;
        if (s < 0 || e < s || (int64_t)e > nnzA) {
            bad |= ERR_ROWPTR;
        } else {
            for (int k = s + l; k < e; k += G) {
                const int j = Acol[k];
                if ((unsigned)j >= (unsigned long long)K) {
                    bad |= ERR_ACOL;
                    continue;
                }
                const int b0 = Brp[j], b1 = Brp[j + 1];
                const int t0 = tptr[j], t1 = tptr[j + 1];
                if (b1 > b0 && t1 > t0) {
                    p += b1 - b0;
                    tp += t1 - t0;
                    const int f = tcol[t0], z = tcol[t1 - 1];
                    mn = f < mn ? f : mn;
                    mx = z > mx ? z : mx;
                }
            }
        }
        p = group_sum<G>(p, (long long *)nullptr);
        tp = group_sum<G>(tp, (int *)nullptr);
        mn = group_min<G>(mn);
        mx = group_max<G>(mx);
        if (l == 0) {
            rprod[i] = p;
            rtp[i] = tp;
            rmin[i] = mn;
            rmax[i] = mx;
            if (p == 0)
                rnnz[i] = 0;
            tot_p += (unsigned long long)p;
            tot_tp += (unsigned long long)tp;
            if (p > 0) {
                ++nonempty;
                const int span = mx - mn + 1;
                span_max = span > span_max ? span : span_max;
            }
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        tot_p += __shfl_xor(tot_p, o, 64);
        tot_tp += __shfl_xor(tot_tp, o, 64);
        nonempty += __shfl_xor(nonempty, o, 64);
        bad |= __shfl_xor(bad, o, 64);
        const int y = __shfl_xor(span_max, o, 64);
        span_max = y > span_max ? y : span_max;
    }
    if ((threadIdx.x & 63) == 0) {
        if (tot_p)
            atomicAdd(&st->products, tot_p);
        if (tot_tp)
            atomicAdd(&st->tile_products, tot_tp);
        if (nonempty)
            atomicAdd(&st->rows_nonempty, nonempty);
        if (bad)
            atomicOr(&st->err, bad);
        if (span_max)
            atomicMax(&st->max_span, span_max);
    }
}

hipError_t launch_row_info(int64_t M, int64_t K, int64_t nnzA, const int *Arp, const int *Acol, const int *Brp,
                           const int *tptr, const int *tcol, long long *rprod, int *rtp, int *rmin, int *rmax, int *rnnz,
                           DevStats *st, int avg_a, hipStream_t s)
{
#define ROW_INFO(G)                                                                                                    \
    {                                                                                                                  \
        constexpr int RPB = 256 / G;                                                                                   \
        int64_t blocks = (M + RPB - 1) / RPB;                                                                          \
        if (blocks > 65536)                                                                                            \
            blocks = 65536;                                                                                            \
        if (blocks < 1)                                                                                                \
            blocks = 1;                                                                                                \
        hipLaunchKernelGGL((k_row_info<G>), dim3((unsigned)blocks), dim3(256), 0, s, M, K, nnzA, Arp, Acol, Brp, tptr, \
                           tcol, rprod, rtp, rmin, rmax, rnnz, st);                                                    \
    }
    if (avg_a <= 12)
        ROW_INFO(8)
    else if (avg_a <= 24)
        ROW_INFO(16)
    else if (avg_a <= 48)
        ROW_INFO(32)
    else
        ROW_INFO(64)
#undef ROW_INFO
    return hipGetLastError();
}

// ---- binning -----------------------------------------------------------------------------

// Bin of one row: lanes from work (per_lane items per lane, 8..1024 lanes), then fewer
// rows per workgroup until the workgroup's LDS fits the budget; rows whose own need
// exceeds lds_row_max go to the global-memory bin.  Reference ranges being replaced:
// src/bining.cuh:12 (symbolic), :121 (tile OR), :228 (numeric).
DEVI int choose_bin(long long work, long long need, const BinParams &bp)
{
    if (bp.force_global || need > bp.lds_row_max)
        return BIN_GLOBAL;
    const long long lanes = (work + bp.per_lane - 1) / bp.per_lane;
    int b = 0;
    while (b < NLDS - 1 && (8LL << b) < lanes)
        ++b;
    while (b < NLDS - 1) {
        const int G = 8 << b;
        const long long rpb = G >= BLOCK ? 1 : BLOCK / G;
        if (need * rpb <= bp.lds_block_budget)
            break;
        ++b;
    }
    return b;
}

// One thread per row; appends are aggregated per workgroup in LDS, then one global
// atomic per bin per workgroup reserves the range.  Also sums each bin's work, nnz(A)
// and (numeric) nnz(C) for the per-kernel algorithmic byte counts.
__global__ __launch_bounds__(1024) void k_bin(int numeric, int64_t M, const int *__restrict__ Arp,
                                              const long long *__restrict__ rprod, const int *__restrict__ rtp,
                                              const int *__restrict__ rmin, const int *__restrict__ rmax,
                                              const int *__restrict__ rnnz, BinParams bp, int *__restrict__ binrows,
                                              DevStats *st)
{
    __shared__ int s_cnt[NBINS], s_need[NBINS], s_span[NBINS], s_base[NBINS];
    __shared__ unsigned long long s_work[NBINS], s_nc[NBINS], s_na[NBINS];
    if (threadIdx.x < NBINS) {
        s_cnt[threadIdx.x] = 0;
        s_need[threadIdx.x] = 0;
        s_span[threadIdx.x] = 0;
        s_work[threadIdx.x] = 0;
        s_nc[threadIdx.x] = 0;
        s_na[threadIdx.x] = 0;
    }
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * 1024 + threadIdx.x;
    int b = -1, pos = 0;
    if (i < M && rprod[i] > 0) {
        const int S = rmax[i] - rmin[i] + 1;
        long long work, need;
        int nc = 0;
        if (numeric) {
            nc = rnnz[i];
            need = num_need(S, nc);
            work = rprod[i];
        } else {
            need = sym_need(S);
            work = rtp[i];
        }
        b = choose_bin(work, need, bp);
        pos = atomicAdd(&s_cnt[b], 1);
        atomicMax(&s_need[b], (int)(need < INT_MAX ? need : INT_MAX));
        atomicMax(&s_span[b], S);
        atomicAdd(&s_work[b], (unsigned long long)work);
        atomicAdd(&s_nc[b], (unsigned long long)nc);
        atomicAdd(&s_na[b], (unsigned long long)(Arp[i + 1] - Arp[i]));
    }
    __syncthreads();
    BinStats *bs = numeric ? &st->num : &st->sym;
    if (threadIdx.x < NBINS) {
        const int c = s_cnt[threadIdx.x];
        s_base[threadIdx.x] = c ? atomicAdd(&bs->count[threadIdx.x], c) : 0;
        if (c) {
            atomicMax(&bs->max_need[threadIdx.x], s_need[threadIdx.x]);
            atomicMax(&bs->max_span[threadIdx.x], s_span[threadIdx.x]);
            atomicAdd(&bs->work[threadIdx.x], s_work[threadIdx.x]);
            atomicAdd(&bs->nnz_c[threadIdx.x], s_nc[threadIdx.x]);
            atomicAdd(&bs->nnz_a[threadIdx.x], s_na[threadIdx.x]);
        }
    }
    __syncthreads();
    if (b >= 0)
        binrows[(int64_t)b * M + s_base[b] + pos] = (int)i;
}

hipError_t launch_bin(int numeric, int64_t M, const int *Arp, const long long *rprod, const int *rtp, const int *rmin,
                      const int *rmax, const int *rnnz, BinParams bp, int *binrows, DevStats *st, hipStream_t s)
{
    const int64_t blocks = (M + 1023) / 1024;
    if (blocks == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_bin, dim3((unsigned)blocks), dim3(1024), 0, s, numeric, M, Arp, rprod, rtp, rmin, rmax, rnnz,
                       bp, binrows, st);
    return hipGetLastError();
}

// ---- symbolic ------------------------------------------------------------------------------

// nnz(C_i) for rows kept in LDS: a dense 64-bit mask per tile of the row's span,
// OR-accumulated with ds_or_b64 over every (a_ij, tile of B_j), then popcounted.
// One group of G lanes per row; within it, sub-groups of 2^tshift lanes walk one A
// entry's tiles each.  Reference: k_tileOR_shared_hash_tb,
// src/compressed/compressed_tileOR.cuh:20-104 (hash keyed by tile there).
template <int G>
__global__ __launch_bounds__(G < BLOCK ? BLOCK : G) void k_symbolic_lds(const int *__restrict__ rows,
                                                                       const int *__restrict__ count_p, int slice_bytes,
                                                                       Operands op, int *__restrict__ rnnz, int tshift)
{
    constexpr int BS = G < BLOCK ? BLOCK : G;
    constexpr int RPB = BS / G;
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ int red[16];
    const int g = threadIdx.x / G, l = threadIdx.x & (G - 1);
    u64 *dm = (u64 *)(lds + (size_t)g * slice_bytes);
    const int nrows = *count_p;
    const int SG = 1 << tshift, sl = l & (SG - 1), sg = l >> tshift, NSG = G >> tshift;
    for (int r0 = xcd_block(blockIdx.x, gridDim.x) * RPB; r0 < nrows; r0 += gridDim.x * RPB) {
        const int idx = r0 + g;
        const int row = idx < nrows ? rows[idx] : -1;
        int S = 0, mn = 0;
        if (row >= 0) {
            mn = op.rmin[row];
            S = op.rmax[row] - mn + 1;
        }
        for (int t = l; t < S; t += G)
            dm[t] = 0;
        __syncthreads();
        if (row >= 0) {
            const int a1 = op.Arp[row + 1];
            for (int e = op.Arp[row] + sg; e < a1; e += NSG) {
                const int j = op.Acol[e];
                const int t1 = op.tptr[j + 1];
                for (int t = op.tptr[j] + sl; t < t1; t += SG)
                    atomicOr(&dm[op.tcol[t] - mn], op.tmask[t]);
            }
        }
        __syncthreads();
        int cnt = 0;
        for (int t = l; t < S; t += G)
            cnt += __popcll(dm[t]);
        cnt = group_sum<G>(cnt, red);
        if (row >= 0 && l == 0)
            rnnz[row] = cnt;
        __syncthreads();
    }
}

// Rows too wide for LDS: the same dense directory in a per-workgroup global workspace,
// updated with L2 atomics.  One 1024-lane workgroup per row.
__global__ __launch_bounds__(1024) void k_symbolic_global(const int *__restrict__ rows, const int *__restrict__ count_p,
                                                          Operands op, int *__restrict__ rnnz, int tshift, u64 *gws,
                                                          long long gws_stride)
{
    __shared__ int red[16];
    const int l = threadIdx.x;
    u64 *dm = gws + (size_t)blockIdx.x * gws_stride;
    const int nrows = *count_p;
    const int SG = 1 << tshift, sl = l & (SG - 1), sg = l >> tshift, NSG = 1024 >> tshift;
    for (int idx = blockIdx.x; idx < nrows; idx += gridDim.x) {
        const int row = rows[idx];
        const int mn = op.rmin[row];
        const int S = op.rmax[row] - mn + 1;
        for (int t = l; t < S; t += 1024)
            dm[t] = 0;
        block_sync_global();
        const int a1 = op.Arp[row + 1];
        for (int e = op.Arp[row] + sg; e < a1; e += NSG) {
            const int j = op.Acol[e];
            const int t1 = op.tptr[j + 1];
            for (int t = op.tptr[j] + sl; t < t1; t += SG)
                atomicOr(&dm[op.tcol[t] - mn], op.tmask[t]);
        }
        block_sync_global();
        int cnt = 0;
        for (int t = l; t < S; t += 1024)
            cnt += __popcll(dm[t]);
        cnt = group_sum<1024>(cnt, red);
        if (l == 0)
            rnnz[row] = cnt;
        block_sync_global();
    }
}

template <int G>
static void symbolic_g(const int *rows, const int *count_dev, int count, int max_need, const Operands &op, int *rnnz,
                       int tshift, hipStream_t s)
{
    constexpr int BS = G < BLOCK ? BLOCK : G;
    constexpr int RPB = BS / G;
    const int slice = (max_need + 15) & ~15;
    int64_t blocks = ((int64_t)count + RPB - 1) / RPB;
    if (blocks > (1 << 20))
        blocks = 1 << 20;
    if (tshift > __builtin_ctz(G))
        tshift = __builtin_ctz(G);
    hipLaunchKernelGGL((k_symbolic_lds<G>), dim3((unsigned)blocks), dim3(BS), (size_t)slice * RPB, s, rows, count_dev,
                       slice, op, rnnz, tshift);
}

hipError_t launch_symbolic(int bin, const int *rows, const int *count_dev, int count, int max_need, const Operands &op,
                           int *rnnz, int tshift, u64 *gws, long long gws_stride, int num_cus, int gws_blocks,
                           hipStream_t s)
{
    if (count <= 0)
        return hipSuccess;
    switch (bin) {
    case 0: symbolic_g<8>(rows, count_dev, count, max_need, op, rnnz, tshift, s); break;
    case 1: symbolic_g<16>(rows, count_dev, count, max_need, op, rnnz, tshift, s); break;
    case 2: symbolic_g<32>(rows, count_dev, count, max_need, op, rnnz, tshift, s); break;
    case 3: symbolic_g<64>(rows, count_dev, count, max_need, op, rnnz, tshift, s); break;
    case 4: symbolic_g<128>(rows, count_dev, count, max_need, op, rnnz, tshift, s); break;
    case 5: symbolic_g<256>(rows, count_dev, count, max_need, op, rnnz, tshift, s); break;
    case 6: symbolic_g<512>(rows, count_dev, count, max_need, op, rnnz, tshift, s); break;
    case 7: symbolic_g<1024>(rows, count_dev, count, max_need, op, rnnz, tshift, s); break;
    case BIN_GLOBAL: {
        const int blocks = count < gws_blocks ? count : gws_blocks;
        hipLaunchKernelGGL(k_symbolic_global, dim3(blocks), dim3(1024), 0, s, rows, count_dev, op, rnnz,
                           tshift > 10 ? 10 : tshift, gws, gws_stride);
        break;
    }
    default: return hipErrorInvalidValue;
    }
    (void)num_cus;
    return hipGetLastError();
}

// ---- numeric -------------------------------------------------------------------------------

// The HSMU accumulator in LDS.  Per row (slice of `slice_bytes`):
//   dm[S]   64-bit tile masks of the row of C (rebuilt as in the symbolic phase)
//   doff[S] exclusive prefix of popcount(dm) = slot of the tile's first column
//   vals[n] FP64 accumulators, exactly nnz(C_i) of them
// Product a_ij * b_jc goes to slot doff[t] + popcount(dm[t] & below(c & 63)) with
// ds_add_f64.  The row of C is then written in column order: for G >= 64 one wave per
// tile, lane b writing column 64*(rmin+t)+b if present (coalesced); for G < 64 one lane
// per tile walking its set bits.
template <int G>
__global__ __launch_bounds__(G < BLOCK ? BLOCK : G) void k_numeric_lds(const int *__restrict__ rows,
                                                                      const int *__restrict__ count_p, int slice_bytes,
                                                                      Operands op, const long long *__restrict__ crp,
                                                                      int *__restrict__ Ccol, double *__restrict__ Cval,
                                                                      int bshift, int tshift, int *err)
{
    constexpr int BS = G < BLOCK ? BLOCK : G;
    constexpr int RPB = BS / G;
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ unsigned red[16];
    const int g = threadIdx.x / G, l = threadIdx.x & (G - 1);
    unsigned char *slice = lds + (size_t)g * slice_bytes;
    const int nrows = *count_p;
    for (int r0 = xcd_block(blockIdx.x, gridDim.x) * RPB; r0 < nrows; r0 += gridDim.x * RPB) {
        const int idx = r0 + g;
        const int row = idx < nrows ? rows[idx] : -1;
        int S = 0, mn = 0, n = 0;
        long long cb = 0;
        if (row >= 0) {
            mn = op.rmin[row];
            S = op.rmax[row] - mn + 1;
            cb = crp[row];
            n = (int)(crp[row + 1] - cb);
        }
        u64 *dm = (u64 *)slice;
        unsigned *doff = (unsigned *)(slice + 8 * S);
        double *vals = (double *)(slice + ((12 * S + 7) & ~7));
        for (int t = l; t < S; t += G)
            dm[t] = 0;
        for (int q = l; q < n; q += G)
            vals[q] = 0.0;
        __syncthreads();
        if (row >= 0) {
            const int SG = 1 << tshift, sl = l & (SG - 1), sg = l >> tshift, NSG = G >> tshift;
            const int a1 = op.Arp[row + 1];
            for (int e = op.Arp[row] + sg; e < a1; e += NSG) {
                const int j = op.Acol[e];
                const int t1 = op.tptr[j + 1];
                for (int t = op.tptr[j] + sl; t < t1; t += SG)
                    atomicOr(&dm[op.tcol[t] - mn], op.tmask[t]);
            }
        }
        __syncthreads();
        // slot offsets: each lane scans a contiguous chunk of tiles
        const int chunk = (S + G - 1) / G;
        const int ta = l * chunk;
        const int tb = ta + chunk < S ? ta + chunk : S;
        unsigned loc = 0;
        for (int t = ta; t < tb; ++t)
            loc += (unsigned)__popcll(dm[t]);
        unsigned tot;
        unsigned pre = group_excl_scan<G>(loc, red, tot);
        for (int t = ta; t < tb; ++t) {
            doff[t] = pre;
            pre += (unsigned)__popcll(dm[t]);
        }
        const bool ok = (int)tot == n;
        if (row >= 0 && !ok && l == 0)
            atomicOr(err, ERR_INTERNAL);
        __syncthreads();
        if (row >= 0 && ok) {
            const int SG = 1 << bshift, sl = l & (SG - 1), sg = l >> bshift, NSG = G >> bshift;
            const int a1 = op.Arp[row + 1];
            for (int e = op.Arp[row] + sg; e < a1; e += NSG) {
                const int j = op.Acol[e];
                const double a = op.Aval[e];
                const int b1 = op.Brp[j + 1];
                for (int k = op.Brp[j] + sl; k < b1; k += SG) {
                    const int c = op.Bcol[k];
                    const int t = (c >> 6) - mn;
                    const unsigned slot = doff[t] + (unsigned)__popcll(dm[t] & below(c & 63));
                    atomicAdd(&vals[slot], a * op.Bval[k]);
                }
            }
        }
        __syncthreads();
        if (row >= 0 && ok) {
            if constexpr (G >= 64) {
                constexpr int NW = G / 64;
                const int w = l >> 6, lane = l & 63;
                for (int t = w; t < S; t += NW) {
                    const u64 m = dm[t];
                    if ((m >> lane) & 1ull) {
                        const unsigned pos = doff[t] + (unsigned)__popcll(m & below(lane));
                        Ccol[cb + pos] = ((mn + t) << 6) + lane;
                        Cval[cb + pos] = vals[pos];
                    }
                }
            } else {
                for (int t = l; t < S; t += G) {
                    u64 m = dm[t];
                    unsigned pos = doff[t];
                    const int c0 = (mn + t) << 6;
                    while (m) {
                        const int b = __ffsll((unsigned long long)m) - 1;
                        Ccol[cb + pos] = c0 + b;
                        ++pos;
                        m &= m - 1;
                    }
                }
                for (int q = l; q < n; q += G)
                    Cval[cb + q] = vals[q];
            }
        }
        __syncthreads();
    }
}

// Rows that do not fit LDS: directory (masks + offsets) in a per-workgroup global
// workspace, FP64 accumulation with global atomics straight into C's value array.
__global__ __launch_bounds__(1024) void k_numeric_global(const int *__restrict__ rows, const int *__restrict__ count_p,
                                                         Operands op, const long long *__restrict__ crp,
                                                         int *__restrict__ Ccol, double *Cval, int bshift, int tshift,
                                                         u64 *gws, long long gws_stride, long long gws_tiles, int *err)
{
    __shared__ unsigned red[16];
    __shared__ int ok_s;
    const int l = threadIdx.x;
    // workgroup region: gws_tiles masks, then gws_tiles 32-bit offsets
    u64 *dm = gws + (size_t)blockIdx.x * gws_stride;
    unsigned *doff = (unsigned *)(dm + gws_tiles);
    const int nrows = *count_p;
    for (int idx = blockIdx.x; idx < nrows; idx += gridDim.x) {
        const int row = rows[idx];
        const int mn = op.rmin[row];
        const int S = op.rmax[row] - mn + 1;
        const long long cb = crp[row];
        const int n = (int)(crp[row + 1] - cb);
        for (int t = l; t < S; t += 1024)
            dm[t] = 0;
        for (int q = l; q < n; q += 1024)
            Cval[cb + q] = 0.0;
        block_sync_global();
        {
            const int SG = 1 << tshift, sl = l & (SG - 1), sg = l >> tshift, NSG = 1024 >> tshift;
            const int a1 = op.Arp[row + 1];
            for (int e = op.Arp[row] + sg; e < a1; e += NSG) {
                const int j = op.Acol[e];
                const int t1 = op.tptr[j + 1];
                for (int t = op.tptr[j] + sl; t < t1; t += SG)
                    atomicOr(&dm[op.tcol[t] - mn], op.tmask[t]);
            }
        }
        block_sync_global();
        const int chunk = (S + 1023) / 1024;
        const int ta = l * chunk;
        const int tb = ta + chunk < S ? ta + chunk : S;
        unsigned loc = 0;
        for (int t = ta; t < tb; ++t)
            loc += (unsigned)__popcll(dm[t]);
        unsigned tot;
        unsigned pre = group_excl_scan<1024>(loc, red, tot);
        for (int t = ta; t < tb; ++t) {
            doff[t] = pre;
            pre += (unsigned)__popcll(dm[t]);
        }
        if (l == 0) {
            ok_s = (int)tot == n;
            if ((int)tot != n)
                atomicOr(err, ERR_INTERNAL);
        }
        block_sync_global();
        if (ok_s) {
            const int SG = 1 << bshift, sl = l & (SG - 1), sg = l >> bshift, NSG = 1024 >> bshift;
            const int a1 = op.Arp[row + 1];
            for (int e = op.Arp[row] + sg; e < a1; e += NSG) {
                const int j = op.Acol[e];
                const double a = op.Aval[e];
                const int b1 = op.Brp[j + 1];
                for (int k = op.Brp[j] + sl; k < b1; k += SG) {
                    const int c = op.Bcol[k];
                    const int t = (c >> 6) - mn;
                    const unsigned slot = doff[t] + (unsigned)__popcll(dm[t] & below(c & 63));
                    unsafeAtomicAdd(&Cval[cb + slot], a * op.Bval[k]);
                }
            }
            const int w = l >> 6, lane = l & 63;
            for (int t = w; t < S; t += 16) {
                const u64 m = dm[t];
                if ((m >> lane) & 1ull) {
                    const unsigned pos = doff[t] + (unsigned)__popcll(m & below(lane));
                    Ccol[cb + pos] = ((mn + t) << 6) + lane;
                }
            }
        }
        block_sync_global();
    }
}

template <int G>
static void numeric_g(const int *rows, const int *count_dev, int count, int max_need, const Operands &op,
                      const long long *crp, int *Ccol, double *Cval, int bshift, int tshift, int *err, hipStream_t s)
{
    constexpr int BS = G < BLOCK ? BLOCK : G;
    constexpr int RPB = BS / G;
    const int slice = (max_need + 15) & ~15;
    int64_t blocks = ((int64_t)count + RPB - 1) / RPB;
    if (blocks > (1 << 20))
        blocks = 1 << 20;
    const int lg = __builtin_ctz(G);
    if (bshift > lg)
        bshift = lg;
    if (tshift > lg)
        tshift = lg;
    hipLaunchKernelGGL((k_numeric_lds<G>), dim3((unsigned)blocks), dim3(BS), (size_t)slice * RPB, s, rows, count_dev,
                       slice, op, crp, Ccol, Cval, bshift, tshift, err);
}

hipError_t launch_numeric(int bin, const int *rows, const int *count_dev, int count, int max_need, const Operands &op,
                          const long long *crp, int *Ccol, double *Cval, int bshift, int tshift, u64 *gws,
                          long long gws_stride, long long gws_tiles, int gws_blocks, int *err, hipStream_t s)
{
    if (count <= 0)
        return hipSuccess;
    switch (bin) {
    case 0: numeric_g<8>(rows, count_dev, count, max_need, op, crp, Ccol, Cval, bshift, tshift, err, s); break;
    case 1: numeric_g<16>(rows, count_dev, count, max_need, op, crp, Ccol, Cval, bshift, tshift, err, s); break;
    case 2: numeric_g<32>(rows, count_dev, count, max_need, op, crp, Ccol, Cval, bshift, tshift, err, s); break;
    case 3: numeric_g<64>(rows, count_dev, count, max_need, op, crp, Ccol, Cval, bshift, tshift, err, s); break;
    case 4: numeric_g<128>(rows, count_dev, count, max_need, op, crp, Ccol, Cval, bshift, tshift, err, s); break;
    case 5: numeric_g<256>(rows, count_dev, count, max_need, op, crp, Ccol, Cval, bshift, tshift, err, s); break;
    case 6: numeric_g<512>(rows, count_dev, count, max_need, op, crp, Ccol, Cval, bshift, tshift, err, s); break;
    case 7: numeric_g<1024>(rows, count_dev, count, max_need, op, crp, Ccol, Cval, bshift, tshift, err, s); break;
    case BIN_GLOBAL: {
        const int blocks = count < gws_blocks ? count : gws_blocks;
        hipLaunchKernelGGL(k_numeric_global, dim3(blocks), dim3(1024), 0, s, rows, count_dev, op, crp, Ccol, Cval,
                           bshift > 10 ? 10 : bshift, tshift > 10 ? 10 : tshift, gws, gws_stride, gws_tiles, err);
        break;
    }
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---- misc ------------------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_i64_to_i32(const long long *__restrict__ in, int64_t n, int *__restrict__ out)
{
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride)
        out[i] = (int)in[i];
}

hipError_t launch_rowptr_i64_to_i32(const long long *in, int64_t n, int *out, hipStream_t s)
{
    int64_t blocks = (n + 255) / 256;
    if (blocks > 8192)
        blocks = 8192;
    if (blocks < 1)
        blocks = 1;
    hipLaunchKernelGGL(k_i64_to_i32, dim3((unsigned)blocks), dim3(256), 0, s, in, n, out);
    return hipGetLastError();
}

hipError_t configure_kernels(int lds_max)
{
    hipError_t e = hipSuccess;
#define CFG(K)                                                                                                         \
    {                                                                                                                  \
        hipError_t r = hipFuncSetAttribute((const void *)(K), hipFuncAttributeMaxDynamicSharedMemorySize, lds_max);    \
        if (r != hipSuccess)                                                                                           \
            e = r;                                                                                                     \
    }
    CFG(k_symbolic_lds<8>) CFG(k_symbolic_lds<16>) CFG(k_symbolic_lds<32>) CFG(k_symbolic_lds<64>)
    CFG(k_symbolic_lds<128>) CFG(k_symbolic_lds<256>) CFG(k_symbolic_lds<512>) CFG(k_symbolic_lds<1024>)
    CFG(k_numeric_lds<8>) CFG(k_numeric_lds<16>) CFG(k_numeric_lds<32>) CFG(k_numeric_lds<64>)
    CFG(k_numeric_lds<128>) CFG(k_numeric_lds<256>) CFG(k_numeric_lds<512>) CFG(k_numeric_lds<1024>)
#undef CFG
    return e;
}

} // namespace hsmu
