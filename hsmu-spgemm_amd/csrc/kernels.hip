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
//   6. Numeric (k_numeric_lds / k_numeric_window): the HSMU accumulator.  The row's
//      value array is exactly nnz(C_i) long; the slot of column c is
//        off[t] + popcount(mask[t] & ((1 << (c & 63)) - 1)),  t = (c >> 6) - rmin,
//      i.e. the rank of c among C's sorted columns, found in O(1) from the tile
//      directory where the reference binary-searches a sorted column list
//      (src/compressed/compressed_Cval.cuh:5-69, src/small/numeric.cuh:233-253).
//      The same kernel writes C's sorted columns by expanding the tile masks, which
//      replaces the reference's separate O(n^2) rank-sort pass
//      (src/compressed/compressed_Ccol_symbolic.cuh:4-63).
//   Rows whose directory + values do not fit LDS: the symbolic tile OR runs on a global
//   workspace, the numeric phase slides a dense LDS window over the row's columns (the
//   reference's bin-13 dense spECK path, src/external/spECK/dense_acc.cu:246-267, whose
//   early return leaves rows unwritten; every row is written here).

#include <hip/hip_runtime.h>

#include <climits>

#include "device_common.h"
#include "hsmu_internal.h"

namespace hsmu {

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

// Each thread scans SCAN_IPT consecutive elements; the block's tile passes through LDS
// (padded by one element per SCAN_IPT, so a thread's run and a wave's coalesced row both
// spread over the banks) so that global loads and stores stay coalesced.
DEVI int scan_pad(int k) { return k + k / SCAN_IPT; }

template <typename Tin, typename Tout>
__global__ __launch_bounds__(SCAN_BS) void k_scan_apply(const Tin *__restrict__ in, int64_t n, const long long *__restrict__ bsum,
                                                         int64_t nb, Tout *__restrict__ out, long long *total_copy)
{
    __shared__ long long red[SCAN_BS / 64];
    __shared__ Tout tile[SCAN_TILE + SCAN_TILE / SCAN_IPT];
    const int64_t b0 = (int64_t)blockIdx.x * SCAN_TILE;
    const int t = threadIdx.x;
#pragma unroll
    for (int i = 0; i < SCAN_IPT; ++i) {
        const int q = i * SCAN_BS + t;
        tile[scan_pad(q)] = b0 + q < n ? (Tout)in[b0 + q] : (Tout)0;
    }
    __syncthreads();
    long long v[SCAN_IPT];
    long long loc = 0;
#pragma unroll
    for (int i = 0; i < SCAN_IPT; ++i) {
        v[i] = (long long)tile[scan_pad(t * SCAN_IPT + i)];
        loc += v[i];
    }
    long long total;
    long long pre = group_excl_scan<SCAN_BS>(loc, red, total) + bsum[blockIdx.x];
#pragma unroll
    for (int i = 0; i < SCAN_IPT; ++i) {
        tile[scan_pad(t * SCAN_IPT + i)] = (Tout)pre;
        pre += v[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < SCAN_IPT; ++i) {
        const int q = i * SCAN_BS + t;
        if (b0 + q < n)
            out[b0 + q] = tile[scan_pad(q)];
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
//
// Element-parallel, so a row of any length costs the same per element (R-MAT's hub rows
// made a row-per-group conversion wait on its longest row):
//   1. k_bmark:  per row, validate the row pointers and mark the row's first element in a
//                bitmap (only the bitmap is cleared beforehand, not the flags);
//   2. k_bhead:  per element, validate the column and sortedness and flag tile heads
//                (first element of a row, or tile column different from the left neighbour);
//   3. exclusive scan of the flags: the tile index of every head;
//   4. k_bfill:  per head, write its tile column and OR the bits of its run (k_bcarry
//                adds the parts of runs that cross a wave);
//      k_btptr:  per row, tptr[j] = tile index at the row's first element.
// Reference: compute_tile64_nums / compute_tile64_col / compute_tile64_mask_num,
// src/compressed/compressed_mask64.cuh:61-342.

__global__ __launch_bounds__(256) void k_bmark(int64_t K, int64_t nnz, const int *__restrict__ rp,
                                               unsigned *__restrict__ starts, DevStats *st)
{
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    int len = 0;
    if (j < K) {
        const int s = rp[j], e = rp[j + 1];
        if (s < 0 || e < s || (int64_t)e > nnz) {
            atomicOr(&st->err, ERR_ROWPTR);
        } else {
            len = e - s;
            if (e > s)
                atomicOr(&starts[s >> 5], 1u << (s & 31));
        }
    }
    // longest row, counted only from LONG_TILES on (all the kernels need to know): no
    // atomic at all for inputs without long rows, one per wave that has one otherwise
    if (len >= LONG_TILES) {
        const int m = wave_max<int>(len);
        if (len == m)
            atomicMax(&st->max_brow, len);
    }
}

__global__ __launch_bounds__(256) void k_bhead(int64_t nnz, int64_t N, const int *__restrict__ col,
                                               const unsigned *__restrict__ starts, int *__restrict__ flag, DevStats *st)
{
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    int bad = 0;
    if (k < nnz) {
        const int c = col[k];
        if ((unsigned)c >= (unsigned long long)N)
            bad |= ERR_BCOL;
        const bool start = (starts[k >> 5] >> (k & 31)) & 1u;
        int head = 1;
        if (!start && k > 0) {
            const int pc = col[k - 1];
            if (pc > c)
                bad |= ERR_UNSORTED;
            head = (pc >> 6) != (c >> 6);
        }
        flag[k] = head;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        bad |= __shfl_xor(bad, o, 64);
    if (bad && (threadIdx.x & 63) == 0)
        atomicOr(&st->err, bad);
}

// The bits of a run are OR-ed across the wave with shuffles (a segmented suffix OR: lane
// i gathers [i, end of its run within the wave) in six doubling steps).  A run that goes
// on past the wave's last lane is finished by k_bcarry: every wave whose first element is
// not a head leaves the OR of its lanes up to its first head in part[wave].
__global__ __launch_bounds__(256) void k_bfill(int64_t nnz, const int *__restrict__ col, const int *__restrict__ flag,
                                               const int *__restrict__ tix, int *__restrict__ tcol, u64 *__restrict__ tmask,
                                               u64 *__restrict__ part)
{
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool valid = k < nnz;
    const int c = valid ? col[k] : 0;
    const bool head = !valid || flag[k] != 0; // past the end counts as a head: runs stop there
    const u64 H = __ballot(head);
    // first head after this lane within the wave (64: none)
    const u64 after = lane < 63 ? (H >> (lane + 1)) : 0ull;
    const int send = after ? lane + 1 + (int)__builtin_ctzll(after) : 64;
    u64 acc = valid ? (1ull << (c & 63)) : 0ull;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const u64 other = __shfl_down(acc, o, 64);
        if (lane + o < send)
            acc |= other;
    }
    if (lane == 0 && !head)
        part[k >> 6] = acc;
    if (!valid || !head)
        return;
    const int t = tix[k];
    tcol[t] = c >> 6;
    tmask[t] = acc;
}

// Per wave of k_bfill whose first element continues a run: OR its part into that run's
// tile (number tix - 1: the heads before it, less one).  Several waves may continue one
// run (a tile with repeated columns), hence the atomic.
__global__ __launch_bounds__(256) void k_bcarry(int64_t nnz, const int *__restrict__ flag, const int *__restrict__ tix,
                                                const u64 *__restrict__ part, u64 *__restrict__ tmask)
{
    const int64_t w = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t k = w << 6;
    if (k >= nnz || flag[k] != 0)
        return;
    atomicOr((unsigned long long *)&tmask[tix[k] - 1], (unsigned long long)part[w]);
}

__global__ __launch_bounds__(256) void k_btptr(int64_t K, int64_t nnz, const int *__restrict__ rp,
                                               const int *__restrict__ col, const int *__restrict__ tix,
                                               int *__restrict__ tptr, BRow *__restrict__ brow, DevStats *st)
{
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j > K)
        return;
    // a row pointer outside [0, nnz] has already raised ERR_ROWPTR in k_bmark; keep
    // the reads in bounds until the host sees the error
    auto clamp = [&](int r) { return r < 0 ? 0 : (r > nnz ? (int)nnz : r); };
    const int r = clamp(rp[j]);
    tptr[j] = tix[r];
    if (j == K) {
        st->ntiles_b = tix[r];
    } else if (brow) {
        const int r1 = clamp(rp[j + 1]);
        BRow q = {0, 0, 0, -1};
        if (r1 > r) {
            q.len = r1 - r;
            q.ntiles = tix[r1] - tix[r];
            q.first = col[r] >> 6; // rows are sorted: their first and last tiles
            q.last = col[r1 - 1] >> 6;
        }
        brow[j] = q;
    }
}

size_t btiles_ws_bytes(int64_t nnz)
{
    return (size_t)(2 * nnz + 2) * sizeof(int) + (size_t)(nnz / 64 + 1) * sizeof(u64) +
           (size_t)(nnz / 32 + 1) * sizeof(unsigned);
}

hipError_t launch_btiles(int64_t K, int64_t N, int64_t nnz, const int *rp, const int *col, int *ws, int *tptr,
                         int *tcol, u64 *tmask, BRow *brow, void *scan_ws, DevStats *st, int num_cus, hipStream_t s)
{
    (void)num_cus;
    int *flag = ws;           // nnz + 1
    int *tix = ws + nnz + 1;  // nnz + 1
    u64 *part = (u64 *)(ws + 2 * nnz + 2); // nnz / 64 + 1 (8-byte aligned: 2 nnz + 2 ints before it)
    unsigned *starts = (unsigned *)(part + nnz / 64 + 1); // nnz / 32 + 1 bits of row starts
    const unsigned gk = (unsigned)((K + 256) / 256);
    const unsigned ge = (unsigned)((nnz + 255) / 256 > 0 ? (nnz + 255) / 256 : 1);
    hipError_t e = launch_fill_zero(starts, (size_t)(nnz / 32 + 1) * sizeof(unsigned), s);
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_bmark, dim3(gk), dim3(256), 0, s, K, nnz, rp, starts, st);
    hipLaunchKernelGGL(k_bhead, dim3(ge), dim3(256), 0, s, nnz, N, col, starts, flag, st);
    e = scan_excl<int, int>(flag, nnz, tix, scan_ws, nullptr, s);
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_bfill, dim3(ge), dim3(256), 0, s, nnz, col, flag, tix, tcol, tmask, part);
    hipLaunchKernelGGL(k_bcarry, dim3((unsigned)((nnz / 64 + 256) / 256)), dim3(256), 0, s, nnz, flag, tix, part, tmask);
    hipLaunchKernelGGL(k_btptr, dim3(gk), dim3(256), 0, s, K, nnz, rp, col, tix, tptr, brow, st);
    return hipGetLastError();
}

// ---- row analysis of A ---------------------------------------------------------------

// Per row i of A: products sum_j nnz(B_j), tile-products sum_j tiles(B_j), and the span
// [rmin, rmax] of tile columns of row i of C (rows of B are sorted, so each B row's
// first and last tile bound it).  Reference: k_compute_flop, src/small/mask.cuh:14-51.
// One group of G lanes per row, one row per group (no grid-stride loop, so the latency
// of the dependent loads Acol -> Brp/tptr -> tcol is paid once per row in parallel).
// Totals are summed per bin by k_bin; only the rare error bits use a global atomic.
constexpr int LONG_A = 2048; // rows of A longer than this are analysed by a whole workgroup

template <int G>
__global__ __launch_bounds__(256) void k_row_info(int64_t M, int64_t K, int64_t nnzA, const int *__restrict__ Arp,
                                                  const int *__restrict__ Acol, const BRow *__restrict__ brow,
                                                  long long *__restrict__ rprod, int *__restrict__ rtp,
                                                  int *__restrict__ rmin, int *__restrict__ rmax, int *__restrict__ rnnz,
                                                  int *__restrict__ longrows, DevStats *st)
{
    constexpr int RPB = 256 / G;
    const int g = threadIdx.x / G, l = threadIdx.x & (G - 1);
    const int64_t i = (int64_t)blockIdx.x * RPB + g;
    if (i >= M)
        return;
    int bad = 0;
    const int s = Arp[i], e = Arp[i + 1];
    long long p = 0;
    int tp = 0, mn = INT_MAX, mx = -1;
    if (s < 0 || e < s || (int64_t)e > nnzA) {
        bad |= ERR_ROWPTR;
    } else if (e - s > LONG_A) {
        if (l == 0)
            longrows[atomicAdd(&st->nlong, 1)] = (int)i;
        return;
    } else {
        for (int k = s + l; k < e; k += G) {
            const int j = Acol[k];
            if ((unsigned)j >= (unsigned long long)K) {
                bad |= ERR_ACOL;
                continue;
            }
            const BRow q = brow[j];
            if (q.len > 0) {
                p += q.len;
                tp += q.ntiles;
                mn = q.first < mn ? q.first : mn;
                mx = q.last > mx ? q.last : mx;
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
    }
    if (bad)
        atomicOr(&st->err, bad);
}

// The rows k_row_info set aside (more than LONG_A entries): one 1024-lane workgroup per
// row, grid-stride over the device-side list.
__global__ __launch_bounds__(1024) void k_row_info_long(int64_t K, const int *__restrict__ longrows,
                                                        const int *__restrict__ Arp, const int *__restrict__ Acol,
                                                        const BRow *__restrict__ brow, long long *__restrict__ rprod,
                                                        int *__restrict__ rtp, int *__restrict__ rmin,
                                                        int *__restrict__ rmax, int *__restrict__ rnnz, DevStats *st)
{
    __shared__ long long s_p[16];
    __shared__ int s_tp[16], s_mn[16], s_mx[16], s_bad[16];
    const int n = st->nlong;
    const int w = threadIdx.x >> 6;
    for (int idx = blockIdx.x; idx < n; idx += gridDim.x) {
        const int i = longrows[idx];
        const int s = Arp[i], e = Arp[i + 1];
        long long p = 0;
        int tp = 0, mn = INT_MAX, mx = -1, bad = 0;
        for (int k = s + threadIdx.x; k < e; k += 1024) {
            const int j = Acol[k];
            if ((unsigned)j >= (unsigned long long)K) {
                bad |= ERR_ACOL;
                continue;
            }
            const BRow q = brow[j];
            if (q.len > 0) {
                p += q.len;
                tp += q.ntiles;
                mn = q.first < mn ? q.first : mn;
                mx = q.last > mx ? q.last : mx;
            }
        }
        p = group_sum<64>(p, (long long *)nullptr);
        tp = group_sum<64>(tp, (int *)nullptr);
        mn = group_min<64>(mn);
        mx = group_max<64>(mx);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1)
            bad |= __shfl_xor(bad, o, 64);
        if ((threadIdx.x & 63) == 0) {
            s_p[w] = p;
            s_tp[w] = tp;
            s_mn[w] = mn;
            s_mx[w] = mx;
            s_bad[w] = bad;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int k = 1; k < 16; ++k) {
                p += s_p[k];
                tp += s_tp[k];
                mn = s_mn[k] < mn ? s_mn[k] : mn;
                mx = s_mx[k] > mx ? s_mx[k] : mx;
                bad |= s_bad[k];
            }
            rprod[i] = p;
            rtp[i] = tp;
            rmin[i] = mn;
            rmax[i] = mx;
            if (p == 0)
                rnnz[i] = 0;
            if (bad)
                atomicOr(&st->err, bad);
        }
        __syncthreads();
    }
}

hipError_t launch_row_info(int64_t M, int64_t K, int64_t nnzA, const int *Arp, const int *Acol, const BRow *brow,
                           long long *rprod, int *rtp, int *rmin, int *rmax, int *rnnz, int *longrows, DevStats *st,
                           int avg_a, int num_cus, hipStream_t s)
{
#define ROW_INFO(G)                                                                                                    \
    {                                                                                                                  \
        constexpr int RPB = 256 / G;                                                                                   \
        const int64_t blocks = (M + RPB - 1) / RPB;                                                                    \
        hipLaunchKernelGGL((k_row_info<G>), dim3((unsigned)(blocks > 0 ? blocks : 1)), dim3(256), 0, s, M, K, nnzA,    \
                           Arp, Acol, brow, rprod, rtp, rmin, rmax, rnnz, longrows, st);                               \
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
    hipLaunchKernelGGL(k_row_info_long, dim3((unsigned)num_cus), dim3(1024), 0, s, K, longrows, Arp, Acol, brow, rprod,
                       rtp, rmin, rmax, rnnz, st);
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
    // the budget limits only workgroups that pack several rows: a one-row workgroup needs
    // the row's LDS whatever its width
    while (b < NLDS - 1) {
        const int G = 8 << b;
        if (G >= BLOCK || need * (BLOCK / G) <= bp.lds_block_budget)
            break;
        ++b;
    }
    // a one-row workgroup's LDS decides how many fit a CU: widen the row's group until
    // those workgroups hold min_waves waves, to hide the latency of the gathers from B
    while (b < NLDS - 1) {
        const int G = 8 << b;
        if (G < BLOCK || need <= 0 || (long long)(G / 64) * (bp.lds_cu / need) >= bp.min_waves)
            break;
        ++b;
    }
    return b;
}

// One thread per row; appends are aggregated per workgroup in LDS, then one global
// atomic per bin per workgroup reserves the range.  Also sums each bin's work, nnz(A)
// and (numeric) nnz(C) for the per-kernel algorithmic byte counts.  Workgroups of BS
// lanes: 256 while that makes at most four per CU (62K rows then occupy every CU, where
// 1024-lane ones left three quarters idle), 1024 beyond (millions of rows would otherwise
// queue four times as many atomics on the same bin counters).
template <int BS>
__global__ __launch_bounds__(BS) void k_bin(int numeric, int64_t M, const int *__restrict__ Arp,
                                              const long long *__restrict__ rprod, const int *__restrict__ rtp,
                                              const int *__restrict__ rmin, const int *__restrict__ rmax,
                                              const int *__restrict__ rnnz, BinParams bp, int *__restrict__ binrows,
                                              long long *__restrict__ rdir, DevStats *st)
{
    __shared__ int s_cnt[NBINS], s_need[NBINS], s_span[NBINS], s_base[NBINS], s_sparse[NBINS], s_na_max[NBINS];
    __shared__ unsigned long long s_work[NBINS], s_nc[NBINS], s_na[NBINS];
    __shared__ unsigned long long s_dir, s_dir_base;
    if (threadIdx.x < NBINS) {
        s_cnt[threadIdx.x] = 0;
        s_need[threadIdx.x] = 0;
        s_span[threadIdx.x] = 0;
        s_work[threadIdx.x] = 0;
        s_nc[threadIdx.x] = 0;
        s_na[threadIdx.x] = 0;
        s_sparse[threadIdx.x] = 0;
        s_na_max[threadIdx.x] = 0;
    }
    if (threadIdx.x == 0)
        s_dir = 0;
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
    const bool with_long = st->max_brow >= LONG_TILES;
    int b = -1, S = 0, nA = 0, nc = 0, sp = 0;
    bool wide = false;
    long long work = 0, need = 0, E = 0;
    if (i < M && rprod[i] > 0) {
        S = rmax[i] - rmin[i] + 1;
        nA = Arp[i + 1] - Arp[i];
        const int tp = rtp[i];
        if (numeric) {
            nc = rnnz[i];
            need = num_need(S, nc, nA, tp, with_long);
            work = rprod[i];
        } else {
            need = sym_need(S, nA, tp, with_long);
            work = tp;
        }
        b = choose_bin(work, need, bp);
        sp = use_sparse(S, tp) ? 1 : 0;
        if (!numeric) {
            // the numeric need with nnz(C_i) <= min(products, 64 * span)
            const long long nub = rprod[i] < 64LL * S ? rprod[i] : 64LL * S;
            wide = num_need(S, nub, nA, tp, with_long) > bp.lds_row_max;
        }
        // directory entries kept for the numeric phase: the row's form in an LDS bin, the
        // flat span for rows the symbolic phase handles in global memory
        if (!numeric)
            E = b < BIN_GLOBAL ? dir_entries(S, tp) : S;
    }
    // Aggregate per wave first: the lanes that chose one bin (neighbouring rows usually
    // share it) add their counts and sums with one LDS atomic per bin per wave, instead
    // of one per lane on the same LDS word.
    const int lane = threadIdx.x & 63;
    const u64 wides = __ballot(wide);
    if (lane == 0 && wides)
        atomicAdd(&st->num_wide_hint, (unsigned long long)__popcll(wides));
    // nonzeros of the wide rows whose work is split into several items (their compensated
    // FP32 partial sums meet in a scratch buffer of this size)
    if (numeric) {
        const unsigned long long sn =
            wave_sum<unsigned long long>(b == BIN_GLOBAL && nA > wide_split(nc) ? (unsigned long long)nc : 0ull);
        if (lane == 0 && sn)
            atomicAdd(&st->wide_split_nnz, sn);
        const unsigned long long ss = wave_sum<unsigned long long>(b == BIN_GLOBAL ? (unsigned long long)S : 0ull);
        if (lane == 0 && ss)
            atomicAdd(&st->wide_span, ss);
    }
    int pos = 0;
    long long dloc = -1;
    u64 pending = __ballot(b >= 0);
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const int bb = __shfl(b, leader, 64);
        const u64 same = __ballot(b == bb);
        pending &= ~same;
        const bool mine = b == bb;
        const unsigned long long wsum = wave_sum<unsigned long long>(mine ? (unsigned long long)work : 0ull);
        const unsigned long long ncs = wave_sum<unsigned long long>(mine ? (unsigned long long)nc : 0ull);
        const unsigned long long nas = wave_sum<unsigned long long>(mine ? (unsigned long long)nA : 0ull);
        const int needm = wave_max<int>(mine ? (int)(need < INT_MAX ? need : INT_MAX) : 0);
        const int spanm = wave_max<int>(mine ? S : 0);
        const int nam = wave_max<int>(mine ? nA : 0);
        const int sps = wave_sum<int>(mine ? sp : 0);
        const unsigned long long dincl = wave_incl_scan<unsigned long long>(mine ? (unsigned long long)E : 0ull);
        const unsigned long long dtot = __shfl(dincl, 63, 64);
        int base = 0;
        unsigned long long dbase = 0;
        if (lane == leader) {
            base = atomicAdd(&s_cnt[bb], __popcll(same));
            atomicMax(&s_need[bb], needm);
            atomicMax(&s_span[bb], spanm);
            atomicMax(&s_na_max[bb], nam);
            atomicAdd(&s_work[bb], wsum);
            atomicAdd(&s_nc[bb], ncs);
            atomicAdd(&s_na[bb], nas);
            if (sps)
                atomicAdd(&s_sparse[bb], sps);
            if (dtot)
                dbase = atomicAdd(&s_dir, dtot);
        }
        base = __shfl(base, leader, 64);
        dbase = __shfl(dbase, leader, 64);
        if (mine) {
            pos = base + mbcnt(same);
            if (E > 0)
                dloc = (long long)(dbase + dincl) - E;
        }
    }
    __syncthreads();
    BinStats *bs = numeric ? &st->num : &st->sym;
    if (threadIdx.x < NBINS) {
        const int c = s_cnt[threadIdx.x];
        s_base[threadIdx.x] = c ? atomicAdd(&bs->count[threadIdx.x], c) : 0;
        if (c) {
            atomicMax(&bs->max_need[threadIdx.x], s_need[threadIdx.x]);
            atomicMax(&bs->max_span[threadIdx.x], s_span[threadIdx.x]);
            atomicMax(&bs->max_na[threadIdx.x], s_na_max[threadIdx.x]);
            atomicAdd(&bs->work[threadIdx.x], s_work[threadIdx.x]);
            atomicAdd(&bs->nnz_c[threadIdx.x], s_nc[threadIdx.x]);
            atomicAdd(&bs->nnz_a[threadIdx.x], s_na[threadIdx.x]);
            if (s_sparse[threadIdx.x])
                atomicAdd(&bs->nsparse[threadIdx.x], s_sparse[threadIdx.x]);
        }
    }
    if (!numeric && threadIdx.x == NBINS)
        s_dir_base = s_dir ? atomicAdd(&st->dir_total, s_dir) : 0;
    __syncthreads();
    if (b >= 0)
        binrows[(int64_t)b * M + s_base[b] + pos] = (int)i;
    if (!numeric && i < M)
        rdir[i] = dloc >= 0 ? (long long)s_dir_base + dloc : -1;
}

hipError_t launch_bin(int numeric, int64_t M, const int *Arp, const long long *rprod, const int *rtp, const int *rmin,
                      const int *rmax, const int *rnnz, BinParams bp, int *binrows, long long *rdir, DevStats *st,
                      int num_cus, hipStream_t s)
{
    if (M <= 0)
        return hipSuccess;
    const int64_t small = (M + 255) / 256;
    if (small <= 4LL * num_cus)
        hipLaunchKernelGGL(k_bin<256>, dim3((unsigned)small), dim3(256), 0, s, numeric, M, Arp, rprod, rtp, rmin, rmax,
                           rnnz, bp, binrows, rdir, st);
    else
        hipLaunchKernelGGL(k_bin<1024>, dim3((unsigned)((M + 1023) / 1024)), dim3(1024), 0, s, numeric, M, Arp, rprod,
                           rtp, rmin, rmax, rnnz, bp, binrows, rdir, st);
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

// ---- row panels and row sums ------------------------------------------------------------

__global__ __launch_bounds__(256) void k_rowptr_rebase(const int *__restrict__ in, int64_t n, int *__restrict__ out)
{
    const int base = in[0];
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride)
        out[i] = in[i] - base;
}

hipError_t launch_rowptr_rebase(const int *in, int64_t n, int *out, hipStream_t s)
{
    int64_t blocks = (n + 255) / 256;
    blocks = blocks > 8192 ? 8192 : (blocks < 1 ? 1 : blocks);
    hipLaunchKernelGGL(k_rowptr_rebase, dim3((unsigned)blocks), dim3(256), 0, s, in, n, out);
    return hipGetLastError();
}

// One wave per row: sum and sum of magnitudes of the row's values.
template <typename RP, typename V>
__global__ __launch_bounds__(256) void k_row_sums(int64_t M, const RP *__restrict__ rp, const V *__restrict__ val,
                                                  double *__restrict__ sum, double *__restrict__ asum)
{
    const int lane = threadIdx.x & 63;
    const int64_t stride = (int64_t)gridDim.x * 4;
    for (int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); i < M; i += stride) {
        double s = 0.0, a = 0.0;
        for (long long k = (long long)rp[i] + lane; k < (long long)rp[i + 1]; k += 64) {
            const double v = (double)val[k];
            s += v;
            a += v < 0 ? -v : v;
        }
        s = wave_sum<double>(s);
        a = wave_sum<double>(a);
        if (lane == 0) {
            sum[i] = s;
            asum[i] = a;
        }
    }
}

hipError_t launch_row_sums(int64_t M, const void *rp, int rp64, const void *val, int val32, double *sum, double *asum,
                           hipStream_t s)
{
    int64_t blocks = (M + 3) / 4;
    blocks = blocks > 16384 ? 16384 : (blocks < 1 ? 1 : blocks);
    const dim3 g((unsigned)blocks), b(256);
    if (rp64 && val32)
        hipLaunchKernelGGL((k_row_sums<long long, float>), g, b, 0, s, M, (const long long *)rp, (const float *)val, sum, asum);
    else if (rp64)
        hipLaunchKernelGGL((k_row_sums<long long, double>), g, b, 0, s, M, (const long long *)rp, (const double *)val, sum,
                           asum);
    else if (val32)
        hipLaunchKernelGGL((k_row_sums<int, float>), g, b, 0, s, M, (const int *)rp, (const float *)val, sum, asum);
    else
        hipLaunchKernelGGL((k_row_sums<int, double>), g, b, 0, s, M, (const int *)rp, (const double *)val, sum, asum);
    return hipGetLastError();
}

// ---- batched form: stacking the instances and splitting C ----------------------------------

// Largest i in [0, count) with off[i] + i * step <= x (off nondecreasing, off[0] = 0 <= x):
// the instance whose range holds x, empty instances skipped.
DEVI int find_instance(const long long *__restrict__ off, int count, long long x, int step)
{
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] + (long long)mid * step <= x)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_batch_rowptr(int which, int count, const BatchInst *__restrict__ inst,
                                                      const long long *__restrict__ roff,
                                                      const long long *__restrict__ eoff, int64_t nrows, int64_t nnz,
                                                      int *__restrict__ out, DevStats *st)
{
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x; g <= nrows; g += stride) {
        if (g == nrows) {
            out[g] = (int)nnz;
            continue;
        }
        const int i = find_instance(roff, count, g, 0);
        const BatchInst I = inst[i];
        const int *rp = which ? I.Brp : I.Arp;
        const long long r = g - roff[i];
        const long long e0 = eoff[i];
        out[g] = (int)(e0 + rp[r]);
        if (r == 0) {
            const int rows = which ? I.K : I.M;
            if (rp[0] != 0 || (long long)rp[rows] != eoff[i + 1] - e0)
                atomicOr(&st->err, ERR_ROWPTR);
        }
    }
}

// Entries are stacked in chunks of STACK_CHUNK per workgroup: one binary search per chunk
// for its first instance, then each thread steps forward from there (a chunk spans few
// instances), so the lookup costs no chain of dependent loads per entry.
constexpr int STACK_CHUNK = 4096;

__global__ __launch_bounds__(256) void k_batch_entries(int which, int count, const BatchInst *__restrict__ inst,
                                                       const long long *__restrict__ eoff,
                                                       const long long *__restrict__ koff, int64_t nnz, int val32,
                                                       int *__restrict__ col, void *__restrict__ val, DevStats *st)
{
    __shared__ int s_i0;
    int bad = 0;
    for (int64_t base = (int64_t)blockIdx.x * STACK_CHUNK; base < nnz; base += (int64_t)gridDim.x * STACK_CHUNK) {
        if (threadIdx.x == 0)
            s_i0 = find_instance(eoff, count, base, 0);
        __syncthreads();
        int i = s_i0;
        for (int k = threadIdx.x; k < STACK_CHUNK; k += 256) {
            const int64_t e = base + k;
            if (e >= nnz)
                break;
            while (eoff[i + 1] <= e) // ends: e < nnz = eoff[count]
                ++i;
            const BatchInst &I = inst[i];
            const long long q = e - eoff[i];
            const int c = which ? I.Bcol[q] : I.Acol[q];
            const int lim = which ? I.N : I.K;
            int cs = c;
            if (c < 0 || c >= lim) {
                bad |= which ? ERR_BCOL : ERR_ACOL;
                cs = 0;
            }
            col[e] = which ? cs : (int)(cs + koff[i]);
            if (val) {
                const void *src = which ? I.Bval : I.Aval;
                if (val32)
                    ((float *)val)[e] = ((const float *)src)[q];
                else
                    ((double *)val)[e] = ((const double *)src)[q];
            }
        }
        __syncthreads(); // s_i0 is rewritten for the next chunk
    }
    if (bad)
        atomicOr(&st->err, bad);
}

__global__ __launch_bounds__(256) void k_batch_split(int count, const long long *__restrict__ mo,
                                                     const long long *__restrict__ crp, void *__restrict__ out,
                                                     int rp64, long long *__restrict__ bases)
{
    const int64_t n = mo[count] + count; // sum over instances of M_i + 1
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < n; q += stride) {
        const int i = find_instance(mo, count, q, 1);
        const long long r0 = mo[i];
        const long long v = crp[r0 + (q - r0 - i)] - crp[r0];
        if (rp64)
            ((long long *)out)[q] = v;
        else
            ((int *)out)[q] = (int)v;
    }
    if (blockIdx.x == 0)
        for (int i = threadIdx.x; i <= count; i += 256)
            bases[i] = crp[mo[i]];
}

static unsigned grid_for(int64_t n)
{
    int64_t blocks = (n + 255) / 256;
    if (blocks > 16384)
        blocks = 16384;
    return (unsigned)(blocks < 1 ? 1 : blocks);
}

hipError_t launch_batch_rowptr(int which, int64_t count, const BatchInst *inst, BatchOffsets off, int64_t nrows,
                               int64_t nnz, int *out, DevStats *st, hipStream_t s)
{
    hipLaunchKernelGGL(k_batch_rowptr, dim3(grid_for(nrows + 1)), dim3(256), 0, s, which, (int)count, inst,
                       which ? off.ko : off.mo, which ? off.eb : off.ea, nrows, nnz, out, st);
    return hipGetLastError();
}

hipError_t launch_batch_entries(int which, int64_t count, const BatchInst *inst, BatchOffsets off, int64_t nnz,
                                int val32, int *col, void *val, DevStats *st, hipStream_t s)
{
    if (nnz <= 0)
        return hipSuccess;
    int64_t chunks = (nnz + STACK_CHUNK - 1) / STACK_CHUNK;
    if (chunks > 16384)
        chunks = 16384;
    hipLaunchKernelGGL(k_batch_entries, dim3((unsigned)chunks), dim3(256), 0, s, which, (int)count, inst,
                       which ? off.eb : off.ea, off.ko, nnz, val32, col, val, st);
    return hipGetLastError();
}

hipError_t launch_batch_split(int64_t count, BatchOffsets off, const long long *crp, void *out, int rp64,
                              long long *bases, int64_t nout, hipStream_t s)
{
    hipLaunchKernelGGL(k_batch_split, dim3(grid_for(nout)), dim3(256), 0, s, (int)count, off.mo, crp, out, rp64, bases);
    return hipGetLastError();
}

// ---- replayed plans --------------------------------------------------------------------------

__global__ void k_plan_check(DevStats *st, const DevStats *expect, int stage)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    const BinStats &a = stage == 0 ? st->sym : st->num;
    const BinStats &b = stage == 0 ? expect->sym : expect->num;
    bool same = true;
    for (int k = 0; k < NBINS; ++k)
        same = same && a.count[k] == b.count[k] && a.max_need[k] == b.max_need[k] && a.max_span[k] == b.max_span[k] &&
               a.nsparse[k] == b.nsparse[k] && a.max_na[k] == b.max_na[k];
    if (stage == 0)
        same = same && st->dir_total == expect->dir_total && st->ntiles_b == expect->ntiles_b &&
               st->max_brow == expect->max_brow;
    else
        same = same && st->nnz_c == expect->nnz_c && st->wide_split_nnz == expect->wide_split_nnz;
    if (!same)
        st->err |= ERR_PLAN;
    if (st->err) {
        BinStats &w = stage == 0 ? st->sym : st->num;
        for (int k = 0; k < NBINS; ++k)
            w.count[k] = 0;
    }
}

hipError_t launch_plan_check(DevStats *st, const DevStats *expect, int stage, hipStream_t s)
{
    hipLaunchKernelGGL(k_plan_check, dim3(1), dim3(64), 0, s, st, expect, stage);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_plan_bases(const long long *__restrict__ got, const long long *__restrict__ want,
                                                    int64_t n, DevStats *st)
{
    bool diff = false;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        diff = diff || got[i] != want[i];
    if (diff)
        atomicOr(&st->err, ERR_PLAN);
}

hipError_t launch_plan_bases(const long long *got, const long long *want, int64_t n, DevStats *st, hipStream_t s)
{
    hipLaunchKernelGGL(k_plan_bases, dim3(grid_for(n)), dim3(256), 0, s, got, want, n, st);
    return hipGetLastError();
}

} // namespace hsmu
