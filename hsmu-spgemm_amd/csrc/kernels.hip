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
// One group of G lanes per row, one row per group (no grid-stride loop, so the latency
// of the dependent loads Acol -> Brp/tptr -> tcol is paid once per row in parallel).
// Totals are summed per bin by k_bin; only the rare error bits use a global atomic.
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
    const int64_t i = (int64_t)blockIdx.x * RPB + g;
    if (i >= M)
        return;
    int bad = 0;
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
    }
    if (bad)
        atomicOr(&st->err, bad);
}

hipError_t launch_row_info(int64_t M, int64_t K, int64_t nnzA, const int *Arp, const int *Acol, const int *Brp,
                           const int *tptr, const int *tcol, long long *rprod, int *rtp, int *rmin, int *rmax, int *rnnz,
                           DevStats *st, int avg_a, hipStream_t s)
{
#define ROW_INFO(G)                                                                                                    \
    {                                                                                                                  \
        constexpr int RPB = 256 / G;                                                                                   \
        const int64_t blocks = (M + RPB - 1) / RPB;                                                                    \
        hipLaunchKernelGGL((k_row_info<G>), dim3((unsigned)(blocks > 0 ? blocks : 1)), dim3(256), 0, s, M, K, nnzA,    \
                           Arp, Acol, Brp, tptr, tcol, rprod, rtp, rmin, rmax, rnnz, st);                              \
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
                                              long long *__restrict__ rdir, DevStats *st)
{
    __shared__ int s_cnt[NBINS], s_need[NBINS], s_span[NBINS], s_base[NBINS];
    __shared__ unsigned long long s_work[NBINS], s_nc[NBINS], s_na[NBINS];
    __shared__ unsigned long long s_dir, s_dir_base;
    if (threadIdx.x < NBINS) {
        s_cnt[threadIdx.x] = 0;
        s_need[threadIdx.x] = 0;
        s_span[threadIdx.x] = 0;
        s_work[threadIdx.x] = 0;
        s_nc[threadIdx.x] = 0;
        s_na[threadIdx.x] = 0;
    }
    if (threadIdx.x == 0)
        s_dir = 0;
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * 1024 + threadIdx.x;
    int b = -1, pos = 0;
    long long dloc = -1;
    if (i < M && rprod[i] > 0) {
        const int S = rmax[i] - rmin[i] + 1;
        const int nA = Arp[i + 1] - Arp[i];
        long long work, need;
        int nc = 0;
        if (numeric) {
            nc = rnnz[i];
            need = num_need(S, nc, nA);
            work = rprod[i];
        } else {
            need = sym_need(S, nA);
            work = rtp[i];
        }
        b = choose_bin(work, need, bp);
        pos = atomicAdd(&s_cnt[b], 1);
        atomicMax(&s_need[b], (int)(need < INT_MAX ? need : INT_MAX));
        atomicMax(&s_span[b], S);
        atomicAdd(&s_work[b], (unsigned long long)work);
        atomicAdd(&s_nc[b], (unsigned long long)nc);
        atomicAdd(&s_na[b], (unsigned long long)nA);
        if (!numeric && b < BIN_GLOBAL)
            dloc = (long long)atomicAdd(&s_dir, (unsigned long long)S);
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
                      hipStream_t s)
{
    const int64_t blocks = (M + 1023) / 1024;
    if (blocks == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_bin, dim3((unsigned)blocks), dim3(1024), 0, s, numeric, M, Arp, rprod, rtp, rmin, rmax, rnnz,
                       bp, binrows, rdir, st);
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

} // namespace hsmu
