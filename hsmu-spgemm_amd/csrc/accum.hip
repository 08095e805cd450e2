// accum.hip -- symbolic and numeric row kernels of the HSMU SpGEMM path for gfx950.
//
// Symbolic: nnz(C_i) from the OR of B's 64-column tile masks over row i of A, in a dense
// mask directory over the row's tile span (reference: compressed_tilePtr_symbolic.cuh
// and compressed_tileOR.cuh, two hash passes over tiles there; one pass and no hashing
// here).  Numeric: the HSMU accumulator -- exactly nnz(C_i) FP64 slots in LDS, the slot
// of column c being off[t] + popcount(mask[t] & below(c & 63)), t = (c >> 6) - rmin,
// i.e. c's rank among C's sorted columns found in O(1) (reference: binary search over
// the sorted column list, src/compressed/compressed_Cval.cuh:5-69 and
// src/small/numeric.cuh:233-253).  The numeric kernel also writes C's sorted columns,
// replacing the reference's rank-sort pass (compressed_Ccol_symbolic.cuh:4-63).
//
// LDS kernels first stage the row's A entries in LDS (B row start, tile start, a_ij, and
// prefix sums of nnz(B_j) and tiles(B_j)), then walk the row's products (or
// tile-products) as one flat sequence: each wave takes 64 consecutive items, so lanes
// load consecutive elements of B (coalesced) and no lane waits on a per-entry chain of
// dependent loads.  Rows whose slice does not fit LDS use the global-memory kernels.

#include <hip/hip_runtime.h>

#include <climits>

#include "device_common.h"
#include "hsmu_internal.h"

namespace hsmu {

// Stage the row's A entries a0 .. a0+nA-1 into LDS: first tile and tile count of B_j,
// and (numeric) first element and length of B_j and a_ij.  One parallel step per row,
// so the dependent loads Acol -> tptr/Brp are paid once per row rather than once per
// entry inside the accumulation loops.
template <int G, bool NUMERIC>
DEVI void stage_row(const Operands &op, int a0, int nA, int *t0, int *tlen, int *b0, int *blen, double *av)
{
    const int l = threadIdx.x & (G - 1);
    for (int k = l; k < nA; k += G) {
        const int j = op.Acol[a0 + k];
        if (t0) {
            const int tt = op.tptr[j];
            t0[k] = tt;
            tlen[k] = op.tptr[j + 1] - tt;
        }
        if constexpr (NUMERIC) {
            const int bb = op.Brp[j];
            b0[k] = bb;
            blen[k] = op.Brp[j + 1] - bb;
            av[k] = op.val32 ? (double)((const float *)op.Aval)[a0 + k] : ((const double *)op.Aval)[a0 + k];
        }
    }
}

constexpr int UNROLL = 4;

// Compact index of tile t (relative to rmin) in a two-level directory: the rank of t
// among the row's non-empty tiles.
DEVI unsigned fine_index(const u64 *cm, const unsigned *coff, int t)
{
    const int w = t >> 6;
    return coff[w] + (unsigned)__popcll(cm[w] & below(t & 63));
}

// Walks over the row's tile-products (every tile of every B_j of the row):
//   OR_FLAT    dm[t] |= mask                  flat directory
//   OR_COARSE  cm[t >> 6] |= bit (t & 63)     two-level, pass 1: which tiles are non-empty
//   OR_FINE    fm[fine_index(t)] |= mask      two-level, pass 2
enum { OR_FLAT, OR_COARSE, OR_FINE };

// Sub-groups of 2^tshift lanes take one entry j each and walk the tiles of B_j, UNROLL
// tiles per lane per iteration with all loads issued before the ds_or_b64s.
template <int G, int MODE>
DEVI void or_tiles(const Operands &op, const int *t0, const int *tlen, int nA, u64 *dm, const unsigned *coff, u64 *fm,
                   int mn, int tshift)
{
    const int l = threadIdx.x & (G - 1);
    const int SG = 1 << tshift, sl = l & (SG - 1), NSG = G >> tshift;
    for (int e = l >> tshift; e < nA; e += NSG) {
        const int tb = t0[e], te = tb + tlen[e];
        for (int t = tb + sl; t < te; t += UNROLL * SG) {
            int c[UNROLL];
            u64 m[UNROLL];
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) {
                const int tu = t + u * SG;
                c[u] = -1;
                m[u] = 0;
                if (tu < te) {
                    c[u] = op.tcol[tu];
                    if constexpr (MODE != OR_COARSE)
                        m[u] = op.tmask[tu];
                }
            }
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) {
                if (c[u] >= 0) {
                    const int tt = c[u] - mn;
                    if constexpr (MODE == OR_FLAT)
                        atomicOr(&dm[tt], m[u]);
                    else if constexpr (MODE == OR_COARSE)
                        atomicOr(&dm[tt >> 6], 1ull << (tt & 63));
                    else
                        atomicOr(&fm[fine_index(dm, coff, tt)], m[u]);
                }
            }
        }
    }
}

// Each lane scans a contiguous chunk of the n directory entries: the group's exclusive
// prefix of the popcounts at the chunk start, the group total in tot.  Every thread of the
// workgroup must call it (G > 64 synchronises the workgroup).
template <int G>
DEVI unsigned scan_popcounts(const u64 *m, int n, unsigned *redu, unsigned &tot, int &ta, int &tb)
{
    const int l = threadIdx.x & (G - 1);
    const int chunk = (n + G - 1) / G;
    ta = l * chunk;
    tb = ta + chunk < n ? ta + chunk : n;
    unsigned loc = 0;
    for (int t = ta; t < tb; ++t)
        loc += (unsigned)__popcll(m[t]);
    return group_excl_scan<G>(loc, redu, tot);
}

// ---- numeric forms ---------------------------------------------------------------------
//
// A slot is 8 bytes in both forms, so the LDS layout and the bins are the same.
// PrecF64 (plain): FP64 values, slot = one double, ds_add_f64 per product.
// PrecF32C (compensated FP32): FP32 values; the product is split exactly into
// p = fl(a*b) and e = fma(a, b, -p) (TwoProduct); the slot holds a float pair (s, c):
// p is added to s by a compare-and-swap loop (so the stored sum is the VALU's round-to-
// nearest-even fl(s + p)), the exact rounding error of that addition is recomputed
// (TwoSum, Knuth) and err + e go into c with a plain float atomic add.  The result is
// fl(s + c): c carries every rounding error of the products and of the additions into
// s, so the sum is accurate to about one FP32 rounding of the result plus O(n u^2) of
// sum |a*b| (u = 2^-24).
struct PrecF64 {
    typedef double V;
    static DEVI void add(double *vals, unsigned slot, double a, V b) { atomicAdd(&vals[slot], a * b); }
    static DEVI V result(const double *vals, unsigned slot) { return vals[slot]; }
    // global-memory accumulation (rows outside LDS): straight into C
    static DEVI void gadd(V *Cval, float *, long long q, double a, V b) { unsafeAtomicAdd(&Cval[q], a * b); }
};

struct PrecF32C {
    typedef float V;
    // s += p atomically with the sum rounded by the VALU (round to nearest even): a
    // compare-and-swap loop, because the rounding of the memory units' float atomic add
    // does not match, and the TwoSum error below must describe the addition actually
    // stored.  The error of that addition plus e then goes into c, whose own rounding is
    // second order (|c| ~ u |s|).
    template <bool LDS> static DEVI void two_add(float *s, float *c, float p, float e)
    {
        unsigned *si = (unsigned *)s;
        float old = *s;
        float sum;
        while (true) {
            sum = old + p;
            const unsigned want = __float_as_uint(old);
            const unsigned prev = atomicCAS(si, want, __float_as_uint(sum));
            if (prev == want)
                break;
            old = __uint_as_float(prev);
        }
        const float bb = sum - old;
        const float err = (old - (sum - bb)) + (p - bb);
        if constexpr (LDS)
            atomicAdd(c, err + e);
        else
            unsafeAtomicAdd(c, err + e);
    }
    static DEVI void add(double *vals, unsigned slot, double a, V b)
    {
        const float af = (float)a;
        const float pr = __fmul_rn(af, b);
        const float e = __builtin_fmaf(af, b, -pr);
        float *f = (float *)vals + 2 * slot;
        two_add<true>(f, f + 1, pr, e);
    }
    static DEVI V result(const double *vals, unsigned slot)
    {
        const float *f = (const float *)vals + 2 * slot;
        return f[0] + f[1];
    }
    // global-memory accumulation: s in C, c in a compensation array of nnz(C) floats
    static DEVI void gadd(V *Cval, float *comp, long long q, double a, V b)
    {
        const float af = (float)a;
        const float pr = __fmul_rn(af, b);
        const float e = __builtin_fmaf(af, b, -pr);
        two_add<false>(&Cval[q], &comp[q], pr, e);
    }
};

// Slot of column c in the row's accumulator, by directory form.
enum { SLOT_FLAT, SLOT_DENSE, SLOT_TWO_LEVEL };

// Add every product a_ij * b_jc of the row into its slot: sub-groups of 2^bshift lanes
// take one entry each and walk B_j, UNROLL elements per lane per iteration with all
// loads issued before the ds_add_f64s.  fm/foff: the (fine) tile masks and first slots;
// cm/coff: the coarse level of a two-level directory.
template <int G, int FORM, typename P>
DEVI void add_products(const Operands &op, const int *b0, const int *blen, const double *av, int nA, const u64 *cm,
                       const unsigned *coff, const u64 *fm, const unsigned *foff, double *vals, int mn, int bshift)
{
    typedef typename P::V V;
    const V *__restrict__ Bval = (const V *)op.Bval;
    const int l = threadIdx.x & (G - 1);
    const int SG = 1 << bshift, sl = l & (SG - 1), NSG = G >> bshift;
    for (int e = l >> bshift; e < nA; e += NSG) {
        const int kb = b0[e], ke = kb + blen[e];
        const double a = av[e];
        for (int k = kb + sl; k < ke; k += UNROLL * SG) {
            int c[UNROLL];
            V v[UNROLL];
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) {
                const int ku = k + u * SG;
                c[u] = -1;
                v[u] = 0;
                if (ku < ke) {
                    c[u] = op.Bcol[ku];
                    v[u] = Bval[ku];
                }
            }
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) {
                if (c[u] >= 0) {
                    unsigned slot;
                    if constexpr (FORM == SLOT_DENSE) {
                        slot = (unsigned)(c[u] - (mn << 6));
                    } else {
                        const int t = (c[u] >> 6) - mn;
                        const unsigned k = FORM == SLOT_TWO_LEVEL ? fine_index(cm, coff, t) : (unsigned)t;
                        slot = foff[k] + (unsigned)__popcll(fm[k] & below(c[u] & 63));
                    }
                    P::add(vals, slot, a, v[u]);
                }
            }
        }
    }
}

// ---- symbolic ------------------------------------------------------------------------------

// nnz(C_i) for rows kept in LDS: stage the row, OR the tile masks with ds_or_b64 into the
// row's tile directory, popcount.  Flat rows: one pass into a directory over the whole
// span.  Two-level rows: pass 1 marks the non-empty tiles in the coarse bitmap, a scan
// gives their compact indices, pass 2 ORs the masks into the compact fine level.
// two_level = 0 when no row of the launch is two-level (its extra steps are skipped).
// When the directory is kept, it goes out with each entry's exclusive prefix (coarse:
// first compact index; fine or flat: first slot) for the numeric phase.
template <int G>
__global__ __launch_bounds__(G < BLOCK ? BLOCK : G) void k_symbolic_lds(const int *__restrict__ rows,
                                                                       const int *__restrict__ count_p, int slice_bytes,
                                                                       Operands op, int *__restrict__ rnnz, int tshift,
                                                                       int two_level, unsigned *queue)
{
    constexpr int BS = G < BLOCK ? BLOCK : G;
    constexpr int RPB = BS / G;
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ unsigned redu[16];
    __shared__ int grp_s[2];
    const int g = threadIdx.x / G, l = threadIdx.x & (G - 1);
    unsigned char *slice = lds + (size_t)g * slice_bytes;
    const int nrows = *count_p;
    for (GroupWalk gw(queue, (nrows + RPB - 1) / RPB, grp_s); gw.more(); gw.next()) {
        const int idx = gw.grp * RPB + g;
        const int row = idx < nrows ? rows[idx] : -1;
        int S = 0, mn = 0, a0 = 0, nA = 0, tp = 0;
        if (row >= 0) {
            mn = op.rmin[row];
            S = op.rmax[row] - mn + 1;
            a0 = op.Arp[row];
            nA = op.Arp[row + 1] - a0;
            tp = op.rtp[row];
        }
        const bool sparse = use_sparse(S, tp);
        const int W = sparse ? (S + 63) >> 6 : S; // entries of the first level
        const RowLayout L = row_layout(0, S, 0, nA, tp);
        u64 *dm = (u64 *)(slice + L.dm);
        unsigned *doff = (unsigned *)(slice + L.doff);
        int *tlen = (int *)(slice + L.tlen);
        int *t0 = (int *)(slice + L.t0);
        for (int t = l; t < W; t += G)
            dm[t] = 0;
        stage_row<G, false>(op, a0, nA, t0, tlen, nullptr, nullptr, nullptr);
        __syncthreads();
        if (sparse)
            or_tiles<G, OR_COARSE>(op, t0, tlen, nA, dm, nullptr, nullptr, mn, tshift);
        else
            or_tiles<G, OR_FLAT>(op, t0, tlen, nA, dm, nullptr, nullptr, mn, tshift);
        __syncthreads();
        int ta, tb;
        unsigned tot;
        unsigned pre = scan_popcounts<G>(dm, W, redu, tot, ta, tb);
        const long long rd = (row >= 0 && op.dir_mask) ? op.rdir[row] : -1;
        unsigned nnz = tot;
        if (!sparse) {
            if (rd >= 0) {
                for (int t = ta; t < tb; ++t) {
                    const u64 m = dm[t];
                    op.dir_mask[rd + t] = m;
                    op.dir_off[rd + t] = pre;
                    pre += (unsigned)__popcll(m);
                }
            }
        } else {
            for (int t = ta; t < tb; ++t) {
                const u64 m = dm[t];
                doff[t] = pre;
                if (rd >= 0) {
                    op.dir_mask[rd + t] = m;
                    op.dir_off[rd + t] = pre;
                }
                pre += (unsigned)__popcll(m);
            }
        }
        if (two_level) {
            // two-level rows: tot = non-empty tiles; build and count the fine level
            u64 *fm = dm + W;
            __syncthreads();
            if (sparse)
                for (int k = l; k < (int)tot; k += G)
                    fm[k] = 0;
            __syncthreads();
            if (sparse)
                or_tiles<G, OR_FINE>(op, t0, tlen, nA, dm, doff, fm, mn, tshift);
            __syncthreads();
            unsigned tot2;
            unsigned pre2 = scan_popcounts<G>(fm, sparse ? (int)tot : 0, redu, tot2, ta, tb);
            if (sparse) {
                nnz = tot2;
                if (rd >= 0) {
                    for (int k = ta; k < tb; ++k) {
                        const u64 m = fm[k];
                        op.dir_mask[rd + W + k] = m;
                        op.dir_off[rd + W + k] = pre2;
                        pre2 += (unsigned)__popcll(m);
                    }
                }
            }
        }
        if (row >= 0 && l == 0)
            rnnz[row] = (int)nnz;
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

// ---- numeric -------------------------------------------------------------------------------

// The HSMU accumulator in LDS (slice layout: row_layout).  The row's tile directory is
// loaded (kept by the symbolic phase) or rebuilt as there, every product is added with
// ds_add_f64 into its slot, and the row of C is written in column order: for G >= 64 one
// wave per tile, lane b writing column 64*(rmin+t)+b if present (coalesced); for G < 64
// one lane per tile walking its set bits.  Directory forms: flat (slot from the tile's
// entry), dense (one slot per column of the span) and two-level (the tile's entry found
// through the coarse bitmap first); two_level = 0 when no row of the launch uses it.
template <int G, typename P>
__global__ __launch_bounds__(G < BLOCK ? BLOCK : G) void k_numeric_lds(const int *__restrict__ rows,
                                                                      const int *__restrict__ count_p, int slice_bytes,
                                                                      Operands op, const long long *__restrict__ crp,
                                                                      int *__restrict__ Ccol, typename P::V *__restrict__ Cval,
                                                                      int bshift, int tshift, int two_level, int *err,
                                                                      unsigned *queue)
{
    constexpr int BS = G < BLOCK ? BLOCK : G;
    constexpr int RPB = BS / G;
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ unsigned redu[16];
    __shared__ int grp_s[2];
    const int g = threadIdx.x / G, l = threadIdx.x & (G - 1);
    unsigned char *slice = lds + (size_t)g * slice_bytes;
    const int nrows = *count_p;
    for (GroupWalk gw(queue, (nrows + RPB - 1) / RPB, grp_s); gw.more(); gw.next()) {
        const int idx = gw.grp * RPB + g;
        const int row = idx < nrows ? rows[idx] : -1;
        int S = 0, mn = 0, n = 0, a0 = 0, nA = 0, tp = 0;
        long long cb = 0;
        if (row >= 0) {
            mn = op.rmin[row];
            S = op.rmax[row] - mn + 1;
            cb = crp[row];
            n = (int)(crp[row + 1] - cb);
            a0 = op.Arp[row];
            nA = op.Arp[row + 1] - a0;
            tp = op.rtp[row];
        }
        const bool sparse = use_sparse(S, tp);
        const int W = sparse ? (S + 63) >> 6 : S; // entries of the first level
        const RowLayout L = row_layout(1, S, n, nA, tp);
        u64 *dm = (u64 *)(slice + L.dm);
        unsigned *doff = (unsigned *)(slice + L.doff);
        unsigned *ft = (unsigned *)(slice + L.ft);
        double *vals = (double *)(slice + L.vals);
        double *av = (double *)(slice + L.av);
        int *tlen = (int *)(slice + L.tlen);
        int *t0 = (int *)(slice + L.t0);
        int *blen = (int *)(slice + L.blen);
        int *b0 = (int *)(slice + L.b0);
        // fine level: the flat directory itself, or the compact entries after the coarse words
        u64 *fm = sparse ? dm + W : dm;
        unsigned *foff = sparse ? doff + W : doff;
        const bool dense = !sparse && use_dense(S, n);
        const int nv = dense ? 64 * S : n;
        const long long rd = (row >= 0 && op.dir_mask) ? op.rdir[row] : -1;
        for (int q = l; q < nv; q += G)
            vals[q] = 0.0;
        if (rd >= 0) {
            // directory kept by the symbolic phase: load its first level, stage only B's ranges
            for (int t = l; t < W; t += G) {
                dm[t] = op.dir_mask[rd + t];
                doff[t] = op.dir_off[rd + t];
            }
            stage_row<G, true>(op, a0, nA, nullptr, nullptr, b0, blen, av);
        } else {
            for (int t = l; t < W; t += G)
                dm[t] = 0;
            stage_row<G, true>(op, a0, nA, t0, tlen, b0, blen, av);
        }
        __syncthreads();
        if (rd < 0) {
            if (sparse)
                or_tiles<G, OR_COARSE>(op, t0, tlen, nA, dm, nullptr, nullptr, mn, tshift);
            else
                or_tiles<G, OR_FLAT>(op, t0, tlen, nA, dm, nullptr, nullptr, mn, tshift);
        }
        __syncthreads();
        // first level: flat slot offsets (rebuilt) or coarse compact indices (rebuilt), and
        // the total (flat: nnz(C_i); two-level: non-empty tiles)
        int ta, tb;
        unsigned tot;
        unsigned pre = scan_popcounts<G>(dm, W, redu, tot, ta, tb);
        if (rd < 0) {
            for (int t = ta; t < tb; ++t) {
                doff[t] = pre;
                pre += (unsigned)__popcll(dm[t]);
            }
        }
        unsigned nnz = tot;
        const int F = sparse ? (int)tot : S; // fine entries
        if (two_level) {
            __syncthreads();
            if (sparse) {
                if (rd >= 0) {
                    for (int k = l; k < F; k += G) {
                        fm[k] = op.dir_mask[rd + W + k];
                        foff[k] = op.dir_off[rd + W + k];
                    }
                } else {
                    for (int k = l; k < F; k += G)
                        fm[k] = 0;
                }
                // tile of every fine entry, from the coarse bitmap
                for (int w = l; w < W; w += G) {
                    u64 m = dm[w];
                    unsigned k = doff[w];
                    while (m) {
                        ft[k++] = (unsigned)((w << 6) + __ffsll((unsigned long long)m) - 1);
                        m &= m - 1;
                    }
                }
            }
            __syncthreads();
            if (sparse && rd < 0)
                or_tiles<G, OR_FINE>(op, t0, tlen, nA, dm, doff, fm, mn, tshift);
            __syncthreads();
            unsigned tot2;
            unsigned pre2 = scan_popcounts<G>(fm, sparse ? F : 0, redu, tot2, ta, tb);
            if (sparse) {
                nnz = tot2;
                if (rd < 0) {
                    for (int k = ta; k < tb; ++k) {
                        foff[k] = pre2;
                        pre2 += (unsigned)__popcll(fm[k]);
                    }
                }
            }
        }
        const bool ok = (int)nnz == n;
        if (row >= 0 && !ok && l == 0)
            atomicOr(err, ERR_INTERNAL);
        __syncthreads();
        if (row >= 0 && ok) {
            if (dense)
                add_products<G, SLOT_DENSE, P>(op, b0, blen, av, nA, nullptr, nullptr, fm, foff, vals, mn, bshift);
            else if (sparse)
                add_products<G, SLOT_TWO_LEVEL, P>(op, b0, blen, av, nA, dm, doff, fm, foff, vals, mn, bshift);
            else
                add_products<G, SLOT_FLAT, P>(op, b0, blen, av, nA, nullptr, nullptr, fm, foff, vals, mn, bshift);
        }
        __syncthreads();
        if (row >= 0 && ok) {
            if constexpr (G >= 64) {
                constexpr int NW = G / 64;
                const int w = l >> 6, lane = l & 63;
                for (int k = w; k < F; k += NW) {
                    const u64 m = fm[k];
                    if ((m >> lane) & 1ull) {
                        const int t = sparse ? (int)ft[k] : k;
                        const unsigned pos = foff[k] + (unsigned)__popcll(m & below(lane));
                        Ccol[cb + pos] = ((mn + t) << 6) + lane;
                        Cval[cb + pos] = P::result(vals, dense ? (t << 6) + lane : pos);
                    }
                }
            } else {
                for (int k = l; k < F; k += G) {
                    u64 m = fm[k];
                    unsigned pos = foff[k];
                    const int t = sparse ? (int)ft[k] : k;
                    const int c0 = (mn + t) << 6;
                    while (m) {
                        const int b = __ffsll((unsigned long long)m) - 1;
                        Ccol[cb + pos] = c0 + b;
                        if (dense)
                            Cval[cb + pos] = P::result(vals, (t << 6) + b);
                        ++pos;
                        m &= m - 1;
                    }
                }
                if (!dense)
                    for (int q = l; q < n; q += G)
                        Cval[cb + q] = P::result(vals, q);
            }
        }
        __syncthreads();
    }
}

// Rows that do not fit LDS: directory (masks + offsets) in a per-workgroup global
// workspace, accumulation with global atomics straight into C's value array (FP32C: the
// compensation terms go to comp[], folded into C at the end of the row).
template <typename P>
__global__ __launch_bounds__(1024) void k_numeric_global(const int *__restrict__ rows, const int *__restrict__ count_p,
                                                         Operands op, const long long *__restrict__ crp,
                                                         int *__restrict__ Ccol, typename P::V *Cval, float *comp,
                                                         int bshift, int tshift, u64 *gws, long long gws_stride,
                                                         long long gws_tiles, int *err)
{
    typedef typename P::V V;
    const V *__restrict__ Aval = (const V *)op.Aval;
    const V *__restrict__ Bval = (const V *)op.Bval;
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
        for (int q = l; q < n; q += 1024) {
            Cval[cb + q] = 0;
            if (comp)
                comp[cb + q] = 0.0f;
        }
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
                const double a = (double)Aval[e];
                const int b1 = op.Brp[j + 1];
                for (int k = op.Brp[j] + sl; k < b1; k += SG) {
                    const int c = op.Bcol[k];
                    const int t = (c >> 6) - mn;
                    const unsigned slot = doff[t] + (unsigned)__popcll(dm[t] & below(c & 63));
                    P::gadd(Cval, comp, cb + slot, a, Bval[k]);
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
            if (comp) {
                block_sync_global();
                for (int q = l; q < n; q += 1024)
                    Cval[cb + q] = (V)(Cval[cb + q] + comp[cb + q]);
            }
        }
        block_sync_global();
    }
}

// ---- launchers -------------------------------------------------------------------------------

// Grid of an LDS row kernel: one workgroup per row group, or, with a work queue, as many
// workgroups as can be resident at once (occupancy for this LDS size x CUs).
template <typename Kern>
static int64_t row_grid(Kern kernel, int bs, size_t lds, int64_t groups, int num_cus, const unsigned *queue)
{
    int64_t blocks = groups;
    if (queue) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, bs, lds) != hipSuccess || per_cu < 1) {
            (void)hipGetLastError();
            per_cu = 1;
        }
        const int64_t resident = (int64_t)per_cu * (num_cus > 0 ? num_cus : 1);
        blocks = blocks < resident ? blocks : resident;
    }
    if (blocks > (1 << 20))
        blocks = 1 << 20;
    return blocks < 1 ? 1 : blocks;
}

template <int G>
static void symbolic_g(const int *rows, const int *count_dev, int count, int max_need, const Operands &op, int *rnnz,
                       int tshift, int two_level, int num_cus, unsigned *queue, hipStream_t s)
{
    constexpr int BS = G < BLOCK ? BLOCK : G;
    constexpr int RPB = BS / G;
    const int slice = (max_need + 15) & ~15;
    const size_t lds = (size_t)slice * RPB;
    const int64_t blocks = row_grid(k_symbolic_lds<G>, BS, lds, ((int64_t)count + RPB - 1) / RPB, num_cus, queue);
    const int lg = __builtin_ctz(G);
    hipLaunchKernelGGL((k_symbolic_lds<G>), dim3((unsigned)blocks), dim3(BS), lds, s, rows, count_dev, slice, op, rnnz,
                       tshift > lg ? lg : tshift, two_level, queue);
}

hipError_t launch_symbolic(int bin, const int *rows, const int *count_dev, int count, int max_need, const Operands &op,
                           int *rnnz, int tshift, int two_level, u64 *gws, long long gws_stride, int num_cus,
                           int gws_blocks, unsigned *queue, hipStream_t s)
{
    if (count <= 0)
        return hipSuccess;
    switch (bin) {
    case 0: symbolic_g<8>(rows, count_dev, count, max_need, op, rnnz, tshift, two_level, num_cus, queue, s); break;
    case 1: symbolic_g<16>(rows, count_dev, count, max_need, op, rnnz, tshift, two_level, num_cus, queue, s); break;
    case 2: symbolic_g<32>(rows, count_dev, count, max_need, op, rnnz, tshift, two_level, num_cus, queue, s); break;
    case 3: symbolic_g<64>(rows, count_dev, count, max_need, op, rnnz, tshift, two_level, num_cus, queue, s); break;
    case 4: symbolic_g<128>(rows, count_dev, count, max_need, op, rnnz, tshift, two_level, num_cus, queue, s); break;
    case 5: symbolic_g<256>(rows, count_dev, count, max_need, op, rnnz, tshift, two_level, num_cus, queue, s); break;
    case 6: symbolic_g<512>(rows, count_dev, count, max_need, op, rnnz, tshift, two_level, num_cus, queue, s); break;
    case 7: symbolic_g<1024>(rows, count_dev, count, max_need, op, rnnz, tshift, two_level, num_cus, queue, s); break;
    case BIN_GLOBAL: {
        const int blocks = count < gws_blocks ? count : gws_blocks;
        hipLaunchKernelGGL(k_symbolic_global, dim3(blocks), dim3(1024), 0, s, rows, count_dev, op, rnnz,
                           tshift > 10 ? 10 : tshift, gws, gws_stride);
        break;
    }
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int G, typename P>
static void numeric_g(const int *rows, const int *count_dev, int count, int max_need, const Operands &op,
                      const long long *crp, int *Ccol, void *Cval, int bshift, int tshift, int two_level, int *err,
                      int num_cus, unsigned *queue, hipStream_t s)
{
    constexpr int BS = G < BLOCK ? BLOCK : G;
    constexpr int RPB = BS / G;
    const int slice = (max_need + 15) & ~15;
    const size_t lds = (size_t)slice * RPB;
    const int64_t blocks = row_grid(k_numeric_lds<G, P>, BS, lds, ((int64_t)count + RPB - 1) / RPB, num_cus, queue);
    const int lg = __builtin_ctz(G);
    hipLaunchKernelGGL((k_numeric_lds<G, P>), dim3((unsigned)blocks), dim3(BS), lds, s, rows, count_dev, slice, op, crp,
                       Ccol, (typename P::V *)Cval, bshift > lg ? lg : bshift, tshift > lg ? lg : tshift, two_level, err,
                       queue);
}

template <typename P>
static hipError_t numeric_p(int bin, const int *rows, const int *count_dev, int count, int max_need, const Operands &op,
                            const long long *crp, int *Ccol, void *Cval, float *comp, int bshift, int tshift,
                            int two_level, u64 *gws, long long gws_stride, long long gws_tiles, int gws_blocks, int *err,
                            int num_cus, unsigned *queue, hipStream_t s)
{
    switch (bin) {
    case 0: numeric_g<8, P>(rows, count_dev, count, max_need, op, crp, Ccol, Cval, bshift, tshift, two_level, err, num_cus, queue, s); break;
    case 1: numeric_g<16, P>(rows, count_dev, count, max_need, op, crp, Ccol, Cval, bshift, tshift, two_level, err, num_cus, queue, s); break;
    case 2: numeric_g<32, P>(rows, count_dev, count, max_need, op, crp, Ccol, Cval, bshift, tshift, two_level, err, num_cus, queue, s); break;
    case 3: numeric_g<64, P>(rows, count_dev, count, max_need, op, crp, Ccol, Cval, bshift, tshift, two_level, err, num_cus, queue, s); break;
    case 4: numeric_g<128, P>(rows, count_dev, count, max_need, op, crp, Ccol, Cval, bshift, tshift, two_level, err, num_cus, queue, s); break;
    case 5: numeric_g<256, P>(rows, count_dev, count, max_need, op, crp, Ccol, Cval, bshift, tshift, two_level, err, num_cus, queue, s); break;
    case 6: numeric_g<512, P>(rows, count_dev, count, max_need, op, crp, Ccol, Cval, bshift, tshift, two_level, err, num_cus, queue, s); break;
    case 7: numeric_g<1024, P>(rows, count_dev, count, max_need, op, crp, Ccol, Cval, bshift, tshift, two_level, err, num_cus, queue, s); break;
    case BIN_GLOBAL: {
        const int blocks = count < gws_blocks ? count : gws_blocks;
        hipLaunchKernelGGL(k_numeric_global<P>, dim3(blocks), dim3(1024), 0, s, rows, count_dev, op, crp, Ccol,
                           (typename P::V *)Cval, comp, bshift > 10 ? 10 : bshift, tshift > 10 ? 10 : tshift, gws,
                           gws_stride, gws_tiles, err);
        break;
    }
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_numeric(int bin, int fp32c, const int *rows, const int *count_dev, int count, int max_need,
                          const Operands &op, const long long *crp, int *Ccol, void *Cval, float *comp, int bshift,
                          int tshift, int two_level, u64 *gws, long long gws_stride, long long gws_tiles,
                          int gws_blocks, int *err, int num_cus, unsigned *queue, hipStream_t s)
{
    if (count <= 0)
        return hipSuccess;
    if (fp32c)
        return numeric_p<PrecF32C>(bin, rows, count_dev, count, max_need, op, crp, Ccol, Cval, comp, bshift, tshift,
                                   two_level, gws, gws_stride, gws_tiles, gws_blocks, err, num_cus, queue, s);
    return numeric_p<PrecF64>(bin, rows, count_dev, count, max_need, op, crp, Ccol, Cval, nullptr, bshift, tshift,
                              two_level, gws, gws_stride, gws_tiles, gws_blocks, err, num_cus, queue, s);
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
#define CFGP(P)                                                                                                        \
    CFG((k_numeric_lds<8, P>)) CFG((k_numeric_lds<16, P>)) CFG((k_numeric_lds<32, P>)) CFG((k_numeric_lds<64, P>))     \
    CFG((k_numeric_lds<128, P>)) CFG((k_numeric_lds<256, P>)) CFG((k_numeric_lds<512, P>))                           \
    CFG((k_numeric_lds<1024, P>))
    CFGP(PrecF64) CFGP(PrecF32C)
#undef CFGP
#undef CFG
    return e;
}

} // namespace hsmu
