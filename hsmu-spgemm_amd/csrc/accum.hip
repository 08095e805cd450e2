// accum.hip -- symbolic and numeric row kernels of the HSMU SpGEMM path for gfx950.
//
// Symbolic: nnz(C_i) from the OR of B's 64-column tile masks over row i of A, into the
// row's tile directory in LDS (flat: one mask per tile of the row's span; two-level for
// sparse spans: a coarse bitmap of the non-empty tiles, then their masks), with
// ds_or_b64; a popcount and a group scan give nnz(C_i) and each tile's first slot
// (reference: compressed_tilePtr_symbolic.cuh and compressed_tileOR.cuh, two hash
// passes over tiles there; one pass and no hashing here).
//
// Numeric: the HSMU accumulator -- exactly nnz(C_i) slots in LDS, the slot of column c
// being off[t] + popcount(mask[t] & below(c & 63)) for c's directory entry t, i.e. c's
// rank among C's sorted columns found in O(1) (reference: binary search over the sorted
// column list, src/compressed/compressed_Cval.cuh:5-69 and src/small/numeric.cuh:233-253);
// a dense form (one slot per column of a short, full span) skips the directory.  The
// numeric kernels also write C's sorted columns, replacing the reference's rank-sort pass
// (compressed_Ccol_symbolic.cuh:4-63); for the wide rows of the FP64 path a separate
// streaming kernel (k_wide_cols) writes them from the kept directories.
//
// LDS row kernels stage the row's A entries (B row range, first tile, a_ij) in LDS once;
// sub-groups of lanes then take one entry each and walk B_j, several elements per lane
// with all loads issued first; rows of B longer than LONG_SEG (tiles: LONG_TILES) are
// walked by whole waves.  Rows whose slice does not fit LDS (wide rows) use the windowed
// kernels at the end of this file.

#include <hip/hip_runtime.h>

#include <climits>

#include <type_traits>

#include "device_common.h"
#include "hsmu_internal.h"

namespace hsmu {

// XCDs of the device the row kernels map their row groups onto (configure_kernels sets
// it per device; each device holds its own copy).
static __constant__ int c_nxcd = MAX_XCDS;

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

// A B row with at least LONG_SEG elements (left; LONG_TILES tiles in the tile OR) is
// walked by a whole wave, shorter ones by sub-groups: R-MAT's hub rows of B hold tens of
// thousands of entries and would otherwise leave one sub-group walking while the rest of
// the workgroup waits.  The LDS kernels list such entries only in their LONG variants,
// launched when B has a row that long.

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
template <int MODE>
DEVI void or_put(u64 *dm, const unsigned *coff, u64 *fm, int tt, u64 m)
{
    if constexpr (MODE == OR_FLAT)
        atomicOr(&dm[tt], m);
    else if constexpr (MODE == OR_COARSE)
        atomicOr(&dm[tt >> 6], 1ull << (tt & 63));
    else
        atomicOr(&fm[fine_index(dm, coff, tt)], m);
}

// Tiles [tb, te) of one B row by a group of gw lanes (this lane the gl-th), U tiles per
// lane per step.
template <int MODE, int U = UNROLL>
DEVI void or_walk(const Operands &op, int tb, int te, int gw, int gl, u64 *dm, const unsigned *coff, u64 *fm, int mn)
{
    for (int t = tb + gl; t < te; t += U * gw) {
        int c[U];
        u64 m[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int tu = t + u * gw;
            c[u] = -1;
            m[u] = 0;
            if (tu < te) {
                c[u] = op.tcol[tu];
                if constexpr (MODE != OR_COARSE)
                    m[u] = op.tmask[tu];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (c[u] >= 0)
                or_put<MODE>(dm, coff, fm, c[u] - mn, m[u]);
    }
}

template <int G, int MODE, bool LIST = false>
DEVI void or_tiles(const Operands &op, const int *t0, const int *tlen, int nA, const int *lidx, int nlong, u64 *dm,
                   const unsigned *coff, u64 *fm, int mn, int tshift)
{
    const int l = threadIdx.x & (G - 1);
    // LIST: entries with many tiles (listed in lidx), one wave each (the whole group if
    // smaller); the sub-group walks below skip them
    // 1024-lane groups hold a row per CU (LDS bounds their occupancy), so each lane keeps
    // twice as many tile loads in flight
    constexpr int U = G >= 1024 ? 2 * UNROLL : UNROLL;
    if constexpr (LIST) {
        constexpr int LW = G < 64 ? G : 64;
        for (int i = l / LW; i < nlong; i += G / LW) {
            const int e = lidx[i];
            or_walk<MODE, U>(op, t0[e], t0[e] + tlen[e], LW, l & (LW - 1), dm, coff, fm, mn);
        }
    }
    // sub-groups of 2^tshift lanes take one entry each
    const int SG = 1 << tshift, sl = l & (SG - 1), NSG = G >> tshift;
    for (int e = l >> tshift; e < nA; e += NSG) {
        if (LIST && tlen[e] >= LONG_TILES)
            continue;
        or_walk<MODE, U>(op, t0[e], t0[e] + tlen[e], SG, sl, dm, coff, fm, mn);
    }
}

// Each lane scans a contiguous chunk [ta, tb) of the n directory entries; groups of 256
// lanes or more visit it from s = ta + l mod chunk, wrapping round to ta: with every lane
// starting at its chunk's first entry, a chunk of 16 entries would put all 64 lanes of a
// wave on one LDS bank pair.  Smaller groups have short chunks and walk them from ta
// (s = ta) in one loop.  pre: the group's exclusive prefix of the popcounts at ta; head:
// the popcounts of [ta, s); tot: the group total.  Every thread of the workgroup must
// call it (G > 64 synchronises the workgroup).
struct ChunkScan {
    int ta, tb, s;
    unsigned pre, head, tot;
};
template <int G>
DEVI ChunkScan scan_popcounts(const u64 *m, int n, unsigned *redu)
{
    ChunkScan c;
    const int l = threadIdx.x & (G - 1);
    const int chunk = (n + G - 1) / G;
    c.ta = l * chunk;
    c.tb = c.ta + chunk < n ? c.ta + chunk : n;
    unsigned tail = 0, head = 0;
    if constexpr (G >= 256) {
        c.s = chunk > 1 ? c.ta + l % chunk : c.ta;
        if (c.s > c.tb)
            c.s = c.tb;
        for (int t = c.s; t < c.tb; ++t)
            tail += (unsigned)__popcll(m[t]);
        for (int t = c.ta; t < c.s; ++t)
            head += (unsigned)__popcll(m[t]);
    } else {
        c.s = c.ta;
        for (int t = c.ta; t < c.tb; ++t)
            tail += (unsigned)__popcll(m[t]);
    }
    c.head = head;
    c.pre = group_excl_scan<G>(tail + head, redu, c.tot);
    return c;
}

// The exclusive prefix at every entry of the lane's chunk, out(t, prefix, m[t]), in the
// scan's visiting order.  Each entry is read before out is called for it, so out may
// overwrite m[t] itself.
template <int G, typename Out>
DEVI void chunk_offsets(const u64 *m, const ChunkScan &c, Out out)
{
    if constexpr (G >= 256) {
        unsigned p = c.pre + c.head;
        for (int t = c.s; t < c.tb; ++t) {
            const u64 w = m[t];
            out(t, p, w);
            p += (unsigned)__popcll(w);
        }
        p = c.pre;
        for (int t = c.ta; t < c.s; ++t) {
            const u64 w = m[t];
            out(t, p, w);
            p += (unsigned)__popcll(w);
        }
    } else {
        unsigned p = c.pre;
        for (int t = c.ta; t < c.tb; ++t) {
            const u64 w = m[t];
            out(t, p, w);
            p += (unsigned)__popcll(w);
        }
    }
}

// ---- numeric forms ---------------------------------------------------------------------
//
// A slot is 8 bytes in both forms, so the LDS layout and the bins are the same.
// PrecF64 (plain): FP64 values, slot = one double, ds_add_f64 per product.
// PrecF32C (compensated FP32): FP32 values; the product is split exactly into
// p = fl(a*b) and e = fma(a, b, -p) (TwoProduct); the slot holds a float pair (s, c),
// updated by one 64-bit compare-and-swap: s becomes the VALU's round-to-nearest-even
// fl(s + p), and the exact rounding error of that addition (TwoSum, Knuth) plus e goes
// into c.  The result is fl(s + c): c carries every rounding error of the products and
// of the additions into s, so the sum is accurate to about one FP32 rounding of the
// result plus O(n u^2) of sum |a*b| (u = 2^-24).  (Memory-side float atomic adds return
// old values that are not a serial history of the stored sums when lanes collide, so
// they cannot feed TwoSum.)
struct PrecF64 {
    typedef double V;
    typedef double T; // a product ready to be added
    static DEVI T prep(double a, V b) { return a * b; }
    static DEVI void put(double *vals, unsigned slot, T p) { atomicAdd(&vals[slot], p); }
    static DEVI void add(double *vals, unsigned slot, double a, V b) { put(vals, slot, prep(a, b)); }
    static DEVI V result(const double *vals, unsigned slot) { return vals[slot]; }
};

struct PrecF32C {
    typedef float V;
    // (s, c) += p with one 64-bit compare-and-swap on the pair: the VALU forms
    // t = fl(s + p) (round to nearest even), recovers that addition's exact error
    // (TwoSum, Knuth) and adds it and e into c; the swap succeeds only if no other lane
    // changed the pair meanwhile, so the error always describes the addition stored.
    // c's own rounding is second order (|c| ~ u |s|).
    struct T { // the product split exactly: p + e = a * b
        float p, e;
    };
    static DEVI T prep(double a, V b)
    {
        const float af = (float)a;
        const float p = __fmul_rn(af, b);
        return T{p, __builtin_fmaf(af, b, -p)};
    }
    static DEVI void add(double *vals, unsigned slot, double a, V b) { put(vals, slot, prep(a, b)); }
    static DEVI void put(double *vals, unsigned slot, T pe)
    {
        // no contraction: t must be fl(s + p), not fma(a, b, s), for the error to be exact
#pragma clang fp contract(off)
        const float p = pe.p, e = pe.e;
        unsigned long long *w = (unsigned long long *)vals + slot;
        unsigned long long old = *w;
        while (true) {
            const float sv = __uint_as_float((unsigned)old);
            const float cv = __uint_as_float((unsigned)(old >> 32));
            const float t = sv + p;
            const float bb = t - sv;
            const float err = (sv - (t - bb)) + (p - bb);
            const float c2 = cv + (err + e);
            const unsigned long long nw =
                (unsigned long long)__float_as_uint(t) | ((unsigned long long)__float_as_uint(c2) << 32);
            const unsigned long long prev = atomicCAS(w, old, nw);
            if (prev == old)
                break;
            old = prev;
        }
    }
    static DEVI V result(const double *vals, unsigned slot)
    {
        const float *f = (const float *)vals + 2 * slot;
        return f[0] + f[1];
    }
};

// Slot of column c in the row's accumulator, by directory form.
enum { SLOT_FLAT, SLOT_DENSE, SLOT_TWO_LEVEL };

// Add every product a_ij * b_jc of the row into its slot: sub-groups of 2^bshift lanes
// take one entry each and walk B_j, UNROLL elements per lane per iteration with all
// loads issued before the ds_add_f64s.  fm/foff: the (fine) tile masks and first slots;
// cm/coff: the coarse level of a two-level directory.
// B_j's elements [kb, ke) by a group of gw lanes (this lane the gl-th), each product added
// into its slot.
template <int FORM, typename P>
DEVI void add_walk(const Operands &op, int kb, int ke, int gw, int gl, double a, const u64 *cm, const unsigned *coff,
                   const u64 *fm, const unsigned *foff, double *vals, int mn)
{
    typedef typename P::V V;
    const V *__restrict__ Bval = (const V *)op.Bval;
    for (int k = kb + gl; k < ke; k += UNROLL * gw) {
        int c[UNROLL];
        V v[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            const int ku = k + u * gw;
            c[u] = -1;
            v[u] = 0;
            if (ku < ke) {
                c[u] = op.Bcol[ku];
                v[u] = Bval[ku];
            }
        }
        auto slot_of = [&](int col) -> unsigned {
            if constexpr (FORM == SLOT_DENSE) {
                return (unsigned)(col - (mn << 6));
            } else {
                const int t = (col >> 6) - mn;
                const unsigned kk = FORM == SLOT_TWO_LEVEL ? fine_index(cm, coff, t) : (unsigned)t;
                return foff[kk] + (unsigned)__popcll(fm[kk] & below(col & 63));
            }
        };
        if constexpr (std::is_same<P, PrecF64>::value) {
            // every slot and product of the step first, then the adds: a product computed
            // between two adds reuses the registers the add before it still reads, so each
            // add waited for the previous one to leave the LDS queue (measured on banded:
            // 2.5% off the step, 3% off the dominant kernel)
            unsigned slot[UNROLL];
            double pv[UNROLL];
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) {
                slot[u] = 0;
                pv[u] = 0.0;
                if (c[u] >= 0) {
                    slot[u] = slot_of(c[u]);
                    pv[u] = P::prep(a, v[u]);
                }
            }
#pragma unroll
            for (int u = 0; u < UNROLL; ++u)
                if (c[u] >= 0)
                    P::put(vals, slot[u], pv[u]);
        } else {
            // compensated FP32: each add is a compare-and-swap loop that waits for its own
            // result anyway; separating the steps measured 1.5% slower
#pragma unroll
            for (int u = 0; u < UNROLL; ++u)
                if (c[u] >= 0)
                    P::add(vals, slot_of(c[u]), a, v[u]);
        }
    }
}

template <int G, int FORM, typename P, bool LIST>
DEVI void add_products(const Operands &op, const int *b0, const int *blen, const double *av, int nA, const int *lidx,
                       int nlong, const u64 *cm, const unsigned *coff, const u64 *fm, const unsigned *foff, double *vals,
                       int mn, int bshift)
{
    const int l = threadIdx.x & (G - 1);
    // LIST: long B rows (listed in lidx), one wave each (the whole group if smaller)
    if constexpr (LIST) {
        constexpr int LW = G < 64 ? G : 64;
        for (int i = l / LW; i < nlong; i += G / LW) {
            const int e = lidx[i];
            add_walk<FORM, P>(op, b0[e], b0[e] + blen[e], LW, l & (LW - 1), av[e], cm, coff, fm, foff, vals, mn);
        }
    }
    // sub-groups of SG lanes take one entry each
    const int SG = 1 << bshift, sl = l & (SG - 1), NSG = G >> bshift;
    for (int e = l >> bshift; e < nA; e += NSG) {
        if (LIST && blen[e] >= LONG_SEG)
            continue;
        add_walk<FORM, P>(op, b0[e], b0[e] + blen[e], SG, sl, av[e], cm, coff, fm, foff, vals, mn);
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
template <int G, bool LONG>
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
    for (GroupWalk gw(queue, (nrows + RPB - 1) / RPB, grp_s, c_nxcd); gw.more(); gw.next()) {
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
        if constexpr (G >= 64) { // one row per workgroup: its scalars in scalar registers
            S = uniform(S);
            mn = uniform(mn);
            a0 = uniform(a0);
            nA = uniform(nA);
            tp = uniform(tp);
        }
        const bool sparse = use_sparse(S, tp);
        const int W = sparse ? (S + 63) >> 6 : S; // entries of the first level
        const RowLayout L = row_layout(0, S, 0, nA, tp, LONG);
        u64 *dm = (u64 *)(slice + L.dm);
        unsigned *doff = (unsigned *)(slice + L.doff);
        int *tlen = (int *)(slice + L.tlen);
        int *t0 = (int *)(slice + L.t0);
        int *lidx = (int *)(slice + L.lidx);
        int *lcnt = (int *)(slice + L.lcnt);
        for (int t = l; t < W; t += G)
            dm[t] = 0;
        if (LONG && G >= 128 && l == 0)
            *lcnt = 0;
        stage_row<G, false>(op, a0, nA, t0, tlen, nullptr, nullptr, nullptr);
        __syncthreads();
        // rows of the wide bins: entries with many tiles, one wave each
        int nlong = 0;
        if constexpr (LONG && G >= 128) {
            for (int k = l; k < nA; k += G)
                if (tlen[k] >= LONG_TILES)
                    lidx[atomicAdd(lcnt, 1)] = k;
            __syncthreads();
            nlong = *lcnt;
        }
        // the list, when there is one, takes the long entries out of the sub-group walks
        constexpr bool LIST = LONG && G >= 128;
        if (sparse)
            or_tiles<G, OR_COARSE, LIST>(op, t0, tlen, nA, lidx, nlong, dm, nullptr, nullptr, mn, tshift);
        else
            or_tiles<G, OR_FLAT, LIST>(op, t0, tlen, nA, lidx, nlong, dm, nullptr, nullptr, mn, tshift);
        __syncthreads();
        const ChunkScan cs = scan_popcounts<G>(dm, W, redu);
        const unsigned tot = cs.tot;
        const long long rd = (row >= 0 && op.dir_mask) ? op.rdir[row] : -1;
        // Kept directories of groups of 256 lanes or more (chunks of many entries) go out
        // coalesced: the masks first, then each entry's slot offset (flat rows: written
        // over its mask in LDS; coarse level: into doff, which the fine pass reads) after a
        // barrier.  keep is uniform over the launch, so every thread takes the barriers.
        // Smaller groups store each entry from the chunk walk.
        constexpr bool COAL = G >= 256;
        const bool keep = COAL && op.dir_mask != nullptr;
        unsigned nnz = tot;
        if (COAL && !sparse && rd >= 0)
            for (int t = l; t < W; t += G)
                op.dir_mask[rd + t] = dm[t];
        if (keep)
            __syncthreads();
        if (!COAL) {
            if (sparse)
                chunk_offsets<G>(dm, cs, [&](int t, unsigned p, u64 w) {
                    doff[t] = p;
                    if (rd >= 0) {
                        op.dir_mask[rd + t] = w;
                        op.dir_off[rd + t] = p;
                    }
                });
            else if (rd >= 0)
                chunk_offsets<G>(dm, cs, [&](int t, unsigned p, u64 w) {
                    op.dir_mask[rd + t] = w;
                    op.dir_off[rd + t] = p;
                });
        } else if (sparse) {
            chunk_offsets<G>(dm, cs, [&](int t, unsigned p, u64) { doff[t] = p; });
        } else if (rd >= 0) {
            chunk_offsets<G>(dm, cs, [&](int t, unsigned p, u64) { dm[t] = p; });
        }
        if (keep)
            __syncthreads();
        if (COAL && rd >= 0) {
            if (sparse) {
                for (int t = l; t < W; t += G) {
                    op.dir_mask[rd + t] = dm[t];
                    op.dir_off[rd + t] = doff[t];
                }
            } else {
                for (int t = l; t < W; t += G)
                    op.dir_off[rd + t] = (unsigned)dm[t];
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
                or_tiles<G, OR_FINE, LIST>(op, t0, tlen, nA, lidx, nlong, dm, doff, fm, mn, tshift);
            __syncthreads();
            const int F = sparse ? (int)tot : 0;
            const ChunkScan fs = scan_popcounts<G>(fm, F, redu);
            if (sparse)
                nnz = fs.tot;
            // the fine level goes out as the flat directory does
            if (COAL && rd >= 0)
                for (int k = l; k < F; k += G)
                    op.dir_mask[rd + W + k] = fm[k];
            if (keep)
                __syncthreads();
            if (rd >= 0) {
                if (COAL)
                    chunk_offsets<G>(fm, fs, [&](int k, unsigned p, u64) { fm[k] = p; });
                else
                    chunk_offsets<G>(fm, fs, [&](int k, unsigned p, u64 w) {
                        op.dir_mask[rd + W + k] = w;
                        op.dir_off[rd + W + k] = p;
                    });
            }
            if (keep)
                __syncthreads();
            if (COAL && rd >= 0)
                for (int k = l; k < F; k += G)
                    op.dir_off[rd + W + k] = (unsigned)fm[k];
        }
        if (row >= 0 && l == 0)
            rnnz[row] = (int)nnz;
        __syncthreads();
    }
}

// Rows of the global bin whose flat directory does not fit LDS either (narrower ones run
// in k_symbolic_wide): the directory over the row's span in global memory, updated with
// L2 atomics -- in the kept directory (rdir) when the directories are kept, so the
// numeric phase reads masks and slot offsets from there, else in a per-workgroup
// workspace.  One 1024-lane workgroup per row.
__global__ __launch_bounds__(1024) void k_symbolic_global(const int *__restrict__ rows, const int *__restrict__ count_p,
                                                          Operands op, int *__restrict__ rnnz, int tshift, u64 *gws,
                                                          long long gws_stride)
{
    __shared__ unsigned red[16];
    const int l = threadIdx.x;
    const int nrows = *count_p;
    const int SG = 1 << tshift, sl = l & (SG - 1), sg = l >> tshift, NSG = 1024 >> tshift;
    for (int idx = blockIdx.x; idx < nrows; idx += gridDim.x) {
        const int row = rows[idx];
        const int mn = op.rmin[row];
        const int S = op.rmax[row] - mn + 1;
        const long long rd = op.dir_mask ? op.rdir[row] : -1;
        u64 *dm = rd >= 0 ? op.dir_mask + rd : gws + (size_t)blockIdx.x * gws_stride;
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
        const ChunkScan cs = scan_popcounts<1024>(dm, S, red);
        if (rd >= 0)
            chunk_offsets<1024>(dm, cs, [&](int t, unsigned p, u64) { op.dir_off[rd + t] = p; });
        if (l == 0)
            rnnz[row] = (int)cs.tot;
        block_sync_global();
    }
}

// Rows of the global bin whose flat directory alone fits LDS (their staged A entries did
// not): the directory in LDS, B_j's tiles read from global memory per entry, as in
// k_symbolic_global, but ORed with ds_or_b64 instead of L2 atomics.  A kept directory
// goes out coalesced (masks, then the offsets written over them in LDS).  One 1024-lane
// workgroup per row; dynamic LDS: 8 bytes per tile of the widest span.
__global__ __launch_bounds__(1024) void k_symbolic_wide(const int *__restrict__ rows, const int *__restrict__ count_p,
                                                        Operands op, int *__restrict__ rnnz, int tshift)
{
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ unsigned red[16];
    u64 *dm = (u64 *)lds;
    const int l = threadIdx.x;
    const int nrows = *count_p;
    const int SG = 1 << tshift, sl = l & (SG - 1), sg = l >> tshift, NSG = 1024 >> tshift;
    for (int idx = blockIdx.x; idx < nrows; idx += gridDim.x) {
        const int row = uniform(rows[idx]);
        const int mn = uniform(op.rmin[row]);
        const int S = uniform(op.rmax[row]) - mn + 1;
        const long long rd = op.dir_mask ? uniform64(op.rdir[row]) : -1;
        for (int t = l; t < S; t += 1024)
            dm[t] = 0;
        __syncthreads();
        const int a1 = op.Arp[row + 1];
        for (int e = op.Arp[row] + sg; e < a1; e += NSG) {
            const int j = op.Acol[e];
            or_walk<OR_FLAT, 2 * UNROLL>(op, op.tptr[j], op.tptr[j + 1], SG, sl, dm, nullptr, nullptr, mn);
        }
        __syncthreads();
        const ChunkScan cs = scan_popcounts<1024>(dm, S, red);
        if (rd >= 0)
            for (int t = l; t < S; t += 1024)
                op.dir_mask[rd + t] = dm[t];
        __syncthreads();
        if (rd >= 0)
            chunk_offsets<1024>(dm, cs, [&](int t, unsigned p, u64) { dm[t] = p; });
        __syncthreads();
        if (rd >= 0)
            for (int t = l; t < S; t += 1024)
                op.dir_off[rd + t] = (unsigned)dm[t];
        if (l == 0)
            rnnz[row] = (int)cs.tot;
        __syncthreads();
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
template <int G, typename P, bool LONG>
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
    for (GroupWalk gw(queue, (nrows + RPB - 1) / RPB, grp_s, c_nxcd); gw.more(); gw.next()) {
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
        if constexpr (G >= 64) { // one row per workgroup: its scalars in scalar registers
            S = uniform(S);
            mn = uniform(mn);
            n = uniform(n);
            a0 = uniform(a0);
            nA = uniform(nA);
            tp = uniform(tp);
            cb = uniform64(cb);
        }
        const bool sparse = use_sparse(S, tp);
        const int W = sparse ? (S + 63) >> 6 : S; // entries of the first level
        const RowLayout L = row_layout(1, S, n, nA, tp, LONG);
        u64 *dm = (u64 *)(slice + L.dm);
        unsigned *doff = (unsigned *)(slice + L.doff);
        unsigned *ft = (unsigned *)(slice + L.ft);
        double *vals = (double *)(slice + L.vals);
        double *av = (double *)(slice + L.av);
        int *tlen = (int *)(slice + L.tlen);
        int *t0 = (int *)(slice + L.t0);
        int *blen = (int *)(slice + L.blen);
        int *b0 = (int *)(slice + L.b0);
        int *lidx = (int *)(slice + L.lidx);
        int *lcnt = (int *)(slice + L.lcnt);
        // fine level: the flat directory itself, or the compact entries after the coarse words
        u64 *fm = sparse ? dm + W : dm;
        unsigned *foff = sparse ? doff + W : doff;
        const bool dense = !sparse && use_dense(S, n);
        const int nv = dense ? 64 * S : n;
        const long long rd = (row >= 0 && op.dir_mask) ? op.rdir[row] : -1;
        for (int q = l; q < nv; q += G)
            vals[q] = 0.0;
        if (LONG && l == 0)
            *lcnt = 0;
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
        // entries with long rows of B, walked by whole waves
        if constexpr (LONG) {
            for (int k = l; k < nA; k += G)
                if (blen[k] >= LONG_SEG)
                    lidx[atomicAdd(lcnt, 1)] = k;
        }
        // One row per workgroup with its directory kept (a uniform condition): the slot
        // offsets and totals are loaded, so neither the tile OR with its barrier nor the
        // popcount scans run.
        const bool known = G >= 64 && rd >= 0;
        if (rd < 0) {
            if (sparse)
                or_tiles<G, OR_COARSE>(op, t0, tlen, nA, nullptr, 0, dm, nullptr, nullptr, mn, tshift);
            else
                or_tiles<G, OR_FLAT>(op, t0, tlen, nA, nullptr, 0, dm, nullptr, nullptr, mn, tshift);
        }
        // first level: flat slot offsets (rebuilt) or coarse compact indices (rebuilt), and
        // the total (flat: nnz(C_i); two-level: non-empty tiles)
        unsigned tot;
        if (known) {
            tot = sparse ? (W > 0 ? doff[W - 1] + (unsigned)__popcll(dm[W - 1]) : 0u) : (unsigned)n;
        } else {
            __syncthreads();
            const ChunkScan cs = scan_popcounts<G>(dm, W, redu);
            tot = cs.tot;
            if (rd < 0)
                chunk_offsets<G>(dm, cs, [&](int t, unsigned p, u64) { doff[t] = p; });
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
                or_tiles<G, OR_FINE>(op, t0, tlen, nA, nullptr, 0, dm, doff, fm, mn, tshift);
            if (known) {
                if (sparse)
                    nnz = (unsigned)n;
            } else {
                __syncthreads();
                const ChunkScan fs = scan_popcounts<G>(fm, sparse ? F : 0, redu);
                if (sparse) {
                    nnz = fs.tot;
                    if (rd < 0)
                        chunk_offsets<G>(fm, fs, [&](int k, unsigned p, u64) { foff[k] = p; });
                }
            }
        }
        const bool ok = (int)nnz == n;
        if (row >= 0 && !ok && l == 0)
            atomicOr(err, ERR_INTERNAL);
        __syncthreads();
        if (row >= 0 && ok) {
            const int nlong = LONG ? *lcnt : 0;
            if (dense)
                add_products<G, SLOT_DENSE, P, LONG>(op, b0, blen, av, nA, lidx, nlong, nullptr, nullptr, fm, foff, vals,
                                                     mn, bshift);
            else if (sparse)
                add_products<G, SLOT_TWO_LEVEL, P, LONG>(op, b0, blen, av, nA, lidx, nlong, dm, doff, fm, foff, vals,
                                                         mn, bshift);
            else
                add_products<G, SLOT_FLAT, P, LONG>(op, b0, blen, av, nA, lidx, nlong, nullptr, nullptr, fm, foff, vals,
                                                    mn, bshift);
        }
        __syncthreads();
        if (row >= 0 && ok) {
            if (G >= 64 && (dense || (unsigned)n >= 16u * (unsigned)F)) {
                // tiles holding 16 columns or more on average: a wave per tile, a lane per column
                constexpr int NW = G < 64 ? 1 : G / 64;
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
            } else { // a lane per tile, then the values in slot order
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

// Rows that do not fit LDS (wide rows): a dense accumulator over a window of WIN_COLS
// columns in LDS, slid over the row's column span.  One 1024-lane workgroup per row.  Every
// A entry e keeps, in a global workspace, a cursor into B_j (cur), the end of B_j (kend)
// and the column at the cursor (nxt).  For each window, the entries with nxt inside it are
// collected into an LDS list (ACT_CHUNK at a time), so an entry costs one coalesced load
// in a window it has nothing for; the listed entries' sub-groups walk B_j from the cursor
// while the columns stay in the window, adding each product into its slot (ds_add) and
// marking the column present (ds_or).  Every product is read once and sorted B rows need
// no search.  After a window its present columns are counted and written in column order
// after those of the previous windows.  (Reference: the spECK dense path of the wide-row
// bin, src/external/spECK/dense_acc.cu:3-267, whose early return leaves long rows
// unwritten.)
// Walk B_j from element k while the columns stay below chi, with a group of gw lanes of
// one wave (this lane is gl-th, the group starts at wave lane glo, gmask = gw one-bits):
// U elements per lane per step, all loads of a step issued before any is used.  B_j is
// sorted, so the in-window elements of a step are a prefix of its U * gw elements; add()
// takes each of them.  Returns the new cursor (the same in every lane of the group).
template <int U, typename V, typename F>
DEVI int walk_segment(const int *__restrict__ Bcol, const V *__restrict__ Bval, int k, int ke, int chi, int gw, int gl,
                      int glo, u64 gmask, F &&add)
{
    while (k < ke) {
        int c[U];
        V v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int kk = k + gl + u * gw;
            c[u] = INT_MAX;
            v[u] = 0;
            if (kk < ke) {
                c[u] = Bcol[kk];
                v[u] = Bval[kk];
            }
        }
        int took = 0;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const bool in = c[u] < chi;
            if (in)
                add(c[u], v[u]);
            took += __popcll((__ballot(in) >> glo) & gmask);
        }
        k += took;
        if (took < U * gw)
            break;
    }
    return k;
}

constexpr int WIN_COLS = 8192;
constexpr int WIN_WORDS = WIN_COLS / 64;
constexpr int ACT_CHUNK = 2048;
constexpr int WIN_LDS = WIN_COLS * 8 + WIN_WORDS * 8 + ACT_CHUNK * 4; // slots, presence bits, active list

// List the entries of [b0, min(b0 + ACT_CHUNK, nA)) whose next column is below chi: those
// with at least LONG_SEG elements of B_j left from the back of act (count *nlong), the
// others from the front (count *nact).  Every thread of the workgroup calls it.
DEVI void collect_active(const int *nxt, const int *cur, const int *kend, int b0, int nA, int chi, int *act, int *nact,
                         int *nlong)
{
    const int l = threadIdx.x, lane = l & 63;
    const int lim = b0 + ACT_CHUNK < nA ? b0 + ACT_CHUNK : nA;
    for (int e0 = b0; e0 < lim; e0 += 1024) {
        const int e = e0 + l;
        const bool live = e < lim && nxt[e] < chi;
        const bool lng = live && kend[e] - cur[e] >= LONG_SEG;
        const u64 ms = __ballot(live && !lng), ml = __ballot(lng);
        int s0 = 0, l0 = 0;
        if (lane == 0) {
            if (ms)
                s0 = atomicAdd(nact, __popcll(ms));
            if (ml)
                l0 = atomicAdd(nlong, __popcll(ml));
        }
        s0 = __shfl(s0, 0, 64);
        l0 = __shfl(l0, 0, 64);
        if (live && !lng)
            act[s0 + mbcnt(ms)] = e;
        if (lng)
            act[ACT_CHUNK - 1 - (l0 + mbcnt(ml))] = e;
    }
}

// The directory form the symbolic kernels stored for a row of the wide bin: two-level
// (coarse words, then compact fine entries) or flat (one entry per tile of the span).
DEVI bool wide_two_level(int S, int nA, int tp, long long sym_row_max, int sym_all_global, int long_rows)
{
    return use_sparse(S, tp) && !sym_all_global && sym_need(S, nA, tp, long_rows) <= sym_row_max;
}

template <typename P>
__global__ __launch_bounds__(1024) void k_numeric_window(const int *__restrict__ rows, const int *__restrict__ count_p,
                                                         Operands op, const long long *__restrict__ crp,
                                                         int *__restrict__ Ccol, typename P::V *__restrict__ Cval,
                                                         int bshift, int *cur_ws, long long cur_stride, long long ft_stride,
                                                         int *err)
{
    typedef typename P::V V;
    const V *__restrict__ Aval = (const V *)op.Aval;
    const V *__restrict__ Bval = (const V *)op.Bval;
    extern __shared__ __align__(16) unsigned char lds[];
    double *acc = (double *)lds;
    u64 *bits = (u64 *)(lds + WIN_COLS * 8);
    int *act = (int *)(lds + WIN_COLS * 8 + WIN_WORDS * 8);
    __shared__ unsigned red[16];
    __shared__ unsigned woff[WIN_WORDS];
    __shared__ int nact_s, nlong_s;
    const int l = threadIdx.x, lane = l & 63, w = l >> 6;
    // per-workgroup workspace: cursors, ends of the B rows, columns at the cursors
    int *cur = cur_ws + (size_t)blockIdx.x * (3 * cur_stride + ft_stride);
    int *kend = cur + cur_stride;
    int *nxt = kend + cur_stride;
    const int SG = 1 << bshift, sl = l & (SG - 1), NSG = 1024 >> bshift;
    const int sg0 = lane & ~(SG - 1); // first lane of this lane's sub-group within the wave
    const u64 sgmask = SG >= 64 ? ~0ull : ((1ull << SG) - 1ull);
    const int nrows = *count_p;
    for (int idx = blockIdx.x; idx < nrows; idx += gridDim.x) {
        const int row = rows[idx];
        const int a0 = op.Arp[row], nA = op.Arp[row + 1] - a0;
        const long long cb = crp[row];
        const long long n = crp[row + 1] - cb;
        const int cbeg = op.rmin[row] << 6, cend = (op.rmax[row] + 1) << 6;
        for (int e = l; e < nA; e += 1024) {
            const int j = op.Acol[a0 + e];
            const int k0 = op.Brp[j], k1 = op.Brp[j + 1];
            cur[e] = k0;
            kend[e] = k1;
            nxt[e] = k0 < k1 ? op.Bcol[k0] : INT_MAX;
        }
        block_sync_global();
        long long written = 0;
        for (int c0 = cbeg; c0 < cend; c0 += WIN_COLS) {
            const int c1 = c0 + WIN_COLS;
            for (int q = l; q < WIN_COLS; q += 1024)
                acc[q] = 0.0;
            if (l < WIN_WORDS)
                bits[l] = 0;
            auto add_d = [&](int c, V v, double a) {
                const int q = c - c0;
                P::add(acc, (unsigned)q, a, v);
                atomicOr(&bits[q >> 6], 1ull << (q & 63));
            };
            for (int base = 0; base < nA; base += ACT_CHUNK) {
                if (l == 0) {
                    nact_s = 0;
                    nlong_s = 0;
                }
                __syncthreads();
                collect_active(nxt, cur, kend, base, nA, c1, act, &nact_s, &nlong_s);
                __syncthreads();
                const int nact = nact_s, nlong = nlong_s;
                for (int t = w; t < nlong; t += 16) {
                    const int e = act[ACT_CHUNK - 1 - t];
                    const int ke = kend[e];
                    const double a = (double)Aval[a0 + e];
                    const int k = walk_segment<4>(op.Bcol, Bval, cur[e], ke, c1, 64, lane, 0, ~0ull,
                                                  [&](int c, V v) { add_d(c, v, a); });
                    if (lane == 0) {
                        cur[e] = k;
                        nxt[e] = k < ke ? op.Bcol[k] : INT_MAX;
                    }
                }
                for (int t = l >> bshift; t < nact; t += NSG) {
                    const int e = act[t];
                    const int ke = kend[e];
                    const double a = (double)Aval[a0 + e];
                    const int k = walk_segment<2>(op.Bcol, Bval, cur[e], ke, c1, SG, sl, sg0, sgmask,
                                                  [&](int c, V v) { add_d(c, v, a); });
                    if (sl == 0) {
                        cur[e] = k;
                        nxt[e] = k < ke ? op.Bcol[k] : INT_MAX;
                    }
                }
                __syncthreads(); // act is rewritten by the next chunk
            }
            // column order: word offsets within the window, then one wave per word
            const unsigned cnt = l < WIN_WORDS ? (unsigned)__popcll(bits[l]) : 0u;
            unsigned tot;
            const unsigned pre = group_excl_scan<1024>(cnt, red, tot);
            if (l < WIN_WORDS)
                woff[l] = pre;
            __syncthreads();
            for (int q = w; q < WIN_WORDS; q += 16) {
                const u64 m = bits[q];
                const long long r = written + woff[q] + __popcll(m & below(lane));
                if (((m >> lane) & 1ull) && r < n) { // r >= n only if the counts disagree (flagged below)
                    Ccol[cb + r] = c0 + (q << 6) + lane;
                    Cval[cb + r] = P::result(acc, (unsigned)((q << 6) + lane));
                }
            }
            written += tot;
            __syncthreads();
        }
        if (l == 0 && written != n)
            atomicOr(err, ERR_INTERNAL);
        block_sync_global(); // workspace of the next row
    }
}

// Wide rows whose tile directory was kept (flat, or two-level when the symbolic phase
// held the row in LDS): the HSMU compact accumulator, windowed over the row's OUTPUT
// positions.  A window is a run of directory entries [k0, k1) holding at most CW_VALS of
// C's entries and CW_TILES entries, so its slots and its part of the directory fit LDS;
// a column's slot is found in O(1) from the directory as in the LDS kernels (two-level
// rows read their coarse words, a few KiB per row, from global memory).  Cursors into
// the B rows and active lists as in the dense window kernel; a window spans the row's
// columns from its first to its last non-empty tile, so sparse stretches cost nothing and
// a row needs about max(nnz / CW_VALS, entries / CW_TILES) windows instead of span /
// WIN_COLS.  Values go out coalesced (they are C's entries in order), columns from the
// masks.
constexpr int CW_VALS = 4096;
constexpr int CW_TILES = 2048;
constexpr int CW_LDS = CW_VALS * 8 + CW_TILES * 16 + ACT_CHUNK * 4; // slots, window directory, active list

template <typename P>
__global__ __launch_bounds__(1024) void k_numeric_wcompact(const int *__restrict__ rows, const int *__restrict__ count_p,
                                                           Operands op, const long long *__restrict__ crp,
                                                           int *__restrict__ Ccol, typename P::V *__restrict__ Cval,
                                                           int bshift, int *ws, long long cur_stride, long long ft_stride,
                                                           long long sym_row_max, int sym_all_global, int long_rows,
                                                           int *err)
{
    typedef typename P::V V;
    const V *__restrict__ Aval = (const V *)op.Aval;
    const V *__restrict__ Bval = (const V *)op.Bval;
    extern __shared__ __align__(16) unsigned char lds[];
    double *vals = (double *)lds;
    u64 *wm = (u64 *)(lds + CW_VALS * 8);                        // window masks
    unsigned *wo = (unsigned *)(lds + CW_VALS * 8 + CW_TILES * 8); // slots relative to the window
    unsigned *wt = wo + CW_TILES;                                 // tiles relative to rmin
    int *act = (int *)(wt + CW_TILES);
    __shared__ int nact_s, nlong_s, k1_s;
    const int l = threadIdx.x, lane = l & 63, w = l >> 6;
    int *cur = ws + (size_t)blockIdx.x * (3 * cur_stride + ft_stride);
    int *kend = cur + cur_stride;
    int *nxt = kend + cur_stride;
    int *ft = nxt + cur_stride; // tile of every fine entry of a two-level row
    const int SG = 1 << bshift, sl = l & (SG - 1), NSG = 1024 >> bshift;
    const int sg0 = lane & ~(SG - 1);
    const u64 sgmask = SG >= 64 ? ~0ull : ((1ull << SG) - 1ull);
    const int nrows = *count_p;
    for (int idx = blockIdx.x; idx < nrows; idx += gridDim.x) {
        const int row = rows[idx];
        const int a0 = op.Arp[row], nA = op.Arp[row + 1] - a0;
        const long long cb = crp[row];
        const long long n = crp[row + 1] - cb;
        const int mn = op.rmin[row];
        const int S = op.rmax[row] - mn + 1;
        const int tp = op.rtp[row];
        const long long rd = op.rdir[row];
        // the form the symbolic phase kept: two-level only if the row was in an LDS bin there
        const bool two = wide_two_level(S, nA, tp, sym_row_max, sym_all_global, long_rows);
        const int W = two ? (S + 63) >> 6 : 0;
        const u64 *cmask = op.dir_mask + rd;      // coarse words (two-level)
        const unsigned *coff = op.dir_off + rd;
        const u64 *fm = op.dir_mask + rd + W;     // fine entries (flat: every tile of the span)
        const unsigned *foff = op.dir_off + rd + W;
        const int F = two ? (int)(coff[W - 1] + (unsigned)__popcll(cmask[W - 1])) : S;
        if (two) {
            for (int q = l; q < W; q += 1024) {
                u64 m = cmask[q];
                unsigned k = coff[q];
                while (m) {
                    ft[k++] = (q << 6) + __ffsll((unsigned long long)m) - 1;
                    m &= m - 1;
                }
            }
        }
        for (int e = l; e < nA; e += 1024) {
            const int j = op.Acol[a0 + e];
            const int k0 = op.Brp[j], k1 = op.Brp[j + 1];
            cur[e] = k0;
            kend[e] = k1;
            nxt[e] = k0 < k1 ? op.Bcol[k0] : INT_MAX;
        }
        block_sync_global();
        long long done = 0;
        for (int k0 = 0; k0 < F;) {
            // window end: at most CW_TILES entries and CW_VALS slots (one entry holds <= 64)
            if (l == 0) {
                const unsigned base = foff[k0];
                int lo = k0 + 1, hi = k0 + CW_TILES < F ? k0 + CW_TILES : F;
                while (lo < hi) { // largest k1 in [k0+1, hi] with slots(k0, k1) <= CW_VALS
                    const int mid = (lo + hi + 1) >> 1;
                    const long long end = mid == F ? n : (long long)foff[mid];
                    if (end - base <= CW_VALS)
                        lo = mid;
                    else
                        hi = mid - 1;
                }
                k1_s = lo;
            }
            __syncthreads();
            const int k1 = k1_s;
            const int nk = k1 - k0;
            const long long base = foff[k0];
            const int cnt = (int)((k1 == F ? n : (long long)foff[k1]) - base);
            for (int q = l; q < nk; q += 1024) {
                wm[q] = fm[k0 + q];
                wo[q] = (unsigned)(foff[k0 + q] - base);
                wt[q] = two ? (unsigned)ft[k0 + q] : (unsigned)(k0 + q);
            }
            for (int q = l; q < cnt; q += 1024)
                vals[q] = 0.0;
            __syncthreads();
            const int chi = (mn + (int)wt[nk - 1] + 1) << 6; // end of the window's last tile
            auto add_c = [&](int c, V v, double a) {
                const int tt = (c >> 6) - mn;
                const int ke2 = two ? (int)(coff[tt >> 6] + (unsigned)__popcll(cmask[tt >> 6] & below(tt & 63))) : tt;
                const int r = ke2 - k0;
                P::add(vals, wo[r] + (unsigned)__popcll(wm[r] & below(c & 63)), a, v);
            };
            for (int b0 = 0; b0 < nA; b0 += ACT_CHUNK) {
                if (l == 0) {
                    nact_s = 0;
                    nlong_s = 0;
                }
                __syncthreads();
                collect_active(nxt, cur, kend, b0, nA, chi, act, &nact_s, &nlong_s);
                __syncthreads();
                const int nact = nact_s, nlong = nlong_s;
                // long B rows: one wave each
                for (int t = w; t < nlong; t += 16) {
                    const int e = act[ACT_CHUNK - 1 - t];
                    const int ke = kend[e];
                    const double a = (double)Aval[a0 + e];
                    const int k = walk_segment<4>(op.Bcol, Bval, cur[e], ke, chi, 64, lane, 0, ~0ull,
                                                  [&](int c, V v) { add_c(c, v, a); });
                    if (lane == 0) {
                        cur[e] = k;
                        nxt[e] = k < ke ? op.Bcol[k] : INT_MAX;
                    }
                }
                // the others: sub-groups of SG lanes
                for (int t = l >> bshift; t < nact; t += NSG) {
                    const int e = act[t];
                    const int ke = kend[e];
                    const double a = (double)Aval[a0 + e];
                    const int k = walk_segment<2>(op.Bcol, Bval, cur[e], ke, chi, SG, sl, sg0, sgmask,
                                                  [&](int c, V v) { add_c(c, v, a); });
                    if (sl == 0) {
                        cur[e] = k;
                        nxt[e] = k < ke ? op.Bcol[k] : INT_MAX;
                    }
                }
                __syncthreads();
            }
            // C's entries of this window, in order
            for (int q = l; q < cnt; q += 1024)
                Cval[cb + base + q] = P::result(vals, (unsigned)q);
            for (int q = w; q < nk; q += 16) {
                const u64 m = wm[q];
                if ((m >> lane) & 1ull)
                    Ccol[cb + base + wo[q] + __popcll(m & below(lane))] = ((mn + (int)wt[q]) << 6) + lane;
            }
            done += cnt;
            k0 = k1;
            __syncthreads();
        }
        if (l == 0 && done != n)
            atomicOr(err, ERR_INTERNAL);
        block_sync_global(); // workspace of the next row
    }
}

// Work items of the split wide-row kernel: each row's entries in runs of
// wide_split(nnz(C_i)), listed largest rows first (classes of half an octave of nnz(C_i),
// by a counting sort): the big rows start early (their items bound the tail), and the
// workgroups that run together work on rows of similar size, which move through their
// columns, and so through the shared rows of B, at similar paces.
// ctr: [0] items listed, [1] items taken, [2, 2 + WS_CLASSES) items per class,
// [2 + WS_CLASSES, 2 + 2 WS_CLASSES) first item of each class.
constexpr int WS_CLASSES = 32;
DEVI int wide_class(long long n)
{
    const int lg2 = n > 1 ? 2 * (63 - __builtin_clzll((unsigned long long)n)) +
                                (((unsigned long long)n >> (62 - __builtin_clzll((unsigned long long)n))) & 1)
                          : 0; // 2 log2(n), rounded down to half octaves
    const int c = lg2 - 20; // below 2^10: class 0
    return c < 0 ? 0 : (c >= WS_CLASSES ? WS_CLASSES - 1 : c);
}

DEVI void wide_row_items(const int *rows, int idx, const int *Arp, const long long *crp, int vals, int &row, int &nA,
                         int &sz, int &ns, int &cls)
{
    row = rows[idx];
    nA = Arp[row + 1] - Arp[row];
    const long long n = crp[row + 1] - crp[row];
    sz = wide_split(n, vals);
    ns = nA > 0 ? (nA + sz - 1) / sz : 1;
    cls = wide_class(n);
}

// pass 1: items per class (per-workgroup counts in LDS, one global atomic per class)
__global__ __launch_bounds__(256) void k_wide_count(const int *__restrict__ rows, const int *__restrict__ count_p,
                                                    const int *__restrict__ Arp, const long long *__restrict__ crp,
                                                    int vals, unsigned *ctr)
{
    __shared__ unsigned lc[WS_CLASSES];
    if (threadIdx.x < WS_CLASSES)
        lc[threadIdx.x] = 0;
    __syncthreads();
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx < *count_p) {
        int row, nA, sz, ns, cls;
        wide_row_items(rows, idx, Arp, crp, vals, row, nA, sz, ns, cls);
        atomicAdd(&lc[cls], (unsigned)ns);
    }
    __syncthreads();
    if (threadIdx.x < WS_CLASSES && lc[threadIdx.x])
        atomicAdd(&ctr[2 + threadIdx.x], lc[threadIdx.x]);
}

// pass 2 (one thread): first item of each class, largest class first; the total
__global__ void k_wide_order(unsigned *ctr)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    unsigned o = 0;
    for (int c = WS_CLASSES - 1; c >= 0; --c) {
        ctr[2 + WS_CLASSES + c] = o;
        o += ctr[2 + c];
    }
    ctr[0] = o;
}

// pass 3: the items, at per-workgroup ranges of their class
__global__ __launch_bounds__(256) void k_wide_fill(const int *__restrict__ rows, const int *__restrict__ count_p,
                                                   const int *__restrict__ Arp, const long long *__restrict__ crp,
                                                   int vals, WideItem *__restrict__ items, long long cap,
                                                   unsigned *ctr, unsigned long long *sctr, long long scratch_n,
                                                   int *err)
{
    __shared__ unsigned lc[WS_CLASSES], lb[WS_CLASSES];
    if (threadIdx.x < WS_CLASSES)
        lc[threadIdx.x] = 0;
    __syncthreads();
    const int idx = blockIdx.x * 256 + threadIdx.x;
    const bool have = idx < *count_p;
    int row = 0, nA = 0, sz = 1, ns = 0, cls = 0;
    unsigned local = 0;
    if (have) {
        wide_row_items(rows, idx, Arp, crp, vals, row, nA, sz, ns, cls);
        local = atomicAdd(&lc[cls], (unsigned)ns);
    }
    __syncthreads();
    if (threadIdx.x < WS_CLASSES && lc[threadIdx.x])
        lb[threadIdx.x] = atomicAdd(&ctr[2 + WS_CLASSES + threadIdx.x], lc[threadIdx.x]);
    __syncthreads();
    if (!have)
        return;
    const unsigned base = lb[cls] + local;
    if ((long long)base + ns > cap) { // cannot happen: cap bounds count + nnz(A) / WS_MIN_SPLIT
        atomicOr(err, ERR_INTERNAL);
        return;
    }
    // a split row's range of the compensated scratch (sctr: nullptr for FP64)
    long long sbase = 0;
    if (sctr && ns > 1) {
        const long long n = crp[row + 1] - crp[row];
        sbase = (long long)atomicAdd(sctr, (unsigned long long)n);
        if (sbase + n > scratch_n) { // cannot happen: scratch_n sums these rows' nnz
            atomicOr(err, ERR_INTERNAL);
            return;
        }
    }
    for (int q = 0; q < ns; ++q) {
        const int e0 = q * sz;
        const int e1 = e0 + sz < nA ? e0 + sz : nA;
        items[base + q] = WideItem{row, e0, e1, ns, sbase};
    }
}

constexpr int WC_BS = 256;     // lanes of a k_wide_cols workgroup
constexpr int WC_DENSE = 1024; // dense tiles listed per pass (more are written by their lane)
constexpr int WC_STAGE = 4096; // columns of a flat chunk staged in LDS for coalesced stores

// Columns of the wide rows, from their kept directories, before k_numeric_wsplit adds the
// values; and, for FP64 (Cval given), zeroed values for the rows split into several items
// (their items add into C; the compensated form's meet in a zeroed scratch instead).  One
// workgroup per row.  Flat directory: WC_BS consecutive tiles' columns are one contiguous
// range of C, so a scan of the tiles' popcounts places each lane's columns in an LDS stage
// that the workgroup then stores coalesced.  Chunks of more than WC_STAGE columns, and
// two-level rows, take the direct path: a lane per tile (flat) or per coarse word,
// walking its tiles (two-level); a tile with at most 4 columns is written by its lane,
// denser tiles are listed in LDS and written by a wave each, one lane per column.
__global__ __launch_bounds__(WC_BS) void k_wide_cols(const int *__restrict__ rows, const int *__restrict__ count_p,
                                                     Operands op, const long long *__restrict__ crp,
                                                     int *__restrict__ Ccol, double *__restrict__ Cval,
                                                     long long sym_row_max, int sym_all_global, int long_rows,
                                                     int vals)
{
    // one buffer, used by one path per chunk: the staged chunk's columns, or the list of
    // its dense tiles (mask, first slot, absolute tile), so that the waves writing them
    // read LDS instead of going back to the directory in global memory
    static_assert(WC_DENSE * (8 + 4 + 4) <= WC_STAGE * 4, "dense-tile list fits the stage");
    __shared__ __align__(16) unsigned char buf[WC_STAGE * 4];
    int *stage = (int *)buf;
    u64 *dmask = (u64 *)buf;
    unsigned *doff = (unsigned *)(buf + 8 * WC_DENSE);
    int *dtile = (int *)(buf + 12 * WC_DENSE);
    __shared__ unsigned red[WC_BS / 64];
    __shared__ int nd_s;
    const int l = threadIdx.x, lane = l & 63, w = l >> 6;
    const int nrows = *count_p;
    for (int idx = blockIdx.x; idx < nrows; idx += gridDim.x) {
        const int row = uniform(rows[idx]);
        const int nA = uniform(op.Arp[row + 1]) - uniform(op.Arp[row]);
        const long long cb = uniform64(crp[row]);
        const long long n = uniform64(crp[row + 1]) - cb;
        if (n <= 0)
            continue;
        if (Cval && nA > wide_split(n, vals)) // a split row: its items add into C
            for (long long q = l; q < n; q += WC_BS)
                Cval[cb + q] = 0.0;
        const int mn = uniform(op.rmin[row]);
        const int S = uniform(op.rmax[row]) - mn + 1;
        const bool two = wide_two_level(S, nA, uniform(op.rtp[row]), sym_row_max, sym_all_global, long_rows);
        const int W = two ? (S + 63) >> 6 : 0;
        const long long rd = uniform64(op.rdir[row]);
        const u64 *cmg = op.dir_mask + rd;
        const unsigned *cog = op.dir_off + rd;
        const u64 *fm = op.dir_mask + rd + W;
        const unsigned *foff = op.dir_off + rd + W;
        auto emit = [&](int k, int tile) {
            u64 m = fm[k];
            const unsigned o = foff[k];
            if (__popcll(m) > 4) {
                const int p = atomicAdd(&nd_s, 1);
                if (p < WC_DENSE) {
                    dmask[p] = m;
                    doff[p] = o;
                    dtile[p] = tile;
                    return;
                }
            }
            long long at = cb + o;
            while (m) {
                Ccol[at++] = (tile << 6) + __ffsll((unsigned long long)m) - 1;
                m &= m - 1;
            }
        };
        const int U = two ? W : S; // units: coarse words, or tiles
        // flat rows: each chunk's tile masks and first slot are loaded one chunk ahead, so
        // the load overlaps the previous chunk's scan and stores
        u64 m_next = !two && l < S ? fm[l] : 0ull;
        unsigned base_next = !two && S > 0 ? uniform(foff[0]) : 0u;
        for (int u0 = 0; u0 < U; u0 += WC_BS) {
            const int u = u0 + l;
            if (!two) { // staged: [foff[u0], foff[u0] + tot) holds this chunk's columns
                const u64 m = m_next;
                const unsigned base = base_next;
                m_next = u + WC_BS < S ? fm[u + WC_BS] : 0ull;
                base_next = u0 + WC_BS < S ? uniform(foff[u0 + WC_BS]) : 0u;
                unsigned tot;
                const unsigned pos = group_excl_scan<WC_BS>((unsigned)__popcll(m), red, tot);
                if (uniform(tot) <= (unsigned)WC_STAGE) {
                    const int c0 = (mn + u) << 6;
                    u64 mm = m;
                    unsigned q = pos;
                    while (mm) {
                        stage[q++] = c0 + __ffsll((unsigned long long)mm) - 1;
                        mm &= mm - 1;
                    }
                    __syncthreads();
                    for (unsigned i = l; i < tot; i += WC_BS)
                        Ccol[cb + base + i] = stage[i];
                    __syncthreads(); // the stage of the next chunk
                    continue;
                }
            }
            if (l == 0)
                nd_s = 0;
            __syncthreads();
            if (u < U) {
                if (two) {
                    u64 cw = cmg[u];
                    int k = (int)cog[u];
                    while (cw) {
                        emit(k++, mn + (u << 6) + __ffsll((unsigned long long)cw) - 1);
                        cw &= cw - 1;
                    }
                } else {
                    emit(u, mn + u);
                }
            }
            __syncthreads();
            const int nd = min(uniform(nd_s), WC_DENSE);
            for (int t = w; t < nd; t += WC_BS / 64) {
                const u64 m = dmask[t];
                if ((m >> lane) & 1ull)
                    Ccol[cb + doff[t] + __popcll(m & below(lane))] = (dtile[t] << 6) + lane;
            }
            __syncthreads(); // the list and nd_s of the next pass
        }
    }
}

// Wide rows (FP64, directory kept): the compact windowed accumulator of k_numeric_wcompact,
// on work items pulled from a queue.  An item holds at most SPLIT entries of one row, so
// their cursors into the B rows live in LDS.  Windows are runs of directory entries that
// share both k / (VW / 2) and (slot of k) / (VW - 64), so a window holds at most VW / 2
// entries and VW slots (VW: WS_VALS, or WS_VALS_BIG for sparse rows); their starts are listed in parallel (WS_LV entries
// per lane, a workgroup scan), WS_WIN at a time.  In a window, the elements of the active
// B rows are spread over the whole workgroup: each active entry gets chunks of WS_CH
// elements from its cursor, as many as its density (elements left / columns left to its
// last column) predicts for the window, and every group of WS_GL lanes takes chunks from
// that list.  B rows are sorted, so a chunk's in-window elements are a prefix of it; the
// smallest (position, column) after a prefix, kept by a 64-bit LDS atomicMin, is the
// entry's new cursor and the column there.  An entry whose chunks were all inside the
// window gets twice as many chunks in another round.  A two-level row's coarse words are staged
// in LDS when there are at most WS_COARSE of them.  Items of a split row add their window
// sums into C (zeroed by k_wide_cols, which also writes every wide row's columns) with
// atomics; the other rows store them.
constexpr int WS_COARSE = 256;
constexpr int WS_BS = 768;   // lanes of a workgroup (two fit a CU: 6 waves per SIMD)
constexpr int WS_WIN = 256;  // window starts listed at a time
constexpr int WS_LV = 8;     // directory entries per lane when listing window starts
constexpr int WS_GL = 16;    // lanes of a chunk group
constexpr int WS_CH = 4 * WS_GL; // elements of a chunk (WS_GL lanes x 4)
constexpr int WS_TCAP = 1024; // chunks listed per round
// dynamic LDS of a workgroup with windows of vw slots (vw / 2 directory entries) and items
// of up to wide_split_max(vw) entries
constexpr int ws_lds(int vw)
{
    const int sp = wide_split_max(vw);
    return (vw * 8 + (vw / 2) * 8 + sp * 8 /* av */ + sp * 8 /* pos:col */ + WS_COARSE * 8 +
            sp * 4 * 5 /* cur kend last, two act lists */ + WS_COARSE * 4 + (WS_WIN + 1) * 4 + WS_TCAP * 4 +
            (vw / 2) * 2 + sp * 2 + 15) & ~15;
}
// (the kernel's static LDS is under 256 bytes)
static_assert(2 * (ws_lds(WS_VALS) + 256) <= 160 * 1024, "two workgroups of the default window per CU");
static_assert(ws_lds(WS_VALS_BIG) + 256 <= 160 * 1024, "one workgroup of the large window per CU");
constexpr u64 WS_NOPOS = 0xffffffffull << 32; // (position, column) with no position
static_assert(2 + 2 * WS_CLASSES <= WIDE_CTRS, "wide-row counters");

// Lanes of a workgroup, by form, window and density.  The default window runs two
// workgroups per CU, the large one a single workgroup.  A window's phases share one
// barrier and one LDS whose atomics queue up, so fewer waves than the CU could hold are
// faster, the more so the sparser the rows (R-MAT scale 20, wide-row bin, default window,
// four chunk tasks in flight per group: FP64 192 / 256 / 320 lanes 258 / 246 / 340 ms,
// compensated FP32 256 / 320 / 384 lanes 266 / 261 / 265 ms; one task at a time: FP64
// 128 / 256 / 512 lanes 355 / 263 / 290 ms; large window, 512 lanes: 287 and 301 ms).
// Dense rows want more lanes (a Zipf A·Aᵀ panel's wide rows, FP64 256 / 384 / 512 / 768
// lanes 41.4 / 38.0 / 37.2 / 39.4 ms; compensated 384 / 512 / 768 lanes 38.2 / 35.0 /
// 34.6 ms).
template <typename P, int VW, bool SPARSE> constexpr int ws_bs()
{
    return VW == WS_VALS_BIG ? 512
                             : (std::is_same<P, PrecF64>::value ? (SPARSE ? 256 : 512) : (SPARSE ? 320 : WS_BS));
}
// waves per SIMD the launch holds (its workgroups per CU x waves / 4 SIMDs), for the
// compiler's register budget
template <int VW, int BS> constexpr int ws_waves()
{
    return (BS / 64) * (VW == WS_VALS_BIG ? 1 : 2) / 4 > 0 ? (BS / 64) * (VW == WS_VALS_BIG ? 1 : 2) / 4 : 1;
}

template <typename P, int VW, int BS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu((ws_waves<VW, BS>())))) void k_numeric_wsplit(
    const WideItem *__restrict__ items, const unsigned *__restrict__ ctr_listed, unsigned *ctr_taken, Operands op,
    const long long *__restrict__ crp, typename P::V *__restrict__ Cval, u64 *__restrict__ scratch, int *ws,
    long long ft_stride, long long sym_row_max, int sym_all_global, int long_rows, int *err)
{
    typedef typename P::V V;
    constexpr bool F32C = std::is_same<P, PrecF32C>::value;
    constexpr int WT = VW / 2; // directory entries of a window
    // chunk tasks a 16-lane group loads before it adds any: a narrow (sparse-row)
    // workgroup has few groups, each with many tasks per round, and the registers for
    // four (R-MAT scale 20 wide-row bin 263 -> 244 ms FP64, 275 -> 265 ms compensated;
    // eight: no better for FP64, 384 ms compensated; two for the dense FP64 launch: within
    // 1% on a Zipf panel)
    constexpr int NTS = BS < 512 ? 4 : 1;
    constexpr int SPLIT = wide_split_max(VW); // entries of an item, at most
    constexpr int EPL = (SPLIT + BS - 1) / BS; // entries a lane posts and moves
    static_assert(SPLIT <= 1024 && WS_TCAP <= 1024, "task: entry in bits 0-9, chunk in 10-19");
    const V *__restrict__ Aval = (const V *)op.Aval;
    const V *__restrict__ Bval = (const V *)op.Bval;
    extern __shared__ __align__(16) unsigned char lds[];
    double *vals = (double *)lds;
    u64 *wm = (u64 *)(vals + VW);
    double *av = (double *)(wm + WT);
    u64 *pc = (u64 *)(av + SPLIT); // (position of the next element << 32) | its column
    u64 *cm = pc + SPLIT;
    int *cur = (int *)(cm + WS_COARSE);
    int *kend = cur + SPLIT;
    int *last = kend + SPLIT; // last column of B_j
    int *act = last + SPLIT; // two lists (rounds alternate)
    unsigned *co = (unsigned *)(act + 2 * SPLIT);
    unsigned *wst = co + WS_COARSE;
    unsigned *task = wst + WS_WIN + 1;
    unsigned short *wo = (unsigned short *)(task + WS_TCAP);
    unsigned short *nch = wo + WT; // chunks listed for an entry in this round
    __shared__ int item_s, nact_s[2], ntask_s[2];
    __shared__ unsigned red[16];
    const int l = threadIdx.x, lane = l & 63;
    const int ll = l & (WS_GL - 1), lg0 = lane & ~(WS_GL - 1);
    int *ft = ws + (size_t)blockIdx.x * ft_stride; // tile of every fine entry of a two-level row
    const unsigned nitems = *ctr_listed;
    while (true) {
        if (l == 0)
            item_s = (int)atomicAdd(ctr_taken, 1u);
        __syncthreads();
        // item and row scalars are uniform: keep them in scalar registers
        const unsigned it = (unsigned)uniform(item_s);
        if (it >= nitems)
            break;
        const int row = uniform(items[it].row), ie0 = uniform(items[it].e0);
        const int ne = uniform(items[it].e1) - ie0, nsplit = uniform(items[it].nsplit);
        const long long sbase = uniform64(items[it].sbase);
        const int a0 = uniform(op.Arp[row]), nA = uniform(op.Arp[row + 1]) - a0;
        const long long cb = uniform64(crp[row]);
        const long long n = uniform64(crp[row + 1]) - cb;
        const int mn = uniform(op.rmin[row]);
        const int S = uniform(op.rmax[row]) - mn + 1;
        const int tp = uniform(op.rtp[row]);
        const long long rd = uniform64(op.rdir[row]);
        const bool two = wide_two_level(S, nA, tp, sym_row_max, sym_all_global, long_rows);
        const int W = two ? (S + 63) >> 6 : 0;
        const u64 *cmg = op.dir_mask + rd;
        const unsigned *cog = op.dir_off + rd;
        const u64 *fm = op.dir_mask + rd + W;
        const unsigned *foff = op.dir_off + rd + W;
        const bool clds = W <= WS_COARSE;
        for (int q = l; q < W && clds; q += BS) {
            cm[q] = cmg[q];
            co[q] = cog[q];
        }
        const int F = two ? (int)(uniform(cog[W - 1]) + (unsigned)__popcll(uniform64(cmg[W - 1]))) : S;
        const unsigned off0 = F > 0 ? uniform(foff[0]) : 0u;
        if (two) {
            for (int q = l; q < W; q += BS) {
                u64 m = cmg[q];
                unsigned k = cog[q];
                while (m) {
                    ft[k++] = (q << 6) + __ffsll((unsigned long long)m) - 1;
                    m &= m - 1;
                }
            }
        }
        for (int e = l; e < ne; e += BS) {
            const int j = op.Acol[a0 + ie0 + e];
            const int k0 = op.Brp[j], k1 = op.Brp[j + 1];
            cur[e] = k0;
            kend[e] = k1;
            const int c0 = k0 < k1 ? op.Bcol[k0] : INT_MAX;
            last[e] = k0 < k1 ? op.Bcol[k1 - 1] : 0;
            pc[e] = WS_NOPOS | (unsigned)c0;
            av[e] = Aval[a0 + ie0 + e];
        }
        block_sync_global();
        const bool split = nsplit > 1;
        long long done = 0;
        int kb = 0;         // first entry whose window start is not listed yet
        int nw = 0, wi = 0; // windows listed, current window
        while (true) {
            if (wi == nw) { // list the next window starts (WS_WIN at a time, in order)
                if (kb >= F)
                    break;
                int cnt_st = 0; // starts listed so far
                // WS_LV consecutive entries per lane, their offsets loaded together: one
                // round trip covers BS * WS_LV entries
                for (int kk = kb; kk < F && cnt_st <= WS_WIN; kk += BS * WS_LV) {
                    const int kl = kk + l * WS_LV;
                    unsigned f[WS_LV + 1]; // foff[kl - 1 + i]
#pragma unroll
                    for (int i = 0; i <= WS_LV; ++i) {
                        const int k = kl - 1 + i;
                        f[i] = k >= kb && k < F ? foff[k] : 0u;
                    }
                    unsigned bits = 0;
#pragma unroll
                    for (int i = 0; i < WS_LV; ++i) {
                        const int k = kl + i;
                        bool st = false;
                        if (k < F) {
                            if (k == kb || (k % WT) == 0)
                                st = true;
                            else
                                st = (f[i + 1] - off0) / (unsigned)(VW - 64) != (f[i] - off0) / (unsigned)(VW - 64);
                        }
                        bits |= (st ? 1u : 0u) << i;
                    }
                    unsigned tot;
                    unsigned pos = group_excl_scan<BS>((unsigned)__popc(bits), red, tot);
                    while (bits) { // this lane's starts, in order
                        if (cnt_st + (int)pos <= WS_WIN)
                            wst[cnt_st + pos] = (unsigned)(kl + __ffs(bits) - 1);
                        ++pos;
                        bits &= bits - 1;
                    }
                    cnt_st += (int)uniform(tot);
                    __syncthreads(); // red is reused by the next scan
                }
                // windows: starts 0..m-1 complete; the last listed start (or F) ends them
                int m;
                if (cnt_st > WS_WIN) {
                    m = WS_WIN;
                    kb = (int)uniform(wst[WS_WIN]);
                } else {
                    m = cnt_st;
                    if (l == 0)
                        wst[m] = (unsigned)F;
                    kb = F;
                }
                nw = m;
                wi = 0;
                __syncthreads();
                if (nw == 0)
                    break;
            }
            const int k0 = (int)uniform(wst[wi]), k1 = (int)uniform(wst[wi + 1]);
            ++wi;
            const int nk = k1 - k0;
            const unsigned base = uniform(foff[k0]);
            const int cnt = (int)((k1 == F ? n : (long long)uniform(foff[k1])) - base);
            if (l == 0) {
                nact_s[0] = 0;
                ntask_s[0] = 0;
            }
            for (int q = l; q < nk; q += BS) {
                wm[q] = fm[k0 + q];
                wo[q] = (unsigned short)(foff[k0 + q] - base);
            }
            for (int q = l; q < cnt; q += BS)
                vals[q] = 0.0;
            const int lastt = two ? uniform(ft[k1 - 1]) : k1 - 1;
            const int chi = (mn + lastt + 1) << 6; // end of the window's last tile
            __syncthreads();
            // Chunks of entry e (has) for round r into the lists of buffer r & 1: as many as
            // its density predicts for the window, doubled in every further round (the
            // density was too low for this window).  Every lane of every wave calls it: the
            // chunk counts are laid out with one wave scan and one LDS atomic per wave.
            auto post = [&](bool has, int e, int r) {
                const int buf = r & 1;
                int want = 0;
                if (has) {
                    const int c0 = (int)(unsigned)pc[e];
                    const long long rem = kend[e] - cur[e];
                    long long est = rem;
                    if (last[e] >= chi) { // the row goes past the window: scale by its columns
                        const long long span = (long long)last[e] - c0 + 1;
                        est = rem * (long long)(chi - c0) / span;
                        est = ((est + est / 4 + 8) << (r < 8 ? r : 8));
                        est = est < rem ? est : rem;
                    }
                    want = (int)((est + WS_CH - 1) / WS_CH);
                    want = want < 1 ? 1 : want;
                }
                const int incl = wave_incl_scan(want);
                const int wtot = __shfl(incl, 63, 64);
                const u64 mh = __ballot(has);
                int tb = 0, ab = 0;
                if (lane == 0) {
                    if (wtot)
                        tb = atomicAdd(&ntask_s[buf], wtot);
                    if (mh)
                        ab = atomicAdd(&nact_s[buf], __popcll(mh));
                }
                tb = __shfl(tb, 0, 64);
                ab = __shfl(ab, 0, 64);
                if (has) {
                    const int off = tb + incl - want;
                    const int got = off >= WS_TCAP ? 0 : (off + want > WS_TCAP ? WS_TCAP - off : want);
                    nch[e] = (unsigned short)got;
                    for (int c = 0; c < got; ++c)
                        task[off + c] = (unsigned)e | ((unsigned)c << 10);
                    act[buf * SPLIT + ab + mbcnt(mh)] = e;
                }
            };
            auto add_c = [&](int c, V v, double a) {
                const int tt = (c >> 6) - mn;
                int r = tt - k0;
                if (two) {
                    const int q = tt >> 6;
                    const u64 m = clds ? cm[q] : cmg[q];
                    const unsigned o = clds ? co[q] : cog[q];
                    r = (int)(o + (unsigned)__popcll(m & below(tt & 63))) - k0;
                }
                P::add(vals, (unsigned)wo[r] + (unsigned)__popcll(wm[r] & below(c & 63)), a, v);
            };
            // round 0: the entries with elements in this window; then those that go again
            // (EPL entries per lane)
            bool has[EPL];
            int e[EPL];
#pragma unroll
            for (int q = 0; q < EPL; ++q) {
                e[q] = l + q * BS;
                has[q] = e[q] < ne && (int)(unsigned)pc[e[q]] < chi;
            }
            for (int round = 0;; ++round) {
#pragma unroll
                for (int q = 0; q < EPL; ++q)
                    post(has[q], e[q], round);
                __syncthreads();
                const int buf = round & 1;
                const int nact = uniform(nact_s[buf]);
                if (nact == 0)
                    break;
                const int nt = uniform(ntask_s[buf]);
                const int ntask = nt < WS_TCAP ? nt : WS_TCAP;
                if (l == 0) { // the other buffer's counters, for the next round
                    nact_s[buf ^ 1] = 0;
                    ntask_s[buf ^ 1] = 0;
                }
                // a group loads the chunks of NTS tasks before it adds any of them
                constexpr int NG = BS / WS_GL;
                for (int t = l / WS_GL; t < ntask; t += NTS * NG) {
                    int te[NTS], k[NTS], ke[NTS];
                    double a[NTS];
                    int c[NTS][4];
                    V v[NTS][4];
#pragma unroll
                    for (int h = 0; h < NTS; ++h) {
                        const int th = t + h * NG;
                        te[h] = 0;
                        k[h] = 0;
                        ke[h] = 0; // no task: no element loaded
                        a[h] = 0.0;
                        if (th < ntask) {
                            const unsigned tk = task[th];
                            te[h] = (int)(tk & 1023u);
                            k[h] = cur[te[h]] + (int)(tk >> 10) * WS_CH;
                            ke[h] = kend[te[h]];
                            a[h] = av[te[h]];
                        }
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const int p = k[h] + ll + WS_GL * u;
                            c[h][u] = INT_MAX;
                            v[h][u] = 0;
                            if (p < ke[h]) {
                                c[h][u] = op.Bcol[p];
                                v[h][u] = Bval[p];
                            }
                        }
                    }
#pragma unroll
                    for (int h = 0; h < NTS; ++h) {
                        if (t + h * NG >= ntask) // uniform in the group (t is)
                            break;
                        int took = 0;
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const bool in = c[h][u] < chi;
                            if (in)
                                add_c(c[h][u], v[h][u], a[h]);
                            took += __popcll((__ballot(in) >> lg0) & ((1ull << WS_GL) - 1ull));
                        }
                        if (took < WS_CH) { // the prefix ends here: offer (position, column) after it
#pragma unroll
                            for (int u = 0; u < 4; ++u)
                                if (ll + WS_GL * u == took)
                                    atomicMin((unsigned long long *)&pc[te[h]],
                                              ((u64)(unsigned)(k[h] + took) << 32) | (unsigned)c[h][u]);
                        }
                    }
                }
                __syncthreads();
                // move the cursors; entries whose chunks were all inside the window go again
#pragma unroll
                for (int q = 0; q < EPL; ++q) {
                    const int idx = l + q * BS;
                    bool again = false;
                    int ee = -1;
                    if (idx < nact) {
                        ee = act[buf * SPLIT + idx];
                        const u64 r = pc[ee];
                        if ((r >> 32) != 0xffffffffull) {
                            cur[ee] = (int)(r >> 32);
                            pc[ee] = WS_NOPOS | (r & 0xffffffffull);
                        } else {
                            const int kn = cur[ee] + (int)nch[ee] * WS_CH;
                            if (kn >= kend[ee]) {
                                cur[ee] = kend[ee];
                                pc[ee] = WS_NOPOS | (unsigned)INT_MAX;
                            } else {
                                cur[ee] = kn;
                                again = true;
                            }
                        }
                    }
                    has[q] = again;
                    e[q] = ee;
                }
            }
            // Output: the window's values (its columns were written by k_wide_cols).  Split
            // rows add theirs into C (FP64) or, compensated, s + c into the FP64 scratch.
            if constexpr (F32C) {
                const float *fv = (const float *)vals;
                if (split) {
                    // the item's compensated sums meet in FP64 (s + c is exact in a double),
                    // one atomic add each, no retries; rounded to FP32 once, at the end
                    double *sd = (double *)scratch + sbase + base;
                    for (int q = l; q < cnt; q += BS) {
                        const float sv = fv[2 * q], cv = fv[2 * q + 1];
                        if (sv != 0.0f || cv != 0.0f)
                            unsafeAtomicAdd(&sd[q], (double)sv + (double)cv);
                    }
                } else {
                    for (int q = l; q < cnt; q += BS)
                        Cval[cb + base + q] = fv[2 * q] + fv[2 * q + 1];
                }
            } else if (split) {
                for (int q = l; q < cnt; q += BS) {
                    const double v = vals[q];
                    if (v != 0.0)
                        unsafeAtomicAdd(&Cval[cb + base + q], v);
                }
            } else {
                for (int q = l; q < cnt; q += BS)
                    Cval[cb + base + q] = vals[q];
            }
            done += cnt;
            __syncthreads();
        }
        if (l == 0 && done != n)
            atomicOr(err, ERR_INTERNAL);
        block_sync_global(); // ft and the LDS state of the next item
    }
}

// Compensated FP32, rows split into several items: C's values, rounded once from the
// FP64 sums the items added into the scratch.  A workgroup per row; the row's first item
// names it and its scratch range.
__global__ __launch_bounds__(256) void k_wide_final(const WideItem *__restrict__ items,
                                                    const unsigned *__restrict__ ctr_listed,
                                                    const long long *__restrict__ crp, float *__restrict__ Cval,
                                                    const u64 *__restrict__ scratch)
{
    const unsigned nitems = *ctr_listed;
    for (unsigned idx = blockIdx.x; idx < nitems; idx += gridDim.x) {
        const int e0 = uniform(items[idx].e0), ns = uniform(items[idx].nsplit);
        if (e0 != 0 || ns <= 1)
            continue;
        const int row = uniform(items[idx].row);
        const long long sb = uniform64(items[idx].sbase);
        const long long cb = uniform64(crp[row]), n = uniform64(crp[row + 1]) - cb;
        const double *sd = (const double *)scratch + sb;
        for (long long q = threadIdx.x; q < n; q += 256)
            Cval[cb + q] = (float)sd[q];
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

template <int G, bool LONG>
static void symbolic_g(const int *rows, const int *count_dev, int count, int max_need, const Operands &op, int *rnnz,
                       int tshift, int two_level, int num_cus, unsigned *queue, hipStream_t s)
{
    constexpr int BS = G < BLOCK ? BLOCK : G;
    constexpr int RPB = BS / G;
    const int slice = (max_need + 15) & ~15;
    const size_t lds = (size_t)slice * RPB;
    const int64_t blocks =
        row_grid(k_symbolic_lds<G, LONG>, BS, lds, ((int64_t)count + RPB - 1) / RPB, num_cus, queue);
    const int lg = __builtin_ctz(G);
    hipLaunchKernelGGL((k_symbolic_lds<G, LONG>), dim3((unsigned)blocks), dim3(BS), lds, s, rows, count_dev, slice, op,
                       rnnz, tshift > lg ? lg : tshift, two_level, queue);
}

template <bool LONG>
static void symbolic_l(int bin, const int *rows, const int *count_dev, int count, int max_need, const Operands &op,
                       int *rnnz, int tshift, int two_level, int num_cus, unsigned *queue, hipStream_t s)
{
#define SYM(B, G)                                                                                                      \
    case B: symbolic_g<G, LONG>(rows, count_dev, count, max_need, op, rnnz, tshift, two_level, num_cus, queue, s); break;
    switch (bin) {
        SYM(0, 8) SYM(1, 16) SYM(2, 32) SYM(3, 64) SYM(4, 128) SYM(5, 256) SYM(6, 512) SYM(7, 1024)
    }
#undef SYM
}

hipError_t launch_symbolic(int bin, const int *rows, const int *count_dev, int count, int max_need, const Operands &op,
                           int *rnnz, int tshift, int two_level, int long_rows, u64 *gws, long long gws_stride,
                           int num_cus, int gws_blocks, int lds_dyn, unsigned *queue, hipStream_t s)
{
    if (count <= 0)
        return hipSuccess;
    if (bin == BIN_GLOBAL) {
        const int blocks = count < gws_blocks ? count : gws_blocks;
        const size_t dir_lds = (size_t)gws_stride * sizeof(u64); // >= 8 bytes per tile of the widest span
        if (dir_lds <= (size_t)lds_dyn)
            hipLaunchKernelGGL(k_symbolic_wide, dim3(blocks), dim3(1024), dir_lds, s, rows, count_dev, op, rnnz,
                               tshift > 10 ? 10 : tshift);
        else
            hipLaunchKernelGGL(k_symbolic_global, dim3(blocks), dim3(1024), 0, s, rows, count_dev, op, rnnz,
                               tshift > 10 ? 10 : tshift, gws, gws_stride);
    } else if (bin >= 0 && bin < NLDS) {
        if (long_rows)
            symbolic_l<true>(bin, rows, count_dev, count, max_need, op, rnnz, tshift, two_level, num_cus, queue, s);
        else
            symbolic_l<false>(bin, rows, count_dev, count, max_need, op, rnnz, tshift, two_level, num_cus, queue, s);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int G, typename P, bool LONG>
static void numeric_g(const int *rows, const int *count_dev, int count, int max_need, const Operands &op,
                      const long long *crp, int *Ccol, void *Cval, int bshift, int tshift, int two_level, int *err,
                      int num_cus, unsigned *queue, unsigned long long *lds_bytes, hipStream_t s)
{
    constexpr int BS = G < BLOCK ? BLOCK : G;
    constexpr int RPB = BS / G;
    const int slice = (max_need + 15) & ~15;
    const size_t lds = (size_t)slice * RPB;
    const int64_t blocks =
        row_grid(k_numeric_lds<G, P, LONG>, BS, lds, ((int64_t)count + RPB - 1) / RPB, num_cus, queue);
    *lds_bytes += (unsigned long long)blocks * lds;
    const int lg = __builtin_ctz(G);
    hipLaunchKernelGGL((k_numeric_lds<G, P, LONG>), dim3((unsigned)blocks), dim3(BS), lds, s, rows, count_dev, slice,
                       op, crp, Ccol, (typename P::V *)Cval, bshift > lg ? lg : bshift, tshift > lg ? lg : tshift,
                       two_level, err, queue);
}

template <typename P, bool LONG>
static void numeric_l(int bin, const int *rows, const int *count_dev, int count, int max_need, const Operands &op,
                      const long long *crp, int *Ccol, void *Cval, int bshift, int tshift, int two_level, int *err,
                      int num_cus, unsigned *queue, unsigned long long *lds_bytes, hipStream_t s)
{
#define NUM(B, G)                                                                                                      \
    case B:                                                                                                            \
        numeric_g<G, P, LONG>(rows, count_dev, count, max_need, op, crp, Ccol, Cval, bshift, tshift, two_level, err,   \
                              num_cus, queue, lds_bytes, s);                                                           \
        break;
    switch (bin) {
        NUM(0, 8) NUM(1, 16) NUM(2, 32) NUM(3, 64) NUM(4, 128) NUM(5, 256) NUM(6, 512) NUM(7, 1024)
    }
#undef NUM
}

template <typename P>
static hipError_t numeric_p(int bin, const int *rows, const int *count_dev, int count, int max_need, const Operands &op,
                            const long long *crp, int *Ccol, void *Cval, int bshift, int tshift, int two_level,
                            int long_rows, const WideRows &wide, int *err, int num_cus, unsigned *queue,
                            unsigned long long *lds_bytes, hipStream_t s)
{
    if (bin == BIN_GLOBAL) {
        const int blocks = count < wide.blocks ? count : wide.blocks;
        if (op.dir_mask && wide.items) { // split work items (FP64, or compensated with its scratch)
            constexpr bool F32C = std::is_same<P, PrecF32C>::value;
            hipError_t e = launch_fill_zero(wide.ctr, WIDE_CTRS * sizeof(unsigned), s);
            if (e == hipSuccess && F32C)
                e = launch_fill_zero(wide.sctr, sizeof(unsigned long long), s);
            if (e == hipSuccess && F32C && wide.scratch_n > 0)
                e = launch_fill_zero(wide.scratch, (size_t)wide.scratch_n * sizeof(u64), s);
            if (e != hipSuccess)
                return e;
            const unsigned gr = (unsigned)((count + 255) / 256);
            // (the compensated scratch is sized by wide_split at WS_VALS in the binning; the
            // large window splits a subset of those rows, in fewer items)
            const int vals = wide.vals == WS_VALS_BIG ? WS_VALS_BIG : WS_VALS;
            hipLaunchKernelGGL(k_wide_count, dim3(gr), dim3(256), 0, s, rows, count_dev, op.Arp, crp, vals, wide.ctr);
            hipLaunchKernelGGL(k_wide_order, dim3(1), dim3(64), 0, s, wide.ctr);
            hipLaunchKernelGGL(k_wide_fill, dim3(gr), dim3(256), 0, s, rows, count_dev, op.Arp, crp, vals, wide.items,
                               wide.items_cap, wide.ctr, F32C ? wide.sctr : nullptr, wide.scratch_n, err);
            hipLaunchKernelGGL(k_wide_cols, dim3((unsigned)(count < 8192 ? count : 8192)), dim3(WC_BS), 0, s, rows,
                               count_dev, op, crp, Ccol, F32C ? nullptr : (double *)Cval, wide.sym_row_max,
                               wide.sym_all_global, wide.long_rows, vals);
#define WSPLIT(VW, SPARSE)                                                                                           \
    hipLaunchKernelGGL((k_numeric_wsplit<P, VW, ws_bs<P, VW, SPARSE>()>), dim3((unsigned)wide.blocks),                 \
                       dim3(ws_bs<P, VW, SPARSE>()), ws_lds(VW), s, wide.items, wide.ctr, wide.ctr + 1, op, crp,        \
                       (typename P::V *)Cval, (u64 *)wide.scratch, wide.ws, wide.ft_stride, wide.sym_row_max,           \
                       wide.sym_all_global, wide.long_rows, err)
            if (vals == WS_VALS_BIG)
                WSPLIT(WS_VALS_BIG, false);
            else if (wide.sparse)
                WSPLIT(WS_VALS, true);
            else
                WSPLIT(WS_VALS, false);
#undef WSPLIT
            if (F32C)
                hipLaunchKernelGGL(k_wide_final, dim3(2048), dim3(256), 0, s, wide.items, wide.ctr, crp, (float *)Cval,
                                   (const u64 *)wide.scratch);
        } else if (op.dir_mask)
            hipLaunchKernelGGL(k_numeric_wcompact<P>, dim3(blocks), dim3(1024), CW_LDS, s, rows, count_dev, op, crp,
                               Ccol, (typename P::V *)Cval, bshift > 6 ? 6 : bshift, wide.ws, wide.cur_stride,
                               wide.ft_stride, wide.sym_row_max, wide.sym_all_global, wide.long_rows, err);
        else
            hipLaunchKernelGGL(k_numeric_window<P>, dim3(blocks), dim3(1024), WIN_LDS, s, rows, count_dev, op, crp, Ccol,
                               (typename P::V *)Cval, bshift > 6 ? 6 : bshift, wide.ws, wide.cur_stride, wide.ft_stride,
                               err);
    } else if (bin >= 0 && bin < NLDS) {
        if (long_rows)
            numeric_l<P, true>(bin, rows, count_dev, count, max_need, op, crp, Ccol, Cval, bshift, tshift, two_level,
                               err, num_cus, queue, lds_bytes, s);
        else
            numeric_l<P, false>(bin, rows, count_dev, count, max_need, op, crp, Ccol, Cval, bshift, tshift, two_level,
                                err, num_cus, queue, lds_bytes, s);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_numeric(int bin, int fp32c, const int *rows, const int *count_dev, int count, int max_need,
                          const Operands &op, const long long *crp, int *Ccol, void *Cval, int bshift, int tshift,
                          int two_level, int long_rows, const WideRows &wide, int *err, int num_cus, unsigned *queue,
                          unsigned long long *lds_bytes, hipStream_t s)
{
    if (count <= 0)
        return hipSuccess;
    unsigned long long ignored = 0;
    if (!lds_bytes)
        lds_bytes = &ignored;
    if (fp32c)
        return numeric_p<PrecF32C>(bin, rows, count_dev, count, max_need, op, crp, Ccol, Cval, bshift, tshift, two_level,
                                   long_rows, wide, err, num_cus, queue, lds_bytes, s);
    return numeric_p<PrecF64>(bin, rows, count_dev, count, max_need, op, crp, Ccol, Cval, bshift, tshift, two_level,
                              long_rows, wide, err, num_cus, queue, lds_bytes, s);
}

size_t wide_ws_ints(long long cur_stride, long long ft_stride, int blocks)
{
    return (size_t)blocks * (size_t)(3 * cur_stride + ft_stride);
}

hipError_t configure_kernels(int lds_max, int num_xcds)
{
    const int nx = num_xcds < 1 ? 1 : (num_xcds > MAX_XCDS ? MAX_XCDS : num_xcds);
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_nxcd), &nx, sizeof(int));
#define CFG(K)                                                                                                         \
    {                                                                                                                  \
        hipError_t r = hipFuncSetAttribute((const void *)(K), hipFuncAttributeMaxDynamicSharedMemorySize, lds_max);    \
        if (r != hipSuccess)                                                                                           \
            e = r;                                                                                                     \
    }
#define CFGL(L)                                                                                                        \
    CFG((k_symbolic_lds<8, L>)) CFG((k_symbolic_lds<16, L>)) CFG((k_symbolic_lds<32, L>))                             \
    CFG((k_symbolic_lds<64, L>)) CFG((k_symbolic_lds<128, L>)) CFG((k_symbolic_lds<256, L>))                          \
    CFG((k_symbolic_lds<512, L>)) CFG((k_symbolic_lds<1024, L>))
#define CFGP(P, L)                                                                                                     \
    CFG((k_numeric_lds<8, P, L>)) CFG((k_numeric_lds<16, P, L>)) CFG((k_numeric_lds<32, P, L>))                       \
    CFG((k_numeric_lds<64, P, L>)) CFG((k_numeric_lds<128, P, L>)) CFG((k_numeric_lds<256, P, L>))                    \
    CFG((k_numeric_lds<512, P, L>)) CFG((k_numeric_lds<1024, P, L>))
    CFGL(false) CFGL(true)
    CFGP(PrecF64, false) CFGP(PrecF64, true) CFGP(PrecF32C, false) CFGP(PrecF32C, true)
#undef CFGP
#undef CFGL
    CFG(k_symbolic_wide)
    CFG(k_numeric_window<PrecF64>) CFG(k_numeric_window<PrecF32C>)
    CFG(k_numeric_wcompact<PrecF64>) CFG(k_numeric_wcompact<PrecF32C>) 
#define CFGW(P)                                                                                                        \
    CFG((k_numeric_wsplit<P, WS_VALS, ws_bs<P, WS_VALS, false>()>))                                                    \
    CFG((k_numeric_wsplit<P, WS_VALS, ws_bs<P, WS_VALS, true>()>))                                                     \
    CFG((k_numeric_wsplit<P, WS_VALS_BIG, ws_bs<P, WS_VALS_BIG, false>()>))
    CFGW(PrecF64) CFGW(PrecF32C)
#undef CFGW
#undef CFG
    return e;
}

} // namespace hsmu
