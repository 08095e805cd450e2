// hsmu_internal.h -- declarations shared by the kernels (kernels.hip) and the host
// orchestrator (api.cpp).  Not part of the public ABI.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <mutex>

#include <hip/hip_runtime_api.h>

namespace hsmu {

// Row bins.  Bins 0..7 keep a row in LDS with G = 8 << b lanes per row (rows per
// workgroup = max(1, BLOCK / G)).  BIN_GLOBAL keeps the row's tile directory in global
// memory and accumulates straight into C (rows whose LDS need exceeds the limit).
constexpr int NLDS = 8;
constexpr int BIN_GLOBAL = 8;
constexpr int NBINS = 10;  // slot 9 reserved
// Workgroup size of the bins that pack several rows per workgroup (G < 64 lanes per
// row): one wave, so the phases of a row group are separated by single-wave barriers and
// every wave runs at its own pace (measured: 4-wave workgroups of 8-lane rows spend most
// of their time waiting at barriers for the slowest wave).
constexpr int BLOCK = 64;

// error bits raised on the device, reported at the next host synchronisation
constexpr int ERR_ROWPTR = 1;   // row pointer decreasing or out of [0, nnz]
constexpr int ERR_ACOL = 2;     // column of A outside [0, rows of B)
constexpr int ERR_BCOL = 4;     // column of B outside [0, N)
constexpr int ERR_UNSORTED = 8; // row of B with decreasing columns
constexpr int ERR_INTERNAL = 16;// symbolic and numeric counts disagree (a bug)
constexpr int ERR_PLAN = 32;    // a replayed plan's inputs changed structure

struct BinParams {
    long long lds_block_budget; // LDS bytes per workgroup for bins with several rows per workgroup
    long long lds_row_max;      // LDS bytes one row may use
    long long per_lane;         // work items per lane
    long long lds_cu;           // LDS bytes of a CU
    int force_global;
    int min_waves;              // waves per CU a one-row workgroup's width should reach
};

struct BinStats {
    int count[NBINS];
    int max_need[NBINS]; // LDS bytes of the largest row in the bin
    int max_span[NBINS]; // tiles spanned by the widest row in the bin
    unsigned long long work[NBINS]; // products (numeric) or tile-products (symbolic)
    unsigned long long nnz_c[NBINS];// nnz of C in the bin's rows (numeric)
    unsigned long long nnz_a[NBINS];// nnz of A in the bin's rows
    int nsparse[NBINS];             // rows with a two-level tile directory
    int max_na[NBINS];              // longest row of A in the bin
};

// XCDs (accelerator complex dies, each with its own L2) the kernels map work onto: the
// device's count (hipDeviceAttributeNumberOfXccs; 8 on an MI355X, fewer in partitioned
// modes), at most this many.
constexpr int MAX_XCDS = 8;

// Device-side counters of one call, copied to pinned host memory at the two
// synchronisation points.
struct DevStats {
    unsigned long long dir_total; // directory entries reserved for rows in LDS bins
    long long nnz_c;
    int err;
    int ntiles_b;
    int nlong; // rows of A set aside for the one-workgroup-per-row analysis
    int max_brow; // longest row of B (counted from LONG_TILES on)
    unsigned long long num_wide_hint; // rows whose numeric need may exceed LDS (symbolic binning)
    unsigned long long wide_split_nnz; // nnz(C_i) summed over the wide rows split into several items
    unsigned long long wide_span; // tiles spanned, summed over the wide rows (numeric binning)
    BinStats sym;
    BinStats num;
    // work-queue counters of the persistent LDS kernels: [phase][bin][XCD]
    unsigned queue[2][NBINS][MAX_XCDS];
};

// Batched form: instance i of a batch, as seen by the stacking kernels.  The stacked A
// holds A_i's rows at mo.. and its entries at ea.., with columns shifted by ko (the first
// row of B_i in the stacked B); the stacked B holds B_i's rows at ko.. and entries at
// eb.., columns unchanged (so rows of the stacked C are the rows of C_i, columns local).
struct BatchInst {
    const int *Arp, *Acol;
    const void *Aval;
    const int *Brp, *Bcol;
    const void *Bval;
    int M, K, N, pad;
};

// Offsets of the instances, each array count+1 long (the last entry is the total):
// rows of A (mo), rows of B (ko), entries of A (ea), entries of B (eb).
struct BatchOffsets {
    const long long *mo, *ko, *ea, *eb;
};

#if defined(__HIPCC__)
#define HSMU_HD __host__ __device__
#else
#define HSMU_HD
#endif

// Accumulator form of a row in the numeric LDS kernel.  Compact (the HSMU form): exactly
// nnz(C_i) slots, column c in slot doff[t] + popcount(dm[t] & below(c & 63)).  Dense:
// one slot per column of the row's span, column c in slot c - 64*rmin, used when the
// span is at least two-thirds full (then the dense array costs at most 1.5x the compact
// one and saves the two directory reads per product) and at most 128 tiles wide.
HSMU_HD inline bool use_dense(long long S, long long n) { return S <= 128 && 2 * 64 * S <= 3 * n; }

// Tile directory of a row of C over its span [rmin, rmax] of S tiles.
// Flat: one entry (mask, first slot) per tile of the span, tile t at index t.
// Two-level, for rows whose span is mostly empty (the 3D stencil's A^2 row touches ~10 of
// ~140 tiles): a coarse bitmap of W = ceil(S/64) words marks the non-empty tiles, with
// coff[w] the compact index of the first non-empty tile of word w; the fine level holds
// only the non-empty tiles, tile t at compact index coff[t>>6] + popcount(cm[t>>6] &
// below(t&63)) -- the HSMU rank trick applied to tiles as well as to columns, so a slot
// is still found in O(1).  At most R = min(S, tile-products) fine entries.  Chosen when
// it at least halves the directory (never for spans of up to 64 tiles).
HSMU_HD inline long long fine_cap(long long S, long long tp) { return tp < S ? tp : S; }
HSMU_HD inline bool use_sparse(long long S, long long tp)
{
    return S > 64 && 2 * ((S + 63) / 64 + fine_cap(S, tp)) <= S;
}
// directory entries of a row (12 bytes each when kept between the phases)
HSMU_HD inline long long dir_entries(long long S, long long tp)
{
    return use_sparse(S, tp) ? (S + 63) / 64 + fine_cap(S, tp) : S;
}

// LDS slice of one row in the LDS kernels (byte offsets), for a row whose C row spans
// S tiles, has n = nnz(C_i) entries, nA entries in A and tp tile-products; E =
// dir_entries(S, tp):
//   dm   u64[E]     tile masks (flat), or coarse words then fine masks (two-level)
//   doff u32[E]     slot of each tile's first column (numeric), coarse offsets (two-level)
//   ft   u32[R]     numeric, two-level: tile of each fine entry, relative to rmin
//   vals f64[n]     numeric: accumulators (64*S of them in the dense form)
//   av   f64[nA]    numeric: a_ij of the row's entries
//   tlen i32[nA]    tiles(B_j) for each entry j of the row
//   t0   i32[nA]    first tile of B_j
//   blen i32[nA]    numeric: nnz(B_j)
//   b0   i32[nA]    numeric: first stored element of B_j
//   lidx i32[nA]    with_long: entries whose B_j (numeric) or tile row (symbolic) is
//                   long, walked by whole waves
//   lcnt i32        with_long: their number (16 bytes reserved)
// The size is rounded up to 16 bytes so per-row slices of a workgroup stay aligned.
struct RowLayout {
    long long dm, doff, ft, vals, av, tlen, t0, blen, b0, lidx, lcnt, bytes;
};

// Rows of B at least this long (LONG_TILES for the tile OR) are walked by a whole wave:
// with_long = the longest row of B reaches LONG_TILES, so the kernels list such entries.
constexpr int LONG_SEG = 256;
constexpr int LONG_TILES = 128;

HSMU_HD inline RowLayout row_layout(int numeric, long long S, long long n, long long nA, long long tp, bool with_long)
{
    RowLayout L;
    const bool sparse = use_sparse(S, tp);
    const long long E = dir_entries(S, tp);
    long long o = 0;
    L.dm = o;
    o += 8 * E;
    L.doff = o;
    if (numeric || sparse)
        o += 4 * E;
    L.ft = o;
    if (numeric && sparse)
        o += 4 * fine_cap(S, tp);
    o = (o + 7) & ~7LL;
    L.vals = o;
    if (numeric)
        o += 8 * (!sparse && use_dense(S, n) ? 64 * S : n);
    L.av = o;
    if (numeric)
        o += 8 * nA;
    L.tlen = o;
    o += 4 * nA;
    L.t0 = o;
    o += 4 * nA;
    L.blen = o;
    if (numeric)
        o += 4 * nA;
    L.b0 = o;
    if (numeric)
        o += 4 * nA;
    L.lidx = o;
    if (with_long)
        o += 4 * nA;
    o = (o + 15) & ~15LL;
    L.lcnt = o;
    if (with_long)
        o += 16;
    L.bytes = (o + 15) & ~15LL;
    return L;
}

HSMU_HD inline long long sym_need(long long S, long long nA, long long tp, bool with_long)
{
    return row_layout(0, S, 0, nA, tp, with_long).bytes;
}
HSMU_HD inline long long num_need(long long S, long long n, long long nA, long long tp, bool with_long)
{
    return row_layout(1, S, n, nA, tp, with_long).bytes;
}

// LDS kept back from the dynamic allocation for the kernels' static __shared__ arrays
constexpr int LDS_STATIC_RESERVE = 256;

// ---- launchers (kernels.hip); all enqueue on `s` and return the launch status ----

hipError_t launch_fill_zero(void *p, size_t bytes, hipStream_t s);

// B -> tiles: per-row tile counts, scan to tptr, fill tcol/tmask.
// scan_ws must hold scan_ws_bytes(K) bytes.
size_t scan_ws_bytes(int64_t n);
// ws must hold btiles_ws_bytes(nnz) bytes, scan_ws scan_ws_bytes(nnz + 1).
size_t btiles_ws_bytes(int64_t nnz);
// One row of B as the row analysis needs it: entries, tiles, first and last tile column
// (last = -1 for an empty row).
struct alignas(16) BRow {
    int len, ntiles, first, last;
};
// brow (K entries, may be null): each row's BRow, written beside the tile row pointers.
hipError_t launch_btiles(int64_t K, int64_t N, int64_t nnz, const int *rp, const int *col,
                         int *ws, int *tptr, int *tcol, unsigned long long *tmask, BRow *brow,
                         void *scan_ws, DevStats *st, int num_cus, hipStream_t s);

// Per row of A: products, tile-products, tile span (from the BRow of each B_j it names).
// longrows (M ints) is scratch for the list of long rows analysed one workgroup per row.
hipError_t launch_row_info(int64_t M, int64_t K, int64_t nnzA, const int *Arp, const int *Acol,
                           const BRow *brow, long long *rprod, int *rtp, int *rmin, int *rmax,
                           int *rnnz, int *longrows, DevStats *st, int avg_a, int num_cus, hipStream_t s);

// Bin rows for the symbolic (numeric = 0) or numeric (numeric = 1) phase.  The symbolic
// pass also reserves each LDS-bin row's directory range: rdir[i] = offset (span entries),
// -1 for rows not in an LDS bin; st->dir_total = entries reserved.
hipError_t launch_bin(int numeric, int64_t M, const int *Arp, const long long *rprod,
                      const int *rtp, const int *rmin, const int *rmax, const int *rnnz,
                      BinParams bp, int *binrows, long long *rdir, DevStats *st, int num_cus,
                      hipStream_t s);

struct Operands {
    const int *Arp, *Acol;
    const void *Aval; // double, or float when val32
    const int *Brp, *Bcol;
    const void *Bval;
    int val32;
    const int *tptr, *tcol;
    const unsigned long long *tmask;
    const int *rmin, *rmax;
    const int *rtp; // tile-products per row (chooses the directory form)
    // Row directories kept between the phases (nullptr when not kept): the symbolic LDS
    // kernel stores row i's directory (flat: tile masks and slot offsets; two-level: the
    // coarse words and compact indices, then the fine masks and slot offsets) at
    // dir_mask/dir_off + rdir[i]
    // (rdir < 0: not stored), the numeric kernel loads them instead of rebuilding them.
    const long long *rdir;
    unsigned long long *dir_mask;
    unsigned *dir_off;
};

// Global-memory workspace (bin BIN_GLOBAL): gws_blocks regions of gws_stride u64 each;
// a region holds gws_tiles 64-bit masks followed by gws_tiles 32-bit offsets, with
// gws_tiles >= the widest tile span in the bin.
// queue: nullptr = one workgroup per row group (grid from count); otherwise 8 zeroed
// counters and the LDS kernels run persistent (grid from occupancy x num_cus).
// two_level: some row of the bin has a two-level directory (BinStats::nsparse > 0).
// lds_dyn: LDS a row may use (lds_row_max); BIN_GLOBAL rows keep their flat directory in
// LDS when gws_stride masks fit it (gws is then unused).
hipError_t launch_symbolic(int bin, const int *rows, const int *count_dev, int count,
                           int max_need, const Operands &op, int *rnnz, int tshift, int two_level, int long_rows,
                           unsigned long long *gws, long long gws_stride, int num_cus,
                           int gws_blocks, int lds_dyn, unsigned *queue, hipStream_t s);

hipError_t launch_scan_i32_i64(const int *in, int64_t n, long long *out, void *scan_ws,
                               long long *total_copy, hipStream_t s);

// Wide rows (BIN_GLOBAL) in the numeric phase: blocks workgroups, each with a workspace of
// 3 * cur_stride + ft_stride ints (cursors, ends and next columns of its row's A entries:
// cur_stride >= the longest row of A in the bin; tiles of the fine directory entries of
// a two-level row: ft_stride >= the widest span in the bin).  With kept directories the
// compact windowed kernel runs, and needs to know which rows the symbolic phase held in
// global memory (their directories are flat): sym_row_max (LDS bytes a row may use) and
// sym_all_global (every row was global: force_global).
// Wide rows with kept directories (FP64) are processed as work items: entries [e0, e1) of
// A row `row`.  A row whose entries x windows exceed WS_VISITS is split into several items
// (nsplit > 1); their partial sums meet in C's values by atomic adds (C zeroed first).
struct WideItem {
    int row, e0, e1, nsplit;
    long long sbase; // compensated FP32, split rows: the row's first sum in the scratch
};
constexpr int WIDE_CTRS = 66;
// A window holds at most `vals` value slots and vals / 2 directory entries: WS_VALS (two
// workgroups per CU, the default), or WS_VALS_BIG (one per CU, on request).  Launches
// whose wide rows hold fewer than WS_SPARSE_DENSITY columns of C per tile they span take
// narrower workgroups (the sparse lane counts of k_numeric_wsplit).
constexpr int WS_VALS = 4096;
constexpr int WS_VALS_BIG = 8192;
constexpr int WS_SPARSE_DENSITY = 16;
// A entries of one item, at most: WS_SPLIT with the default window, twice as many with
// the large one (its workgroup has the CU's LDS to itself), so that the fixed cost of a
// window (its directory, its output) is shared by more entries
constexpr int WS_SPLIT = 512;
constexpr int WS_SPLIT_BIG = 1024;
constexpr int WS_MIN_SPLIT = 64;     // A entries of one item, at least (unless the row has fewer)
constexpr long long WS_VISITS = 65536; // target of entries x windows per item (per WS_SPLIT)
HSMU_HD constexpr int wide_split_max(int vals) { return vals == WS_VALS_BIG ? WS_SPLIT_BIG : WS_SPLIT; }
// A entries per item for a row with nA entries and n = nnz(C_i), windows of `vals` slots
HSMU_HD inline int wide_split(long long n, int vals = WS_VALS)
{
    const long long win = n / (vals - 64) + 1; // windows, about
    const long long mx = wide_split_max(vals);
    long long sz = WS_VISITS * (mx / WS_SPLIT) / win;
    sz = sz < WS_MIN_SPLIT ? WS_MIN_SPLIT : (sz > mx ? mx : sz);
    return (int)(sz & ~63LL);
}

struct WideRows {
    int *ws;
    long long cur_stride, ft_stride;
    int blocks;
    long long sym_row_max;
    int sym_all_global;
    int long_rows; // the with_long the symbolic binning used
    int vals;      // split work items: value slots of a window (WS_VALS or WS_VALS_BIG)
    int sparse;    // split work items: the wide rows hold few columns per tile (narrow workgroups)
    // split work items (FP64 with kept directories): capacity, list, WIDE_CTRS counters
    // (ctr[0] = items listed, ctr[1] = items taken, then per size class)
    long long items_cap;
    WideItem *items;
    unsigned *ctr;
    // compensated FP32 split rows: one FP64 sum per nonzero of those rows (zeroed before
    // the items add their windows into it), and the counter that hands out each row's
    // range
    unsigned long long *scratch;
    long long scratch_n;
    unsigned long long *sctr;
};
size_t wide_ws_ints(long long cur_stride, long long ft_stride, int blocks);

// Numeric phase of one bin.  fp32c = 0: FP64 values (Cval double); fp32c = 1: compensated
// FP32 (Cval float).  lds_bytes (may be null): += the dynamic LDS of an LDS-bin launch,
// workgroups x bytes per workgroup (the reference's "practical need",
// compressed_Cval.cuh:202-533); the wide-row kernels' persistent workgroups reuse their LDS
// window by window and are not counted.
hipError_t launch_numeric(int bin, int fp32c, const int *rows, const int *count_dev, int count,
                          int max_need, const Operands &op, const long long *crp, int *Ccol,
                          void *Cval, int bshift, int tshift, int two_level, int long_rows,
                          const WideRows &wide,
                          int *err, int num_cus, unsigned *queue, unsigned long long *lds_bytes,
                          hipStream_t s);

hipError_t launch_rowptr_i64_to_i32(const long long *in, int64_t n, int *out, hipStream_t s);

// out[i] = in[i] - in[0] for i < n (row pointers of a row panel, rebased to 0)
hipError_t launch_rowptr_rebase(const int *in, int64_t n, int *out, hipStream_t s);
// sum[i] and asum[i] = sum of the values / of their magnitudes in row i of a CSR matrix
hipError_t launch_row_sums(int64_t M, const void *rp, int rp64, const void *val, int val32, double *sum,
                           double *asum, hipStream_t s);

// ---- batched form (kernels.hip) ----
// Stacked row pointers of A (which = 0, nrows = total rows of A) or B (which = 1): entry g
// of instance i = e0_i + rp_i[g - r0_i], the last entry = total entries.  Raises
// ERR_ROWPTR when an instance's rp[0] != 0 or rp[rows] != its nnz.
hipError_t launch_batch_rowptr(int which, int64_t count, const BatchInst *inst, BatchOffsets off,
                               int64_t nrows, int64_t nnz, int *out, DevStats *st, hipStream_t s);
// Stacked columns (A: + ko_i; B: unchanged) and values (val may be nullptr: not copied).
// Raises ERR_ACOL / ERR_BCOL for a column outside instance i's range (and writes 0).
hipError_t launch_batch_entries(int which, int64_t count, const BatchInst *inst, BatchOffsets off,
                                int64_t nnz, int val32, int *col, void *val, DevStats *st,
                                hipStream_t s);
// Row pointers of every C_i from the stacked C's (crp, int64): instance i's M_i + 1 entries
// at out + mo_i + i, rebased to start at 0; out is int64 (rp64) or int32.  Also
// bases[i] = crp[mo_i] for i <= count (where C_i's entries start).  nout = rows + count.
hipError_t launch_batch_split(int64_t count, BatchOffsets off, const long long *crp, void *out,
                              int rp64, long long *bases, int64_t nout, hipStream_t s);

// Replayed plans: compare the device's DevStats of one checkpoint (stage 0: after the
// symbolic binning, stage 1: after the numeric binning) with the recorded one; on any
// difference, or an input error already raised, set ERR_PLAN / keep err and zero the
// bin counts of that stage so that no later kernel runs on sizes the plan did not
// provision.
hipError_t launch_plan_check(DevStats *st, const DevStats *expect, int stage, hipStream_t s);
// Replayed batch plans: ERR_PLAN when the start of any C_i moved (its nnz changed).
hipError_t launch_plan_bases(const long long *got, const long long *want, int64_t n, DevStats *st,
                             hipStream_t s);

hipError_t launch_copy(const void *src, void *dst, size_t bytes, int num_cus, hipStream_t s);

// ---- allocation-free entry point (noalloc.hip) ----
// Bytes of the file-scope device array that is hsmu_spgemm_noalloc's whole workspace.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)256 << 20;
// Device address of that array on `device`, which must be the current device: looked up
// once per device (hipGetSymbolAddress), then cached.
hipError_t scratch_setup(int device, void **ptr);
// Held for the whole of an hsmu_spgemm_noalloc call: calls on one device share its array.
std::mutex &scratch_lock(int device);

// Opt every kernel in to the device's full LDS size and tell the row kernels the device's
// XCD count (1..MAX_XCDS); call once per device, with it current.
hipError_t configure_kernels(int lds_max, int num_xcds);

} // namespace hsmu
