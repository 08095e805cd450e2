/*
 * hsmu.h -- C ABI of the MI355X-native HSMU SpGEMM library (libhsmu.so).
 *
 * C = A * B with A, B and C in CSR, on one AMD Instinct MI355X (gfx950).
 * Plain pointers and sizes only: no HIP, torch or CUDA types cross this boundary
 * (streams are passed as void*).
 *
 * Reference interfaces each entry point replaces (wuminqaq/HSMU-SpGEMM):
 *   hsmu_spgemm          <- NHC_spgemm            src/NHC_spgemm.cuh:7-43
 *                           (compressed_NHC_spgemm src/compressed/compressed_NHC_spgemm.cu:22-248,
 *                            small_scale_NHC_spgemm src/small/small_scale_NHC_spgemm.cu:4-121)
 *   hsmu_btiles          <- Initialize_compressed_mask64_matrix
 *                                                 src/compressed/compressed_mask64.cuh:343-424
 *   hsmu_csr_free        <- free_C_matrix         src/loading.cuh:152-166
 *   hsmu_tiles_free      <- realse_NHC_tile       src/loading.cuh:95-104
 *   hsmu_copy_h2d        <- HtD_value_col_ptr     src/loading.cuh:86-94 (one array at a time)
 *   hsmu_read_mtx        <- mmio_allinone         src/mmio_highlevel.h:18-193
 *   hsmu_transpose       <- csr2csc               src/mmio_highlevel.h:194-229
 *   hsmu_ctx_create/destroy <- setup / realse_compressed_bin / free_bin
 *                                                 src/loading.cuh:51-84,105-127,167-174
 * The reference-compatible entry point NHC_spgemm itself is declared in NHC_spgemm.h.
 *
 * Semantics kept from the reference:
 *   - rows of B must be sorted by column (the reference's numeric binary search,
 *     src/compressed/compressed_Cval.cuh:55-61, needs it too); an unsorted row is an
 *     error here (HSMU_ERR_UNSORTED) instead of an out-of-bounds access;
 *   - columns of every row of C are sorted ascending;
 *   - structural nonzeros are kept even when their value cancels to zero;
 *   - repeated column indices in a row of B are summed.
 * Extensions: int64 row pointers for C (nnz(C) >= 2^31), a compensated-FP32 numeric
 * form, and a batched entry point.
 */
#ifndef HSMU_H
#define HSMU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------- */
enum hsmu_status {
    HSMU_OK = 0,
    HSMU_ERR_INVALID = 1,     /* bad argument: null pointer, negative size, A.N != B.M, ... */
    HSMU_ERR_UNSORTED = 2,    /* a row of B has decreasing column indices */
    HSMU_ERR_INDEX = 3,       /* a column index or row pointer out of range */
    HSMU_ERR_OVERFLOW = 4,    /* nnz(C) does not fit the requested row-pointer type */
    HSMU_ERR_NOMEM = 5,       /* device or host allocation failed */
    HSMU_ERR_HIP = 6,         /* HIP runtime error (message in hsmu_last_error) */
    HSMU_ERR_UNSUPPORTED = 7, /* valid request this build does not implement */
    HSMU_ERR_PLAN = 8         /* a plan's inputs no longer have the structure it was made for */
};

/* numeric form */
enum hsmu_precision {
    HSMU_FP64 = 0,      /* plain: FP64 products, FP64 LDS accumulation (the reference's value_t) */
    HSMU_FP32_COMP = 1  /* FP32 products, compensated (two-sum) FP32 accumulation, FP32 values */
};

/* ---- matrices ---------------------------------------------------------------- */

/* A CSR matrix in device memory.
 * Inputs: rowptr is int32 (rowptr64 = 0) and val is double, or float when val32 = 1
 * (FP32_COMP inputs).  Output C: rowptr is int32 or int64 (hsmu_opts.rowptr64), col int32,
 * val double (FP64) or float (FP32_COMP). */
typedef struct hsmu_csr {
    int64_t M;        /* rows */
    int64_t N;        /* columns */
    int64_t nnz;      /* stored entries */
    void *rowptr;     /* device [M+1], int32 or int64 */
    int32_t *col;     /* device [nnz] */
    void *val;        /* device [nnz], double or float */
    int32_t rowptr64; /* 1: rowptr is int64 */
    int32_t val32;    /* 1: val is float */
} hsmu_csr;

/* B in the compressed 64-column tile format of src/compressed/compressed_mask64.cuh:
 * for row j, tiles tptr[j] .. tptr[j+1]-1; tile t covers columns 64*tcol[t] ..
 * 64*tcol[t]+63 and bit b of tmask[t] is set when column 64*tcol[t]+b is stored.
 * Unlike the reference (whose tile order comes from LDS atomics), the tiles of a row
 * are sorted by tcol. */
typedef struct hsmu_tiles {
    int64_t M;        /* rows of B */
    int64_t ntiles;   /* total tiles */
    int32_t *tptr;    /* device [M+1] */
    int32_t *tcol;    /* device [ntiles] */
    uint64_t *tmask;  /* device [ntiles] */
} hsmu_tiles;

/* A CSR matrix in host memory (generators, Matrix Market reader).  Arrays are
 * malloc'd by the library and released with hsmu_host_csr_free. */
typedef struct hsmu_host_csr {
    int64_t M, N, nnz;
    int32_t *rowptr;  /* [M+1] */
    int32_t *col;     /* [nnz] */
    double *val;      /* [nnz] */
} hsmu_host_csr;

/* ---- options and statistics ---------------------------------------------------- */

typedef struct hsmu_opts {
    int32_t precision;        /* enum hsmu_precision */
    int32_t rowptr64;         /* 1: C row pointers are int64 */
    int32_t plain_alloc;      /* 1: allocate C with hipMalloc (release with hipFree, as the
                                 reference's free_C_matrix does); 0: from the context pool */
    int32_t force_global;     /* 1: every nonempty row takes the global-memory accumulator
                                 (exercises the fallback on small inputs) */
    int32_t lds_row_max;      /* bytes of LDS one row may use before it falls back to the
                                 global-memory path; 0 = device maximum */
    int32_t lds_block_budget; /* LDS bytes per workgroup for bins that pack several rows
                                 into one workgroup (one wave); 0 = default (8 KiB) */
    int32_t products_per_lane;/* numeric work per lane that sets a row's lane count; 0 = 32 */
    int32_t tiles_per_lane;   /* symbolic work per lane that sets a row's lane count; 0 = 4 */
    int32_t timing;           /* hipEvents into hsmu_stats (each costs the device a few us):
                                 1 = per stage and per numeric bin (the default), 2 = per
                                 numeric bin only, 0 = none (stage and bin times read 0) */
    int32_t b_lanes;          /* lanes that walk one row of B in the numeric phase (power of
                                 two, 1..64); 0 = a quarter of the mean row length of B */
    int32_t t_lanes;          /* lanes that walk one row of B's tiles in the tile OR (power
                                 of two, 1..64); 0 = from the mean tiles per row of B */
    int32_t rebuild_dirs;     /* 1: the numeric phase rebuilds each row's tile directory
                                 instead of reusing the one the symbolic phase kept */
    int32_t persistent;       /* 1: the LDS row kernels run as persistent workgroups (as
                                 many as fit the device) pulling row groups from per-XCD
                                 device work queues; 0: one workgroup per row group */
    int32_t min_waves;        /* a row that has a workgroup to itself gets enough lanes for
                                 this many waves per CU at the occupancy its LDS allows (at
                                 most 1024 lanes); 0 = 8 */
    int32_t wide_scratch_mb;  /* compensated FP32: MiB the FP64 scratch of wide rows split
                                 into several work items may take; 0 = a quarter of device
                                 memory.  Beyond it those rows take one workgroup each */
    int32_t wide_window;      /* rows wider than LDS, split into work items: value slots of
                                 one window of the windowed accumulator, 4096 (two
                                 workgroups per CU) or 8192 (one); 0 = 4096.  Workgroups
                                 are narrower when those rows hold fewer than 16 columns of
                                 C per 64-column tile they span */
} hsmu_opts;

#define HSMU_NBINS 10 /* 8 LDS bins (8..1024 lanes per row), global fallback, wide rows */

typedef struct hsmu_stats {
    /* stage times in ms (hipEvents on the library stream), named after the
     * reference's stages (src/compressed/compressed_NHC_spgemm.cu:36-170) */
    double ms_conversion;     /* B -> tiles: Initialize_compressed_mask64_matrix */
    double ms_form_tileCptr;  /* row analysis and binning: compressed_Form_tileCptr */
    double ms_tileOR;         /* symbolic tile OR, nnz per row, C row pointers */
    double ms_malloc;         /* allocation of C */
    double ms_get_Ccol;       /* column extraction; fused into the numeric kernels (0) */
    double ms_form_cval;      /* numeric: accumulate, write columns and values */
    double ms_time2;          /* the reference's time2: all of the above but conversion */
    double ms_total;          /* conversion + time2 */
    double ms_numeric_bin[HSMU_NBINS]; /* per numeric bin kernel time (timing 1 or 2) */
    int64_t products;         /* intermediate products, sum over a_ij of nnz(B_j) */
    int64_t tile_products;    /* sum over a_ij of tiles(B_j) */
    int64_t nnz_c;
    int64_t tiles_b;
    int64_t rows_nonempty;
    int64_t sym_bin_rows[HSMU_NBINS];
    int64_t num_bin_rows[HSMU_NBINS];
    double compress_rate;     /* products / nnz(C) (floating, the reference divides integers) */
    double gflops;            /* 2 * products / time2 */
    double space_compressed_mask_mb; /* size of B's tile format, MiB */
    double alg_bytes;         /* algorithmic HBM bytes of the whole SpGEMM (gather model) */
    double num_bin_bytes[HSMU_NBINS];    /* the same model over each numeric bin's rows */
    int64_t num_bin_products[HSMU_NBINS];
    double compulsory_bytes;  /* every input array read once, C written once: the lower bound
                                 of HBM traffic */
    double lds_launched;      /* dynamic LDS of the numeric LDS-bin launches, workgroups x
                                 bytes each (the reference's practical_need,
                                 compressed_Cval.cuh:202-533) */
    double lds_use;           /* 12 * nnz(C) of the LDS-bin rows / lds_launched: the
                                 reference's use_of_shared (compressed_Cval.cuh:564-570) */
} hsmu_stats;

/* ---- context ------------------------------------------------------------------- */

typedef struct hsmu_ctx hsmu_ctx;

/* Create a context on a device: one HIP stream (default priority), a caching device
 * allocator, pinned staging memory.  One context per host thread. */
int hsmu_ctx_create(int device, hsmu_ctx **out);
/* Destroy a context.  Every device block its pool holds is freed, including results
 * (hsmu_csr, batches) that were not released: release results and destroy plans BEFORE
 * their context; releasing them afterwards is invalid. */
void hsmu_ctx_destroy(hsmu_ctx *ctx);
void *hsmu_ctx_stream(hsmu_ctx *ctx);       /* the hipStream_t, as void* */
int hsmu_ctx_trim(hsmu_ctx *ctx);           /* release cached device memory */
int64_t hsmu_ctx_pool_bytes(hsmu_ctx *ctx); /* device bytes held by the pool */
int hsmu_ctx_lds_max(hsmu_ctx *ctx);        /* LDS bytes per workgroup on this device */
int hsmu_ctx_num_cus(hsmu_ctx *ctx);

void hsmu_opts_default(hsmu_opts *opts);
const char *hsmu_status_string(int status);
/* "src_commit=<commit>[+dirty] src=<hash>": the last commit that changed the library's
 * sources and the hash of the sources it was built from (no counterpart in the
 * reference). */
const char *hsmu_build_info(void);
const char *hsmu_last_error(void);          /* message of the last failure on this thread */

/* ---- the operator ---------------------------------------------------------------- */

/* C = A * B.  A and B in device memory (int32 row pointers); the library allocates C's
 * arrays (from the pool unless opts->plain_alloc) and fills every field of *C.
 * Synchronous on return.  opts and stats may be NULL. */
int hsmu_spgemm(hsmu_ctx *ctx, const hsmu_csr *A, const hsmu_csr *B, hsmu_csr *C,
                const hsmu_opts *opts, hsmu_stats *stats);

/* hsmu_spgemm that allocates nothing: no hipMalloc, hipFree or other allocating call.
 * No counterpart in the reference (its NHC_spgemm cudaMallocs C and its workspaces
 * inside the call, src/compressed/compressed_NHC_spgemm.cu:110-113); for callers that
 * cannot allocate device memory: code under stream capture, serving loops, harnesses
 * that forbid it.
 * The caller passes C's arrays: C->rowptr (device, M + 1 entries: int32, or int64 with
 * opts->rowptr64), C->col (int32) and C->val (double, or float for HSMU_FP32_COMP), each
 * with room for cap_nnz entries; on success M, N, nnz, rowptr64 and val32 of *C are set.
 * The workspace is one file-scope device array of HSMU_NOALLOC_SCRATCH_BYTES that every
 * GPU the library runs on holds until the process ends; a product whose workspace needs
 * more fails with HSMU_ERR_NOMEM and a message naming the limit (the banded 62,451-row
 * benchmark product needs about 110 MiB).  If nnz(C) exceeds cap_nnz the call fails with
 * HSMU_ERR_NOMEM and sets C->nnz to the nnz(C) needed.  Every call on one GPU uses that
 * GPU's one array, which keeps the last call's contents: calls that overlapped on a GPU
 * would race, so the library serialises the calls of a process per device (a call waits
 * for the one before it, whichever thread or context made it).  Synchronous on return,
 * like hsmu_spgemm; opts->plain_alloc is ignored. */
#define HSMU_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
int hsmu_spgemm_noalloc(hsmu_ctx *ctx, const hsmu_csr *A, const hsmu_csr *B, hsmu_csr *C, int64_t cap_nnz,
                        const hsmu_opts *opts, hsmu_stats *stats);

/* Release C's device arrays (pool or hipFree, as allocated) and zero the struct. */
int hsmu_csr_free(hsmu_ctx *ctx, hsmu_csr *C);

/* C = A[r0:r1, :] * B: the product for a panel of A's rows (C has r1 - r0 rows), for
 * products too large for device memory, streamed a panel at a time (no counterpart in
 * the reference, which holds the whole C on one GPU: src/NHC_spgemm.cuh:7-43). */
int hsmu_spgemm_rows(hsmu_ctx *ctx, const hsmu_csr *A, int64_t r0, int64_t r1, const hsmu_csr *B,
                     hsmu_csr *C, const hsmu_opts *opts, hsmu_stats *stats);

/* Host arrays sums[i], abs_sums[i] (C.M each) = sum of the values, of their magnitudes,
 * in row i of a device CSR matrix: a check of C that needs no download of C
 * (row sums of A*B equal A times the row sums of B). */
int hsmu_csr_row_sums(hsmu_ctx *ctx, const hsmu_csr *C, double *sums, double *abs_sums);

/* B -> compressed 64-column tiles (the conversion stage on its own). */
int hsmu_btiles(hsmu_ctx *ctx, const hsmu_csr *B, hsmu_tiles *T);
int hsmu_tiles_free(hsmu_ctx *ctx, hsmu_tiles *T);

/* ---- batched form (no counterpart in the reference, which multiplies one pair per
 *      process: src/test.cu:10-94) ----------------------------------------------------- */

/* C[i] = A[i] * B[i] for i < count, as ONE pipeline over all instances: the instances
 * are stacked (rows of every A[i] one after another, columns of A[i] shifted onto the
 * rows of B[i] in a stacked B), so every stage is one launch over the rows of all
 * instances and the bins mix rows of different instances.  Two host synchronisations
 * for the whole batch.  Instances may differ in every size.
 * Output: C[i] are views into three device arrays owned by the batch; C[0]'s pointers
 * are their bases.  Row pointers of every C[i] start at 0.  Release with
 * hsmu_batch_free (never hsmu_csr_free on a C[i]).  stats (may be NULL) covers the
 * whole batch.  With int32 row pointers the whole batch's nnz must fit int32. */
int hsmu_spgemm_batched(hsmu_ctx *ctx, int64_t count, const hsmu_csr *A, const hsmu_csr *B,
                        hsmu_csr *C, const hsmu_opts *opts, hsmu_stats *stats);
int hsmu_batch_free(hsmu_ctx *ctx, int64_t count, hsmu_csr *C);

/* The baseline the batched form is measured against: `passes` times, C = A[i] * B[i] for
 * every i by one hsmu_spgemm call each (each result released right away), timed on the
 * host around the whole loop; *ms_per_pass = the time of one pass over the count
 * instances.  No counterpart in the reference (one product per process, src/test.cu). */
int hsmu_spgemm_each(hsmu_ctx *ctx, int64_t count, const hsmu_csr *A, const hsmu_csr *B,
                     const hsmu_opts *opts, int32_t passes, double *ms_per_pass);

/* ---- plans: inspect once, replay as one HIP graph ----------------------------------
 * hsmu_plan_create runs the batch (count >= 1; count == 1 is a single SpGEMM) once,
 * records every launch size, LDS size and buffer it used, and captures the whole
 * sequence (input stacking, B -> tiles, row analysis, bins, symbolic, scans, numeric,
 * output row pointers) into a HIP graph that contains no host synchronisation, no
 * allocation and no memset.  C[i] are filled by the create call and stay owned by the
 * plan.  hsmu_plan_execute replays the graph: C is recomputed from the CURRENT contents
 * of the bound input arrays, whose values may change but whose structure (row pointers
 * and columns) must not; a changed structure is detected on the device and reported as
 * HSMU_ERR_PLAN.  stats (may be NULL) gets ms_total = the replay's device time. */
typedef struct hsmu_plan hsmu_plan;
int hsmu_plan_create(hsmu_ctx *ctx, int64_t count, const hsmu_csr *A, const hsmu_csr *B,
                     hsmu_csr *C, const hsmu_opts *opts, hsmu_plan **plan);
int hsmu_plan_execute(hsmu_plan *plan, hsmu_stats *stats);
void hsmu_plan_destroy(hsmu_plan *plan); /* releases the plan's buffers, C[i] included */

/* ---- autotuner ------------------------------------------------------------------------
 * Bin thresholds (products_per_lane, tiles_per_lane), sub-group widths (b_lanes,
 * t_lanes), the LDS budget per workgroup, the persistent mode and whether the numeric
 * phase rebuilds the tile directories (rebuild_dirs), chosen per SHAPE CLASS
 * of the input: the class is "c<C>_a<Ea>_b<Eb>_m<Em>" from the instance count bucket C
 * (1, or log2 of the count), the rounded log2 of the mean row length of A and of B, and
 * the floor of log2 of the total rows of A (hsmu_shape_class).  The table is a text
 * file, one line per class: class, the seven values, the measured ms.  When the class is
 * in the table its values are returned; otherwise, if search is nonzero, the candidates
 * are timed on this input (coordinate descent from *base, each candidate run and timed
 * three times, best median kept), the winner is appended to the table (if table_path is
 * not NULL) and returned; if search is zero, *base is returned.  *found = 1 when the class came from the table. */
int hsmu_shape_class(int64_t count, const hsmu_csr *A, const hsmu_csr *B, char *cls, size_t cls_len);
int hsmu_autotune(hsmu_ctx *ctx, int64_t count, const hsmu_csr *A, const hsmu_csr *B,
                  const hsmu_opts *base, const char *table_path, int search, hsmu_opts *out,
                  int32_t *found);

/* ---- device memory helpers (for bindings that carry no device allocator) -------- */
int hsmu_device_alloc(hsmu_ctx *ctx, size_t bytes, void **ptr);
int hsmu_device_free(hsmu_ctx *ctx, void *ptr);
int hsmu_copy_h2d(hsmu_ctx *ctx, void *dst, const void *src, size_t bytes);
int hsmu_copy_d2h(hsmu_ctx *ctx, void *dst, const void *src, size_t bytes);
int hsmu_sync(hsmu_ctx *ctx);

/* Bandwidth probe: copy n bytes device to device with a plain vector copy kernel and
 * return its time in ms (used to scale HBM counters and as the achievable-bandwidth
 * figure beside the roofline). */
int hsmu_copy_kernel(hsmu_ctx *ctx, void *dst, const void *src, size_t bytes, double *ms);

/* ---- host-side inputs ------------------------------------------------------------ */

/* 2D 5-point Laplacian on an n x n grid: row i = y*n + x, diagonal 4, neighbours -1. */
int hsmu_gen_laplace2d(int32_t n, hsmu_host_csr *out);
/* 3D 7-point Laplacian on an n^3 grid: diagonal 6, neighbours -1. */
int hsmu_gen_laplace3d(int32_t n, hsmu_host_csr *out);
/* Banded FEM-like: M x M, per-row nnz uniform in [lo, hi], distinct columns uniform in
 * [i - halfwidth, i + halfwidth] clipped to [0, M), values N(0,1). */
int hsmu_gen_banded(int32_t M, int32_t lo, int32_t hi, int32_t halfwidth, uint64_t seed,
                    hsmu_host_csr *out);
/* R-MAT: 2^scale x 2^scale, edge_factor * 2^scale generated edges with quadrant
 * probabilities a, b, c (d = 1 - a - b - c), duplicates merged, values U(0,1). */
int hsmu_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                  uint64_t seed, hsmu_host_csr *out);
/* Zipf: M x N, per_row distinct columns per row drawn Zipf(s) over the columns,
 * values U(-1,1). */
int hsmu_gen_zipf(int32_t M, int32_t N, int32_t per_row, double s, uint64_t seed,
                  hsmu_host_csr *out);
/* Random sparse M x N with per-row nnz uniform in [0, max_per_row], sorted distinct
 * columns, values U(-1,1) (test inputs). */
int hsmu_gen_random(int32_t M, int32_t N, int32_t max_per_row, uint64_t seed,
                    hsmu_host_csr *out);
/* Transpose (CSR -> CSR of A^T, i.e. csr2csc); rows of the result are sorted. */
int hsmu_transpose(const hsmu_host_csr *A, hsmu_host_csr *AT);
/* Matrix Market coordinate file -> CSR with the reference loader's semantics
 * (symmetric expansion, pattern values 1.0, duplicates kept); rows are then sorted
 * by column (stable), which the reference leaves to the file. */
int hsmu_read_mtx(const char *path, hsmu_host_csr *out, int32_t *is_symmetric);
void hsmu_host_csr_free(hsmu_host_csr *m);

#ifdef __cplusplus
}
#endif

#endif /* HSMU_H */
