/* hsmu_check.h -- the GPU cross-check against rocSPARSE SpGEMM (libhsmu_check.so).
 *
 * The reference checks its result against cuSPARSE in src/cusparse_in_tile.cu
 * (tile_spgemm_cusparse_executor :4-89, tile_cusparse_spgemm :91-272, switched on by
 * CHECK_RESULT in src/test.cu:73-76).  Here the independent implementation is rocSPARSE's
 * generic rocsparse_spgemm on the same device inputs.  This library is test and driver
 * infrastructure: libhsmu.so does not link it.
 *
 * Status codes are hsmu_status values; a rocSPARSE failure is reported as
 * HSMU_ERR_HIP with hsmu_check_last_error() naming the rocSPARSE status.
 */
#ifndef HSMU_CHECK_H
#define HSMU_CHECK_H

#include <stdint.h>

#include "NHC_spgemm.h"
#include "hsmu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* C = A * B by rocSPARSE (int32 indices; FP64, or FP32 when A.val32 and B.val32).
 * Replaces tile_spgemm_cusparse_executor (src/cusparse_in_tile.cu:4-89).  C's arrays
 * are hipMalloc'd here and released by hsmu_check_free.  *ms_stages (optional) receives
 * the device time of rocSPARSE's nnz and compute stages (hipEvents on `stream`), *ms_wall
 * (optional) the host time of the whole call including its buffer and C allocations, as
 * the reference times its cuSPARSE executor (:122-136).  `stream` may be NULL. */
int hsmu_rocsparse_spgemm(void *stream, const hsmu_csr *A, const hsmu_csr *B, hsmu_csr *C, double *ms_stages,
                          double *ms_wall);

/* Release a C returned by hsmu_rocsparse_spgemm (its arrays are zeroed). */
int hsmu_check_free(hsmu_csr *C);

/* Result of comparing two device CSR matrices of the same shape. */
typedef struct hsmu_csr_diff {
    int64_t rowptr_mismatch; /* rows whose row pointer differs */
    int64_t col_mismatch;    /* entries whose column differs (compared only if row pointers agree) */
    int64_t val_mismatch;    /* entries with |x - y| > tol * scale */
    int64_t first_row;       /* first differing row pointer, or -1 */
    int64_t first_entry;     /* first differing column or value, or -1 */
    double max_abs_err;      /* max |x - y| over entries */
    double max_rel_err;      /* max |x - y| / scale over entries with scale > 0 */
} hsmu_csr_diff;

/* Compare X and Y (device CSR, any index or value width): row pointers and columns
 * exactly, values within |x - y| <= tol * scale, where scale is S's value at the same
 * entry when S (same structure, e.g. |A|*|B|) is given, otherwise max(|x|, |y|).
 * Returns HSMU_OK when the comparison ran (whether or not it found differences). */
int hsmu_csr_compare(const hsmu_csr *X, const hsmu_csr *Y, const hsmu_csr *S, double tol, hsmu_csr_diff *diff);

/* The reference's check (tile_cusparse_spgemm, src/cusparse_in_tile.cu:91-272) on the
 * NHC structs: runs rocSPARSE on A and B, prints its nnz, time and GFLOPS, then compares
 * row pointers (against C->rowPtr) and columns and values (against C's device arrays,
 * values within 1e-12 * (|A|*|B|) computed by rocSPARSE), printing the reference's lines
 * with the mismatch counts.  test_C is unused, as in the reference.  Returns 0 when
 * everything agrees, 1 on a mismatch, 2 when the check could not run. */
int tile_rocsparse_spgemm(NHC_CSR *a, NHC_CSR *b, NHC_CSR *test_C, NHC_CSR *c);

const char *hsmu_check_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* HSMU_CHECK_H */
