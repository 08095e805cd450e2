/*
 * NHC_spgemm.h -- drop-in for the reference's operator boundary.
 *
 * Same type names, field order and function names as the reference's
 * inc/NHC_spgemm.h:66-139 and src/loading.cuh, so a caller written against
 * wuminqaq/HSMU-SpGEMM (its driver src/test.cu:10-94 is the only one) compiles against
 * this header and links libhsmu.so instead of the CUDA sources.
 *
 * Differences, all forced by a C ABI without CUDA types:
 *   - every function is extern "C";
 *   - compressed_bin::streams is `void **` (hipStream_t is a pointer) instead of
 *     `cudaStream_t *`;
 *   - the C++ default member initialisers of NHC_mask_matrix::size_arr and
 *     compressed_bin::temp_storage_bytes are dropped (the fields keep their place);
 *   - host arrays owned by these structs are malloc'd/free'd (the reference mixes
 *     malloc in mmio_allinone with delete[] in realse_NHCcsr).
 * The value type is double and the index type int, as in the reference
 * (inc/NHC_spgemm.h:12-13); C's row pointers are therefore int32 and NHC_spgemm fails
 * loudly when nnz(C) >= 2^31 (use hsmu_spgemm with int64 row pointers instead).
 */
#ifndef HSMU_NHC_SPGEMM_H
#define HSMU_NHC_SPGEMM_H

#include <stddef.h>
#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define DateTypeStoreMask unsigned long long
#define DateTypeStoreCompressMask unsigned long long
#define value_t double
#define index_t int
/* B.N below this takes the reference's small (dense bitmap) path; here it only selects
 * which set of stage lines NHC_spgemm prints (inc/NHC_spgemm.h:40). */
#define threshold_value_selected 4096

/* inc/NHC_spgemm.h:66-86 */
typedef struct
{
    index_t *rowPtr;
    index_t *col;
    value_t *val;
    index_t *d_ptr;
    index_t *d_col;
    value_t *d_val;
    index_t *d_tem_ptr;
    int M;
    int N;
    index_t nnz;
    int nnz_max;
    int isSymmetric;
    char *matrix_name;
    index_t *d_tile_ptr;
    index_t *d_tile_col;
    index_t *d_tem_tile_ptr;
    index_t *tile_ptr;
    DateTypeStoreCompressMask *d_mask_num;
} NHC_CSR;

/* inc/NHC_spgemm.h:88-100 */
typedef struct
{
    DateTypeStoreMask *mask_matrix;
    DateTypeStoreMask *after_or_mask_matrix;
    int shift_of_pitch;
    int pitch;
    int *d_nums_over_limit_hashtable;
    int nums_over_limit_hashtable;
    int *d_array_store_Crow_id_over_limit;
    int num_unit;
    int *d_Brow_id;
    int size_arr[5];
} NHC_mask_matrix;

/* inc/NHC_spgemm.h:102-109 */
typedef struct
{
    int *bin;
    int *d_bin;
    int max_rows_partion;
    int num_of_partion;
    int *d_max_nnz_in_onerow;
} NHC_bin;

/* inc/NHC_spgemm.h:111-130 */
typedef struct
{
    int num_of_partion;
    int *bin_size;
    int *bin_offset;
    int *max_row_nnz;
    int *total_nnz;
    int *bins;
    int *d_bins;
    int *d_bin_size;
    int *d_bin_offset;
    int *d_total_nnz;
    int *d_max_row_nnz;
    void **streams;
    int *d_global_mem_pool;
    size_t global_mem_pool_size;
    bool global_mem_pool_malloced;
    void *d_temp_storage;
    size_t temp_storage_bytes;
} compressed_bin;

/* inc/NHC_spgemm.h:132-139 */
typedef struct
{
    int *d_tile_ptr;
    int *d_tile_col;
    int *d_tem_tile_ptr;
    int *tile_ptr;
    DateTypeStoreCompressMask *d_mask_num;
} NHC_tile;

/* C = A * B (src/NHC_spgemm.cuh:7-43).
 * Preconditions as in the reference: A->d_ptr/d_col/d_val and B->d_ptr/d_col/d_val hold
 * the device CSR (HtD_value_col_ptr), rows of B sorted by column, C->M/C->N set by the
 * caller (src/test.cu:65-66).
 * Postconditions: C->d_ptr (int[M+1]), C->d_col, C->d_val on the device, sorted columns
 * in every row; C->rowPtr (host copy of the row pointers) and C->nnz.  B's tile format
 * is workspace of the call and is not left in B (the reference leaves it in
 * B->d_tile_ptr/d_tile_col/d_mask_num); those fields, C's tile fields and every pointer
 * of compressed_bin, c_bin and mask_matrixB are set to NULL so that the release
 * functions below stay valid on them.
 * Prints the reference's stage lines (src/compressed/compressed_NHC_spgemm.cu:55-205,
 * src/NHC_spgemm.cuh:36) with this build's stage times.
 * On failure it prints the reason to stderr, leaves C empty (nnz 0, NULL device arrays)
 * and NHC_last_status() returns the hsmu_status code; the reference has no error path. */
void NHC_spgemm(compressed_bin *compressed_bin, NHC_bin *c_bin, NHC_CSR *A, NHC_CSR *B, NHC_CSR *C,
                NHC_mask_matrix *mask_matrixB);

/* hsmu_status of the last NHC_spgemm call on this thread (0 = success). */
int NHC_last_status(void);

/* Lifecycle functions of src/loading.cuh, same meaning. */
void HtD_value_col_ptr(NHC_CSR *A);              /* :86-94  upload rowPtr/col/val */
void realse_NHC_tile(NHC_CSR *tile);             /* :95-104 free the tile arrays */
void realse_compressed_bin(compressed_bin *bin); /* :105-127 */
void realse_NHCcsr(NHC_CSR *A);                  /* :128-140 free device and host CSR */
void free_mask_matrix(NHC_mask_matrix *mask);    /* :141-151 */
void free_C_matrix(NHC_CSR *C);                  /* :152-166 free C's device arrays, host rowPtr */
void free_bin(NHC_bin *c_bin);                   /* :167-174 */

#ifdef __cplusplus
}
#endif

#endif /* HSMU_NHC_SPGEMM_H */
