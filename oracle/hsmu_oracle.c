/*
 * hsmu_oracle.c -- CPU oracle for HSMU SpGEMM (C = A*B, CSR in, CSR out).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU path is compared
 * against, and the CPU baseline that bench.py times beside the GPU.  Nothing in
 * the product library links, loads or calls it; only tests/, the smoke check in
 * __graft_entry__.py and bench.py's cpu_baseline leg may.
 *
 * What it restates from the reference (wuminqaq/HSMU-SpGEMM):
 *   - products:    P = sum over stored a_ij of nnz(B row j)
 *                  (src/compressed/compressed_NHC_spgemm.cu:191-196)
 *   - tile format: tile = col >> 6, mask bit = col & 63, one (tile, mask) pair per
 *                  distinct tile of a row (src/compressed/compressed_mask64.cuh:61-342);
 *                  the reference's tile order is unordered, the oracle's is ascending.
 *   - C tiles:     OR of B's tile masks over the A row
 *                  (src/compressed/compressed_tileOR.cuh:20-104); nnz(C_i) = sum popcount.
 *   - C structure: columns sorted ascending per row
 *                  (src/compressed/compressed_Ccol_symbolic.cuh:38-62), structural
 *                  nonzeros kept even if the value cancels to zero, duplicate B entries
 *                  summed (src/compressed/compressed_Cval.cuh:51-63 adds every product
 *                  into the slot found by src/small/numeric.cuh:233-253).
 *   - values:      c_ik = sum_j a_ij * b_jk in FP64 (inc/NHC_spgemm.h:12, value_t double).
 *
 * The numeric pass is a row-wise Gustavson (dense accumulator + marker per thread,
 * OpenMP dynamic schedule over rows).  Besides c_ik it returns sum_j |a_ij*b_jk|
 * per entry, the scale of the FP64 tolerance |c - c_ref| <= tol * sum|a*b|.
 *
 * All functions work on a row range [r0, r1) of A so that large products can be
 * checked or timed in chunks; output offsets are relative to row r0.
 * Row pointers of A and B are int32 (the reference's index_t); C offsets are int64.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <omp.h>

#define ORACLE_OK 0
#define ORACLE_ERR_ALLOC 1
#define ORACLE_ERR_RANGE 2

static void set_threads(int nthreads)
{
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
}

int oracle_max_threads(void) { return omp_get_max_threads(); }

/* Per-row intermediate products and their total
 * (src/compressed/compressed_NHC_spgemm.cu:191-196). row_products may be NULL. */
int64_t oracle_products(int64_t r0, int64_t r1, const int32_t *Arp, const int32_t *Acol,
                        const int32_t *Brp, int64_t *row_products)
{
    int64_t total = 0;
    for (int64_t i = r0; i < r1; ++i) {
        int64_t p = 0;
        for (int32_t e = Arp[i]; e < Arp[i + 1]; ++e) {
            int32_t j = Acol[e];
            p += (int64_t)(Brp[j + 1] - Brp[j]);
        }
        if (row_products)
            row_products[i - r0] = p;
        total += p;
    }
    return total;
}

static void sort_i32(int32_t *a, int64_t n)
{
    /* iterative quicksort with insertion sort for short ranges */
    int64_t stack_lo[64], stack_hi[64];
    int sp = 0;
    stack_lo[sp] = 0;
    stack_hi[sp] = n - 1;
    ++sp;
    while (sp > 0) {
        --sp;
        int64_t lo = stack_lo[sp], hi = stack_hi[sp];
        while (hi - lo > 24) {
            int64_t mid = lo + (hi - lo) / 2;
            int32_t x = a[lo], y = a[mid], z = a[hi];
            int32_t piv = (x < y) ? ((y < z) ? y : ((x < z) ? z : x)) : ((x < z) ? x : ((y < z) ? z : y));
            int64_t i = lo, j = hi;
            while (i <= j) {
                while (a[i] < piv)
                    ++i;
                while (a[j] > piv)
                    --j;
                if (i <= j) {
                    int32_t t = a[i];
                    a[i] = a[j];
                    a[j] = t;
                    ++i;
                    --j;
                }
            }
            /* push the larger part, loop on the smaller one */
            if (j - lo > hi - i) {
                stack_lo[sp] = lo;
                stack_hi[sp] = j;
                ++sp;
                lo = i;
            } else {
                stack_lo[sp] = i;
                stack_hi[sp] = hi;
                ++sp;
                hi = j;
            }
        }
        for (int64_t p = lo + 1; p <= hi; ++p) {
            int32_t v = a[p];
            int64_t q = p - 1;
            while (q >= lo && a[q] > v) {
                a[q + 1] = a[q];
                --q;
            }
            a[q + 1] = v;
        }
    }
}

/* Sorted copy of the columns of row j (rows need not be sorted for the oracle). */
static int32_t *sorted_row(const int32_t *Brp, const int32_t *Bcol, int64_t j, int32_t *n_out)
{
    int32_t n = Brp[j + 1] - Brp[j];
    *n_out = n;
    if (n == 0)
        return NULL;
    int32_t *tmp = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    if (!tmp)
        return NULL;
    memcpy(tmp, Bcol + Brp[j], sizeof(int32_t) * (size_t)n);
    sort_i32(tmp, n);
    return tmp;
}

/* Tile count per row of B: number of distinct (col >> 6) values
 * (src/compressed/compressed_mask64.cuh:61-152 counts them with an LDS bitmap).
 * Returns the total, or -1 on allocation failure. */
int64_t oracle_btile_counts(int64_t K, const int32_t *Brp, const int32_t *Bcol, int32_t *tcount)
{
    int64_t total = 0;
    for (int64_t j = 0; j < K; ++j) {
        int32_t n;
        int32_t *s = sorted_row(Brp, Bcol, j, &n);
        if (n > 0 && !s)
            return -1;
        int32_t c = 0;
        for (int32_t e = 0; e < n; ++e)
            if (e == 0 || (s[e] >> 6) != (s[e - 1] >> 6))
                ++c;
        free(s);
        tcount[j] = c;
        total += c;
    }
    return total;
}

/* Fill tiles of B in ascending tile order: tcol = col >> 6, tmask |= 1 << (col & 63)
 * (src/compressed/compressed_mask64.cuh:153-303).  tptr must hold the exclusive
 * scan of oracle_btile_counts.  Returns 0, or -1 on allocation failure. */
int oracle_btile_fill(int64_t K, const int32_t *Brp, const int32_t *Bcol, const int64_t *tptr,
                      int32_t *tcol, uint64_t *tmask)
{
    for (int64_t j = 0; j < K; ++j) {
        int32_t n;
        int32_t *s = sorted_row(Brp, Bcol, j, &n);
        if (n > 0 && !s)
            return -1;
        int64_t t = tptr[j] - 1;
        for (int32_t e = 0; e < n; ++e) {
            if (e == 0 || (s[e] >> 6) != (s[e - 1] >> 6)) {
                ++t;
                tcol[t] = s[e] >> 6;
                tmask[t] = 0;
            }
            tmask[t] |= 1ull << (s[e] & 63);
        }
        free(s);
    }
    return 0;
}

/* Per row of C: number of tiles and nnz, by OR-ing B's tile masks over the A row
 * (src/compressed/compressed_tilePtr_symbolic.cuh:175-231 counts the tiles,
 *  src/compressed/compressed_tileOR.cuh:20-104 ORs the masks and sums popcounts).
 * Uses a dense mask array over the tile range of C's columns. */
int oracle_ctile_rows(int64_t r0, int64_t r1, int64_t N, const int32_t *Arp, const int32_t *Acol,
                      const int64_t *tptr, const int32_t *tcol, const uint64_t *tmask,
                      int64_t *ctiles, int64_t *cnnz, int nthreads)
{
    int64_t ntile = (N + 63) / 64;
    int err = ORACLE_OK;
    set_threads(nthreads);
#pragma omp parallel
    {
        uint64_t *acc = (uint64_t *)calloc((size_t)(ntile > 0 ? ntile : 1), sizeof(uint64_t));
        int32_t *list = (int32_t *)malloc(sizeof(int32_t) * (size_t)(ntile > 0 ? ntile : 1));
        if (!acc || !list) {
#pragma omp atomic write
            err = ORACLE_ERR_ALLOC;
        }
#pragma omp for schedule(dynamic, 64)
        for (int64_t i = r0; i < r1; ++i) {
            if (!acc || !list)
                continue;
            int64_t nl = 0;
            for (int32_t e = Arp[i]; e < Arp[i + 1]; ++e) {
                int32_t j = Acol[e];
                for (int64_t t = tptr[j]; t < tptr[j + 1]; ++t) {
                    int32_t tc = tcol[t];
                    if (acc[tc] == 0)
                        list[nl++] = tc;
                    acc[tc] |= tmask[t];
                }
            }
            int64_t nz = 0;
            for (int64_t q = 0; q < nl; ++q) {
                nz += __builtin_popcountll(acc[list[q]]);
                acc[list[q]] = 0;
            }
            ctiles[i - r0] = nl;
            cnnz[i - r0] = nz;
        }
        free(acc);
        free(list);
    }
    return err;
}

/* Symbolic pass: nnz per row of C for rows [r0, r1) into cnt[0 .. r1-r0).
 * Every B column must lie in [0, N). */
int oracle_count_rows(int64_t r0, int64_t r1, int64_t N, const int32_t *Arp, const int32_t *Acol,
                      const int32_t *Brp, const int32_t *Bcol, int64_t *cnt, int nthreads)
{
    int err = ORACLE_OK;
    set_threads(nthreads);
#pragma omp parallel
    {
        int64_t *mark = (int64_t *)malloc(sizeof(int64_t) * (size_t)(N > 0 ? N : 1));
        if (!mark) {
#pragma omp atomic write
            err = ORACLE_ERR_ALLOC;
        } else {
            for (int64_t c = 0; c < N; ++c)
                mark[c] = -1;
        }
#pragma omp for schedule(dynamic, 64)
        for (int64_t i = r0; i < r1; ++i) {
            if (!mark)
                continue;
            int64_t n = 0;
            for (int32_t e = Arp[i]; e < Arp[i + 1]; ++e) {
                int32_t j = Acol[e];
                for (int32_t k = Brp[j]; k < Brp[j + 1]; ++k) {
                    int32_t c = Bcol[k];
                    if (mark[c] != i) {
                        mark[c] = i;
                        ++n;
                    }
                }
            }
            cnt[i - r0] = n;
        }
        free(mark);
    }
    return err;
}

/* Numeric pass: for rows [r0, r1), write sorted columns, values and sum|a*b| at
 * offsets off[i - r0] (relative to row r0, off has r1-r0+1 entries).  Cabs may be NULL. */
int oracle_fill_rows(int64_t r0, int64_t r1, int64_t N, const int32_t *Arp, const int32_t *Acol,
                     const double *Aval, const int32_t *Brp, const int32_t *Bcol, const double *Bval,
                     const int64_t *off, int32_t *Ccol, double *Cval, double *Cabs, int nthreads)
{
    int err = ORACLE_OK;
    set_threads(nthreads);
#pragma omp parallel
    {
        int64_t *mark = (int64_t *)malloc(sizeof(int64_t) * (size_t)(N > 0 ? N : 1));
        double *acc = (double *)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1));
        double *aacc = (double *)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1));
        int32_t *list = (int32_t *)malloc(sizeof(int32_t) * (size_t)(N > 0 ? N : 1));
        int ok = mark && acc && aacc && list;
        if (!ok) {
#pragma omp atomic write
            err = ORACLE_ERR_ALLOC;
        } else {
            for (int64_t c = 0; c < N; ++c)
                mark[c] = -1;
        }
#pragma omp for schedule(dynamic, 64)
        for (int64_t i = r0; i < r1; ++i) {
            if (!ok)
                continue;
            int64_t n = 0;
            int32_t cmin = INT32_MAX, cmax = -1;
            for (int32_t e = Arp[i]; e < Arp[i + 1]; ++e) {
                int32_t j = Acol[e];
                double a = Aval[e];
                for (int32_t k = Brp[j]; k < Brp[j + 1]; ++k) {
                    int32_t c = Bcol[k];
                    double p = a * Bval[k];
                    if (mark[c] != i) {
                        mark[c] = i;
                        list[n++] = c;
                        acc[c] = p;
                        aacc[c] = fabs(p);
                        if (c < cmin)
                            cmin = c;
                        if (c > cmax)
                            cmax = c;
                    } else {
                        acc[c] += p;
                        aacc[c] += fabs(p);
                    }
                }
            }
            int64_t base = off[i - r0];
            if (off[i - r0 + 1] - base != n) {
#pragma omp atomic write
                err = ORACLE_ERR_RANGE;
                continue;
            }
            if (n > 0 && (int64_t)(cmax - cmin) + 1 <= 4 * n) {
                /* dense enough: walk the column span in order */
                int64_t q = 0;
                for (int32_t c = cmin; c <= cmax; ++c)
                    if (mark[c] == i)
                        list[q++] = c;
            } else {
                sort_i32(list, n);
            }
            for (int64_t q = 0; q < n; ++q) {
                int32_t c = list[q];
                Ccol[base + q] = c;
                Cval[base + q] = acc[c];
                if (Cabs)
                    Cabs[base + q] = aacc[c];
            }
        }
        free(mark);
        free(acc);
        free(aacc);
        free(list);
    }
    return err;
}
