// rocsparse_check.cpp -- libhsmu_check.so: the GPU cross-check of HSMU SpGEMM against
// rocSPARSE's generic SpGEMM, the counterpart of the reference's cuSPARSE check
// (src/cusparse_in_tile.cu:4-272).  Test and driver infrastructure; libhsmu.so does not
// depend on it.

#include <hip/hip_runtime_api.h>
#include <rocsparse/rocsparse.h>

#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/hsmu_check.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_OR_FAIL(x)                                                                          \
    do {                                                                                        \
        hipError_t e_ = (x);                                                                    \
        if (e_ != hipSuccess)                                                                   \
            return fail(e_ == hipErrorOutOfMemory ? HSMU_ERR_NOMEM : HSMU_ERR_HIP, "%s: %s", #x, \
                        hipGetErrorString(e_));                                                 \
    } while (0)

#define RS_OR_FAIL(x)                                                             \
    do {                                                                          \
        rocsparse_status s_ = (x);                                                \
        if (s_ != rocsparse_status_success)                                       \
            return fail(HSMU_ERR_HIP, "%s: rocsparse status %d", #x, (int)s_); \
    } while (0)

// One rocSPARSE handle per device, created on first use and kept until the process ends.
rocsparse_handle handle_for_current_device()
{
    static std::mutex mu;
    static std::map<int, rocsparse_handle> handles;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess)
        return nullptr;
    std::lock_guard<std::mutex> lock(mu);
    auto it = handles.find(dev);
    if (it != handles.end())
        return it->second;
    rocsparse_handle h = nullptr;
    if (rocsparse_create_handle(&h) != rocsparse_status_success)
        return nullptr;
    handles[dev] = h;
    return h;
}

struct DevMem {
    void *p = nullptr;
    ~DevMem()
    {
        if (p)
            (void)hipFree(p);
    }
    void *release()
    {
        void *q = p;
        p = nullptr;
        return q;
    }
};

struct Mat {
    rocsparse_const_spmat_descr c = nullptr;
    rocsparse_spmat_descr m = nullptr;
    ~Mat()
    {
        if (c)
            rocsparse_destroy_spmat_descr(c);
        if (m)
            rocsparse_destroy_spmat_descr(m);
    }
};

struct Events {
    hipEvent_t e[4] = {nullptr, nullptr, nullptr, nullptr};
    ~Events()
    {
        for (auto x : e)
            if (x)
                (void)hipEventDestroy(x);
    }
};

int check_operand(const hsmu_csr *X, const char *name)
{
    if (!X)
        return fail(HSMU_ERR_INVALID, "%s is NULL", name);
    if (X->M < 0 || X->N < 0 || X->nnz < 0)
        return fail(HSMU_ERR_INVALID, "%s has a negative size", name);
    if (X->M > INT32_MAX || X->N > INT32_MAX || X->nnz > INT32_MAX)
        return fail(HSMU_ERR_UNSUPPORTED, "%s is beyond int32 indices (the checker uses 32-bit indices)", name);
    if (X->rowptr64)
        return fail(HSMU_ERR_UNSUPPORTED, "%s has int64 row pointers (the checker uses 32-bit indices)", name);
    if (!X->rowptr || (X->nnz > 0 && (!X->col || !X->val)))
        return fail(HSMU_ERR_INVALID, "%s has a NULL array", name);
    return HSMU_OK;
}

int rocsparse_product(hipStream_t s, const hsmu_csr *A, const hsmu_csr *B, hsmu_csr *C, double *ms_stages,
                      double *ms_wall)
{
    int rc = check_operand(A, "A");
    if (rc == HSMU_OK)
        rc = check_operand(B, "B");
    if (rc != HSMU_OK)
        return rc;
    if (!C)
        return fail(HSMU_ERR_INVALID, "C is NULL");
    if (A->N != B->M)
        return fail(HSMU_ERR_INVALID, "A.N (%lld) != B.M (%lld)", (long long)A->N, (long long)B->M);
    if (A->val32 != B->val32)
        return fail(HSMU_ERR_INVALID, "A and B have different value types");
    memset(C, 0, sizeof(*C));
    rocsparse_handle h = handle_for_current_device();
    if (!h)
        return fail(HSMU_ERR_HIP, "rocsparse_create_handle failed");
    RS_OR_FAIL(rocsparse_set_stream(h, s));

    const rocsparse_datatype dt = A->val32 ? rocsparse_datatype_f32_r : rocsparse_datatype_f64_r;
    const rocsparse_indextype i32 = rocsparse_indextype_i32;
    const rocsparse_index_base zero = rocsparse_index_base_zero;
    const int64_t m = A->M, n = B->N;
    const size_t vsz = A->val32 ? 4 : 8;

    auto t0 = std::chrono::steady_clock::now();
    Mat a, b, d, c;
    RS_OR_FAIL(rocsparse_create_const_csr_descr(&a.c, A->M, A->N, A->nnz, A->rowptr, A->col, A->val, i32, i32, zero,
                                                dt));
    RS_OR_FAIL(rocsparse_create_const_csr_descr(&b.c, B->M, B->N, B->nnz, B->rowptr, B->col, B->val, i32, i32, zero,
                                                dt));
    // D (beta * D is not used): an empty m x n matrix with valid row pointers.
    DevMem d_rp, c_rp, c_col, c_val, buf;
    HIP_OR_FAIL(hipMalloc(&d_rp.p, (m + 1) * sizeof(int32_t)));
    HIP_OR_FAIL(hipMemsetAsync(d_rp.p, 0, (m + 1) * sizeof(int32_t), s));
    RS_OR_FAIL(rocsparse_create_const_csr_descr(&d.c, m, n, 0, d_rp.p, nullptr, nullptr, i32, i32, zero, dt));
    HIP_OR_FAIL(hipMalloc(&c_rp.p, (m + 1) * sizeof(int32_t)));
    RS_OR_FAIL(rocsparse_create_csr_descr(&c.m, m, n, 0, c_rp.p, nullptr, nullptr, i32, i32, zero, dt));

    const double one64 = 1.0, zero64 = 0.0;
    const float one32 = 1.0f, zero32 = 0.0f;
    const void *alpha = A->val32 ? (const void *)&one32 : (const void *)&one64;
    const void *beta = nullptr; // C = alpha * A * B; if rocSPARSE insists on beta, pass 0
    size_t bytes = 0;
    rocsparse_status st = rocsparse_spgemm(h, rocsparse_operation_none, rocsparse_operation_none, alpha, a.c, b.c,
                                           beta, d.c, c.m, dt, rocsparse_spgemm_alg_default,
                                           rocsparse_spgemm_stage_buffer_size, &bytes, nullptr);
    if (st == rocsparse_status_invalid_pointer) {
        beta = A->val32 ? (const void *)&zero32 : (const void *)&zero64;
        st = rocsparse_spgemm(h, rocsparse_operation_none, rocsparse_operation_none, alpha, a.c, b.c, beta, d.c,
                              c.m, dt, rocsparse_spgemm_alg_default, rocsparse_spgemm_stage_buffer_size, &bytes,
                              nullptr);
    }
    RS_OR_FAIL(st);
    HIP_OR_FAIL(hipMalloc(&buf.p, bytes > 256 ? bytes : 256));
    Events ev;
    for (auto &e : ev.e)
        HIP_OR_FAIL(hipEventCreate(&e));
    HIP_OR_FAIL(hipEventRecord(ev.e[0], s));
    RS_OR_FAIL(rocsparse_spgemm(h, rocsparse_operation_none, rocsparse_operation_none, alpha, a.c, b.c, beta, d.c,
                                c.m, dt, rocsparse_spgemm_alg_default, rocsparse_spgemm_stage_nnz, &bytes, buf.p));
    HIP_OR_FAIL(hipEventRecord(ev.e[1], s));
    int64_t rows = 0, cols = 0, nnz = 0;
    RS_OR_FAIL(rocsparse_spmat_get_size(c.m, &rows, &cols, &nnz));
    if (nnz < 0 || nnz > INT32_MAX)
        return fail(HSMU_ERR_OVERFLOW, "nnz(C) = %lld does not fit the checker's 32-bit indices", (long long)nnz);
    HIP_OR_FAIL(hipMalloc(&c_col.p, (nnz > 0 ? nnz : 1) * sizeof(int32_t)));
    HIP_OR_FAIL(hipMalloc(&c_val.p, (nnz > 0 ? nnz : 1) * vsz));
    RS_OR_FAIL(rocsparse_csr_set_pointers(c.m, c_rp.p, c_col.p, c_val.p));
    HIP_OR_FAIL(hipEventRecord(ev.e[2], s));
    RS_OR_FAIL(rocsparse_spgemm(h, rocsparse_operation_none, rocsparse_operation_none, alpha, a.c, b.c, beta, d.c,
                                c.m, dt, rocsparse_spgemm_alg_default, rocsparse_spgemm_stage_compute, &bytes, buf.p));
    HIP_OR_FAIL(hipEventRecord(ev.e[3], s));
    HIP_OR_FAIL(hipEventSynchronize(ev.e[3]));
    auto t1 = std::chrono::steady_clock::now();
    float f01 = 0.f, f23 = 0.f;
    HIP_OR_FAIL(hipEventElapsedTime(&f01, ev.e[0], ev.e[1]));
    HIP_OR_FAIL(hipEventElapsedTime(&f23, ev.e[2], ev.e[3]));
    if (ms_stages)
        *ms_stages = (double)f01 + (double)f23;
    if (ms_wall)
        *ms_wall = std::chrono::duration<double, std::milli>(t1 - t0).count();
    C->M = m;
    C->N = n;
    C->nnz = nnz;
    C->rowptr = c_rp.release();
    C->col = (int32_t *)c_col.release();
    C->val = c_val.release();
    C->rowptr64 = 0;
    C->val32 = A->val32;
    return HSMU_OK;
}

int download_rowptr(const hsmu_csr *X, std::vector<int64_t> &rp)
{
    rp.assign((size_t)X->M + 1, 0);
    if (X->rowptr64) {
        HIP_OR_FAIL(hipMemcpy(rp.data(), X->rowptr, rp.size() * 8, hipMemcpyDeviceToHost));
    } else {
        std::vector<int32_t> r32(rp.size());
        HIP_OR_FAIL(hipMemcpy(r32.data(), X->rowptr, r32.size() * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < r32.size(); ++i)
            rp[i] = r32[i];
    }
    return HSMU_OK;
}

int download_vals(const hsmu_csr *X, int64_t count, std::vector<double> &v)
{
    v.assign((size_t)count, 0.0);
    if (count == 0)
        return HSMU_OK;
    if (X->val32) {
        std::vector<float> f((size_t)count);
        HIP_OR_FAIL(hipMemcpy(f.data(), X->val, f.size() * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < f.size(); ++i)
            v[i] = f[i];
    } else {
        HIP_OR_FAIL(hipMemcpy(v.data(), X->val, v.size() * 8, hipMemcpyDeviceToHost));
    }
    return HSMU_OK;
}

int compare(const hsmu_csr *X, const hsmu_csr *Y, const hsmu_csr *S, double tol, hsmu_csr_diff *out)
{
    if (!X || !Y || !out)
        return fail(HSMU_ERR_INVALID, "NULL argument");
    if (X->M != Y->M || X->N != Y->N || (S && (S->M != X->M || S->N != X->N)))
        return fail(HSMU_ERR_INVALID, "shapes differ");
    hsmu_csr_diff d;
    memset(&d, 0, sizeof(d));
    d.first_row = d.first_entry = -1;
    std::vector<int64_t> rx, ry;
    int rc = download_rowptr(X, rx);
    if (rc == HSMU_OK)
        rc = download_rowptr(Y, ry);
    if (rc != HSMU_OK)
        return rc;
    for (int64_t i = 0; i <= X->M; ++i)
        if (rx[i] != ry[i]) {
            if (d.first_row < 0)
                d.first_row = i;
            ++d.rowptr_mismatch;
        }
    if (d.rowptr_mismatch) {
        *out = d;
        return HSMU_OK;
    }
    const int64_t nnz = rx[X->M] - rx[0];
    std::vector<int32_t> cx((size_t)nnz), cy((size_t)nnz);
    if (nnz > 0) {
        HIP_OR_FAIL(hipMemcpy(cx.data(), X->col + rx[0], nnz * 4, hipMemcpyDeviceToHost));
        HIP_OR_FAIL(hipMemcpy(cy.data(), Y->col + ry[0], nnz * 4, hipMemcpyDeviceToHost));
    }
    std::vector<double> vx, vy, vs;
    hsmu_csr xs = *X, ys = *Y, ss;
    xs.val = (char *)X->val + rx[0] * (X->val32 ? 4 : 8);
    ys.val = (char *)Y->val + ry[0] * (Y->val32 ? 4 : 8);
    rc = download_vals(&xs, nnz, vx);
    if (rc == HSMU_OK)
        rc = download_vals(&ys, nnz, vy);
    if (rc == HSMU_OK && S) {
        std::vector<int64_t> rs;
        rc = download_rowptr(S, rs);
        if (rc == HSMU_OK && (rs[S->M] - rs[0] != nnz))
            return fail(HSMU_ERR_INVALID, "S has a different number of entries");
        if (rc == HSMU_OK) {
            ss = *S;
            ss.val = (char *)S->val + rs[0] * (S->val32 ? 4 : 8);
            rc = download_vals(&ss, nnz, vs);
        }
    }
    if (rc != HSMU_OK)
        return rc;
    for (int64_t k = 0; k < nnz; ++k) {
        if (cx[k] != cy[k]) {
            if (d.first_entry < 0)
                d.first_entry = k;
            ++d.col_mismatch;
            continue;
        }
        const double e = std::fabs(vx[k] - vy[k]);
        const double scale = S ? std::fabs(vs[k]) : std::fmax(std::fabs(vx[k]), std::fabs(vy[k]));
        d.max_abs_err = std::fmax(d.max_abs_err, e);
        if (scale > 0)
            d.max_rel_err = std::fmax(d.max_rel_err, e / scale);
        if (e > tol * scale) {
            if (d.first_entry < 0)
                d.first_entry = k;
            ++d.val_mismatch;
        }
    }
    *out = d;
    return HSMU_OK;
}

// |x| of a device value array into a new device array.
int abs_copy(const hsmu_csr *X, DevMem &out)
{
    const size_t n = (size_t)X->nnz;
    HIP_OR_FAIL(hipMalloc(&out.p, (n ? n : 1) * 8));
    std::vector<double> v;
    int rc = download_vals(X, X->nnz, v);
    if (rc != HSMU_OK)
        return rc;
    for (auto &x : v)
        x = std::fabs(x);
    if (n)
        HIP_OR_FAIL(hipMemcpy(out.p, v.data(), n * 8, hipMemcpyHostToDevice));
    return HSMU_OK;
}

} // namespace

extern "C" {

int hsmu_rocsparse_spgemm(void *stream, const hsmu_csr *A, const hsmu_csr *B, hsmu_csr *C, double *ms_stages,
                          double *ms_wall)
{
    return rocsparse_product((hipStream_t)stream, A, B, C, ms_stages, ms_wall);
}

int hsmu_check_free(hsmu_csr *C)
{
    if (!C)
        return fail(HSMU_ERR_INVALID, "C is NULL");
    if (C->rowptr)
        HIP_OR_FAIL(hipFree(C->rowptr));
    if (C->col)
        HIP_OR_FAIL(hipFree(C->col));
    if (C->val)
        HIP_OR_FAIL(hipFree(C->val));
    memset(C, 0, sizeof(*C));
    return HSMU_OK;
}

int hsmu_csr_compare(const hsmu_csr *X, const hsmu_csr *Y, const hsmu_csr *S, double tol, hsmu_csr_diff *diff)
{
    return compare(X, Y, S, tol, diff);
}

const char *hsmu_check_last_error(void)
{
    return g_err.c_str();
}

int tile_rocsparse_spgemm(NHC_CSR *a, NHC_CSR *b, NHC_CSR * /*test_C*/, NHC_CSR *c)
{
    auto as_csr = [](const NHC_CSR *x, int64_t M, int64_t N) {
        hsmu_csr h;
        memset(&h, 0, sizeof(h));
        h.M = M;
        h.N = N;
        h.nnz = x->nnz;
        h.rowptr = x->d_ptr;
        h.col = x->d_col;
        h.val = x->d_val;
        return h;
    };
    hsmu_csr A = as_csr(a, a->M, a->N), B = as_csr(b, b->M, b->N), R, S;
    double ms_stages = 0, ms_wall = 0;
    int rc = rocsparse_product(nullptr, &A, &B, &R, &ms_stages, &ms_wall);
    if (rc != HSMU_OK) {
        printf("rocSPARSE failed! (%s)\n", g_err.c_str());
        return 2;
    }
    printf("the rocsparse C_nnz is %lld\n", (long long)R.nnz);
    printf("the run time using rocsparse for the %s is:%f[ms]\n", a->matrix_name ? a->matrix_name : "input",
           ms_wall);
    bool ok = true;
    // row pointers against the host copy NHC_spgemm left in c->rowPtr
    std::vector<int64_t> rp;
    rc = download_rowptr(&R, rp);
    if (rc != HSMU_OK || !c->rowPtr) {
        printf("rocSPARSE failed! (%s)\n", c->rowPtr ? g_err.c_str() : "C has no host row pointers");
        hsmu_check_free(&R);
        return 2;
    }
    int64_t wrong = 0;
    for (int64_t i = 0; i <= (int64_t)a->M; ++i)
        if (rp[i] != (int64_t)c->rowPtr[i]) {
            if (wrong < 10)
                printf("the rocsparse ptr[%lld] is %lld,the NHC_spgemm ptr[%lld] result is %d\n", (long long)i,
                       (long long)rp[i], (long long)i, c->rowPtr[i]);
            ++wrong;
        }
    if (wrong) {
        ok = false;
        printf("the cptr is wrong! the wrong times is %lld\n", (long long)wrong);
    } else {
        printf("compute the Cptr is right!\n");
        long long products = 0;
        for (int64_t k = 0; k < (int64_t)a->nnz; ++k) {
            const int j = a->col[k];
            products += b->rowPtr[j + 1] - b->rowPtr[j];
        }
        printf("for %s rocsparse gflops = %lf\n", a->matrix_name ? a->matrix_name : "input",
               ms_stages > 0 ? 2.0 * (double)products / (ms_stages * 1e6) : 0.0);
        // values to 1e-12 of sum |a*b|, which rocSPARSE computes from |A| and |B|
        DevMem absA, absB;
        hsmu_csr Aa = A, Ba = B, C = as_csr(c, c->M, c->N);
        C.nnz = R.nnz;
        rc = abs_copy(&A, absA);
        if (rc == HSMU_OK)
            rc = abs_copy(&B, absB);
        Aa.val = absA.p;
        Ba.val = absB.p;
        if (rc == HSMU_OK)
            rc = rocsparse_product(nullptr, &Aa, &Ba, &S, nullptr, nullptr);
        hsmu_csr_diff d;
        if (rc == HSMU_OK) {
            rc = compare(&C, &R, &S, 1e-12, &d);
            hsmu_check_free(&S);
        }
        if (rc != HSMU_OK) {
            printf("rocSPARSE failed! (%s)\n", g_err.c_str());
            hsmu_check_free(&R);
            return 2;
        }
        if (d.val_mismatch) {
            ok = false;
            printf("the cval is wrong! the wrong times is %lld\n", (long long)d.val_mismatch);
        } else {
            printf("compute the Cval is right!\n");
        }
        if (d.col_mismatch) {
            ok = false;
            printf("the ccol is wrong! the wrong times is %lld (first at entry %lld)\n", (long long)d.col_mismatch,
                   (long long)d.first_entry);
        } else {
            printf("compute the Ccol is right!\n");
        }
        printf("max relative error against rocSPARSE: %.3e of sum |a*b|\n", d.max_rel_err);
    }
    if (ok)
        printf("the result is all right!\n");
    hsmu_check_free(&R);
    return ok ? 0 : 1;
}

} // extern "C"
