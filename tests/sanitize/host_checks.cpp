// host_checks.cpp -- drives the host-side C/C++ of the library (synthetic generators, the
// transpose, the Matrix Market reader, the autotuner's shape classes and table lookup) and
// the CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer.  Built and run by
// tests/test_host_sanitizers.py with g++/gcc only (no HIP): any invalid access aborts the
// run.  The device entry points that tune.cpp also references are never reached here
// (autotune runs with search = 0), so they are defined below to fail loudly.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "hsmu.h"

extern "C" {
int64_t oracle_products(int64_t r0, int64_t r1, const int32_t *Arp, const int32_t *Acol, const int32_t *Brp,
                        int64_t *row_products);
int64_t oracle_btile_counts(int64_t K, const int32_t *Brp, const int32_t *Bcol, int32_t *tcount);
int oracle_btile_fill(int64_t K, const int32_t *Brp, const int32_t *Bcol, const int64_t *tptr, int32_t *tcol,
                      uint64_t *tmask);
int oracle_ctile_rows(int64_t r0, int64_t r1, int64_t N, const int32_t *Arp, const int32_t *Acol, const int64_t *tptr,
                      const int32_t *tcol, const uint64_t *tmask, int64_t *ctiles, int64_t *cnnz, int nthreads);
int oracle_count_rows(int64_t r0, int64_t r1, int64_t N, const int32_t *Arp, const int32_t *Acol, const int32_t *Brp,
                      const int32_t *Bcol, int64_t *cnt, int nthreads);
int oracle_fill_rows(int64_t r0, int64_t r1, int64_t N, const int32_t *Arp, const int32_t *Acol, const double *Aval,
                     const int32_t *Brp, const int32_t *Bcol, const double *Bval, const int64_t *off, int32_t *Ccol,
                     double *Cval, double *Cabs, int nthreads);

// device entry points named by tune.cpp; unreachable with search = 0
int hsmu_spgemm(hsmu_ctx *, const hsmu_csr *, const hsmu_csr *, hsmu_csr *, const hsmu_opts *, hsmu_stats *) { abort(); }
int hsmu_spgemm_batched(hsmu_ctx *, int64_t, const hsmu_csr *, const hsmu_csr *, hsmu_csr *, const hsmu_opts *,
                        hsmu_stats *) { abort(); }
int hsmu_csr_free(hsmu_ctx *, hsmu_csr *) { abort(); }
int hsmu_batch_free(hsmu_ctx *, int64_t, hsmu_csr *) { abort(); }
void hsmu_opts_default(hsmu_opts *o)
{
    memset(o, 0, sizeof(*o));
    o->timing = 1;
}
}

static int failures = 0;
#define CHECK(c)                                                                                                       \
    do {                                                                                                               \
        if (!(c)) {                                                                                                    \
            fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__);                                             \
            ++failures;                                                                                                \
        }                                                                                                              \
    } while (0)

// C = A * B with the oracle (two passes), checked against the product count.
static void oracle_product(const hsmu_host_csr &A, const hsmu_host_csr &B)
{
    const int64_t M = A.M;
    std::vector<int64_t> rp((size_t)M);
    const int64_t prods = oracle_products(0, M, A.rowptr, A.col, B.rowptr, rp.data());
    CHECK(prods >= 0);
    std::vector<int32_t> tc((size_t)B.M + 1);
    const int64_t T = oracle_btile_counts(B.M, B.rowptr, B.col, tc.data());
    CHECK(T >= 0 && T <= B.nnz);
    std::vector<int64_t> tptr((size_t)B.M + 1, 0);
    for (int64_t j = 0; j < B.M; ++j)
        tptr[(size_t)j + 1] = tptr[(size_t)j] + tc[(size_t)j];
    std::vector<int32_t> tcol((size_t)(T > 0 ? T : 1));
    std::vector<uint64_t> tmask((size_t)(T > 0 ? T : 1));
    CHECK(oracle_btile_fill(B.M, B.rowptr, B.col, tptr.data(), tcol.data(), tmask.data()) == 0);
    std::vector<int64_t> ct((size_t)M + 1), cn((size_t)M + 1), cnt((size_t)M + 1);
    CHECK(oracle_ctile_rows(0, M, B.N, A.rowptr, A.col, tptr.data(), tcol.data(), tmask.data(), ct.data(), cn.data(),
                            2) == 0);
    CHECK(oracle_count_rows(0, M, B.N, A.rowptr, A.col, B.rowptr, B.col, cnt.data(), 2) == 0);
    std::vector<int64_t> off((size_t)M + 1, 0);
    for (int64_t i = 0; i < M; ++i) {
        CHECK(cn[(size_t)i] == cnt[(size_t)i]);
        off[(size_t)i + 1] = off[(size_t)i] + cnt[(size_t)i];
    }
    const int64_t nnz = off[(size_t)M];
    std::vector<int32_t> col((size_t)(nnz > 0 ? nnz : 1));
    std::vector<double> val((size_t)(nnz > 0 ? nnz : 1)), ab((size_t)(nnz > 0 ? nnz : 1));
    CHECK(oracle_fill_rows(0, M, B.N, A.rowptr, A.col, A.val, B.rowptr, B.col, B.val, off.data(), col.data(),
                           val.data(), ab.data(), 2) == 0);
    for (int64_t i = 0; i < M; ++i)
        for (int64_t k = off[(size_t)i] + 1; k < off[(size_t)i + 1]; ++k)
            CHECK(col[(size_t)k - 1] < col[(size_t)k]);
}

static std::string write_file(const char *dir, const char *name, const char *text)
{
    std::string p = std::string(dir) + "/" + name;
    FILE *f = fopen(p.c_str(), "w");
    if (!f) {
        perror("fopen");
        exit(2);
    }
    fputs(text, f);
    fclose(f);
    return p;
}

int main(int argc, char **argv)
{
    const char *dir = argc > 1 ? argv[1] : ".";
    hsmu_host_csr m[6];
    CHECK(hsmu_gen_laplace2d(20, &m[0]) == HSMU_OK);
    CHECK(hsmu_gen_laplace3d(6, &m[1]) == HSMU_OK);
    CHECK(hsmu_gen_banded(500, 4, 30, 40, 1, &m[2]) == HSMU_OK);
    CHECK(hsmu_gen_rmat(9, 16, 0.57, 0.19, 0.19, 2, &m[3]) == HSMU_OK);
    CHECK(hsmu_gen_zipf(300, 100, 8, 1.2, 4, &m[4]) == HSMU_OK);
    CHECK(hsmu_gen_random(50, 60, 9, 7, &m[5]) == HSMU_OK);
    for (int i = 0; i < 4; ++i)
        oracle_product(m[i], m[i]);
    hsmu_host_csr zt, rt, rtt;
    CHECK(hsmu_transpose(&m[4], &zt) == HSMU_OK);
    oracle_product(m[4], zt);
    CHECK(hsmu_transpose(&m[5], &rt) == HSMU_OK);
    CHECK(hsmu_transpose(&rt, &rtt) == HSMU_OK);
    CHECK(rtt.nnz == m[5].nnz && memcmp(rtt.col, m[5].col, sizeof(int32_t) * (size_t)m[5].nnz) == 0);
    oracle_product(m[5], rt);

    // Matrix Market reader: accepted forms, then malformed files that must be rejected
    struct Case {
        const char *name, *text;
        int status;
        int64_t nnz;
    } cases[] = {
        {"g.mtx", "%%MatrixMarket matrix coordinate real general\n% c\n3 4 5\n1 4 7\n1 2 3\n3 1 -2\n1 2 5\n2 3 1\n",
         HSMU_OK, 5},
        {"s.mtx", "%%MatrixMarket matrix coordinate pattern symmetric\n3 3 3\n1 1\n3 1\n2 2\n", HSMU_OK, 4},
        {"c.mtx", "%%MatrixMarket matrix coordinate complex hermitian\n2 2 2\n1 1 1.0 0.0\n2 1 2.0 1.0\n", HSMU_OK, 3},
        {"ns.mtx", "%%MatrixMarket matrix coordinate real symmetric\n2 40000 2\n1 1 1.0\n2 39999 2.0\n",
         HSMU_ERR_INVALID, 0},
        {"big.mtx", "%%MatrixMarket matrix coordinate real general\n4 4 9000000000000000\n1 1 1.0\n",
         HSMU_ERR_INVALID, 0},
        {"idx.mtx", "%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n", HSMU_ERR_INDEX, 0},
        {"cut.mtx", "%%MatrixMarket matrix coordinate real general\n5 5 6\n1 1 1.0\n2 2 1.0\n", HSMU_ERR_INVALID, 0},
        {"cut2.mtx", "%%MatrixMarket matrix coordinate real general\n5 5 3\n1 1 1.0\n2 2 1.0\n", HSMU_ERR_INDEX, 0},
        {"arr.mtx", "%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n", HSMU_ERR_UNSUPPORTED, 0},
    };
    for (const Case &c : cases) {
        hsmu_host_csr r;
        int32_t sym = -1;
        const int rc = hsmu_read_mtx(write_file(dir, c.name, c.text).c_str(), &r, &sym);
        if (rc != c.status)
            fprintf(stderr, "%s: status %d, expected %d\n", c.name, rc, c.status);
        CHECK(rc == c.status);
        if (rc == HSMU_OK) {
            CHECK(r.nnz == c.nnz);
            hsmu_host_csr_free(&r);
        }
    }
    hsmu_host_csr r;
    CHECK(hsmu_read_mtx((std::string(dir) + "/missing.mtx").c_str(), &r, nullptr) == HSMU_ERR_INVALID);

    // autotuner: shape classes and the table lookup (lines of every length and form)
    hsmu_csr A = {m[2].M, m[2].N, m[2].nnz, nullptr, nullptr, nullptr, 0, 0};
    char cls[16];
    CHECK(hsmu_shape_class(1, &A, &A, cls, sizeof(cls)) == HSMU_OK);
    char tiny[4];
    CHECK(hsmu_shape_class(1, &A, &A, tiny, sizeof(tiny)) == HSMU_ERR_INVALID);
    std::string table = "# class ppl tpl bl tl budget persistent rebuild ms\nshort line\n";
    table += std::string(600, 'x') + " 1 2 3 4 5 6 7 8.0\n";
    table += std::string(cls) + " 16 8 4 2 4096 1 0 1.25\n";
    const std::string tp = write_file(dir, "table.tsv", table.c_str());
    hsmu_opts base, out;
    hsmu_opts_default(&base);
    int32_t found = 0;
    hsmu_ctx *unused = (hsmu_ctx *)&base; // never dereferenced with search = 0
    CHECK(hsmu_autotune(unused, 1, &A, &A, &base, tp.c_str(), 0, &out, &found) == HSMU_OK);
    CHECK(found == 1 && out.products_per_lane == 16 && out.lds_block_budget == 4096 && out.persistent == 1);

    for (auto &x : m)
        hsmu_host_csr_free(&x);
    hsmu_host_csr_free(&zt);
    hsmu_host_csr_free(&rt);
    hsmu_host_csr_free(&rtt);
    if (failures) {
        fprintf(stderr, "%d checks failed\n", failures);
        return 1;
    }
    printf("HOST_SANITIZE_OK\n");
    return 0;
}
