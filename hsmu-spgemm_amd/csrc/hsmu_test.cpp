// hsmu_test.cpp -- command-line driver, the counterpart of the reference's ./test
// (src/test.cu:10-94): load A, form B = A (or A^T with --aat, the reference's AAT flag),
// upload, run NHC_spgemm, print the reference's lines, release everything.
//
//   hsmu_test <matrix.mtx | gen:SPEC> [--aat] [--csv-dir DIR] [--check]
//
// gen:SPEC names a synthetic input instead of a file:
//   gen:laplace2d:N   gen:laplace3d:N   gen:banded:M[:SEED]   gen:rmat:SCALE[:EF[:SEED]]
//   gen:zipf:M:N[:SEED]
// --csv-dir DIR makes NHC_spgemm append the reference's three result CSVs
// (src/compressed/compressed_NHC_spgemm.cu:220-246) under DIR.
// --check runs the rocSPARSE cross-check after NHC_spgemm (the reference's CHECK_RESULT
// build, src/test.cu:73-76 -> tile_cusparse_spgemm); the exit status is 1 on a mismatch.

#include <sys/time.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/NHC_spgemm.h"
#include "../../include/hsmu.h"
#include "../../include/hsmu_check.h"

namespace {

std::vector<std::string> split(const std::string &s, char sep)
{
    std::vector<std::string> out;
    size_t a = 0;
    while (true) {
        size_t b = s.find(sep, a);
        out.push_back(s.substr(a, b == std::string::npos ? std::string::npos : b - a));
        if (b == std::string::npos)
            break;
        a = b + 1;
    }
    return out;
}

int load(const std::string &spec, hsmu_host_csr *m, int *sym)
{
    *sym = 0;
    if (spec.rfind("gen:", 0) != 0)
        return hsmu_read_mtx(spec.c_str(), m, sym);
    auto f = split(spec.substr(4), ':');
    auto num = [&](size_t i, long long d) { return f.size() > i ? atoll(f[i].c_str()) : d; };
    if (f[0] == "laplace2d")
        return hsmu_gen_laplace2d((int32_t)num(1, 128), m);
    if (f[0] == "laplace3d")
        return hsmu_gen_laplace3d((int32_t)num(1, 32), m);
    if (f[0] == "banded")
        return hsmu_gen_banded((int32_t)num(1, 62451), 40, 88, 400, (uint64_t)num(2, 1), m);
    if (f[0] == "rmat")
        return hsmu_gen_rmat((int32_t)num(1, 20), (int32_t)num(2, 16), 0.57, 0.19, 0.19, (uint64_t)num(3, 2), m);
    if (f[0] == "zipf")
        return hsmu_gen_zipf((int32_t)num(1, 500000), (int32_t)num(2, 100000), 8, 1.2, (uint64_t)num(3, 4), m);
    return HSMU_ERR_INVALID;
}

void to_nhc(hsmu_host_csr *h, NHC_CSR *X, int sym, char *name)
{
    memset(X, 0, sizeof(*X));
    X->M = (int)h->M;
    X->N = (int)h->N;
    X->nnz = (int)h->nnz;
    X->rowPtr = h->rowptr;
    X->col = h->col;
    X->val = h->val;
    X->isSymmetric = sym;
    X->matrix_name = name;
    h->rowptr = nullptr; // ownership moves to X (released by realse_NHCcsr)
    h->col = nullptr;
    h->val = nullptr;
}

} // namespace

int main(int argc, char **argv)
{
    if (argc < 2) {
        printf("Run the code by './test matrix.mtx'.\n");
        return 0;
    }
    std::string spec = argv[1];
    bool aat = false, check = false;
    std::string csv_dir;
    for (int i = 2; i < argc; ++i) {
        if (!strcmp(argv[i], "--aat"))
            aat = true;
        else if (!strcmp(argv[i], "--csv-dir") && i + 1 < argc)
            csv_dir = argv[++i];
        else if (!strcmp(argv[i], "--check"))
            check = true;
    }
    printf("the matrix is: -------------- %s --------------\n", spec.c_str());
    if (spec.rfind("gen:", 0) != 0) {
        FILE *f = fopen(spec.c_str(), "r");
        if (!f) {
            printf("Error opening file %s\n", spec.c_str());
            return 1;
        }
        fclose(f);
    }
    printf("success open file\n");

    struct timeval t1, t2;
    gettimeofday(&t1, nullptr);
    hsmu_host_csr ha, hb;
    memset(&hb, 0, sizeof(hb));
    int sym = 0;
    int rc = load(spec, &ha, &sym);
    if (rc != HSMU_OK) {
        fprintf(stderr, "loading %s failed: %s\n", spec.c_str(), hsmu_status_string(rc));
        return 1;
    }
    // B: A^T when --aat (src/test.cu:39-56), otherwise a second copy of A (:59)
    if (aat) {
        rc = hsmu_transpose(&ha, &hb);
    } else {
        int s2;
        rc = load(spec, &hb, &s2);
    }
    if (rc != HSMU_OK) {
        fprintf(stderr, "forming B failed: %s\n", hsmu_status_string(rc));
        return 1;
    }
    gettimeofday(&t2, nullptr);
    const double ms_load = (t2.tv_sec - t1.tv_sec) * 1000.0 + (t2.tv_usec - t1.tv_usec) / 1000.0;
    std::vector<char> name(spec.begin(), spec.end());
    name.push_back('\0');
    NHC_CSR A, B, C;
    NHC_mask_matrix mask;
    NHC_bin c_bin;
    compressed_bin cbin;
    memset(&C, 0, sizeof(C));
    memset(&mask, 0, sizeof(mask));
    memset(&c_bin, 0, sizeof(c_bin));
    memset(&cbin, 0, sizeof(cbin));
    to_nhc(&ha, &A, sym, name.data());
    to_nhc(&hb, &B, sym, name.data());
    printf("input matrix A: ( %i, %i ) nnz = %lld\nloadfile time    = %4.5f sec\n", A.M, A.N, (long long)A.nnz,
           ms_load / 1000.0);
    C.M = A.M;
    C.N = B.N;
    HtD_value_col_ptr(&A);
    HtD_value_col_ptr(&B);

    if (!csv_dir.empty())
        setenv("HSMU_CSV_DIR", csv_dir.c_str(), 1);
    NHC_spgemm(&cbin, &c_bin, &A, &B, &C, &mask);
    const int status = NHC_last_status();
    int check_rc = 0;
    if (check && status == HSMU_OK)
        check_rc = tile_rocsparse_spgemm(&A, &B, nullptr, &C);

    realse_NHCcsr(&A);
    realse_NHCcsr(&B);
    free_C_matrix(&C);
    if (A.N < threshold_value_selected) {
        free_mask_matrix(&mask);
    } else {
        realse_NHC_tile(&B);
        realse_NHC_tile(&C);
        realse_compressed_bin(&cbin);
    }
    free_bin(&c_bin);
    if (status != HSMU_OK)
        return 2;
    return check_rc == 0 ? 0 : (check_rc == 1 ? 1 : 3);
}
