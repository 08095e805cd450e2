// host_inputs.cpp -- host-side inputs of libhsmu.so: seeded synthetic generators for the
// benchmark workloads, the Matrix Market reader and the CSR transpose.
//
// Generators are deterministic for a given seed and independent of the number of
// threads: random streams are split into fixed chunks whose seeds are derived from
// (seed, chunk index).  Every generated row is sorted by column.

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/hsmu.h"

namespace {

// xoshiro256** (Blackman & Vigna), seeded through splitmix64.
struct Rng {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t &x)
    {
        uint64_t z = (x += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    explicit Rng(uint64_t seed)
    {
        uint64_t x = seed;
        for (auto &v : s)
            v = splitmix(x);
    }
    static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    uint64_t next()
    {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; } // [0, 1)
    uint64_t below(uint64_t n) { return n ? next() % n : 0; }
    double normal()
    {
        // Box-Muller, one value per call
        double u1 = uniform();
        while (u1 <= 0.0)
            u1 = uniform();
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

uint64_t derive_seed(uint64_t seed, uint64_t chunk)
{
    uint64_t x = seed ^ (0x6a09e667f3bcc909ull * (chunk + 1));
    return Rng::splitmix(x);
}

bool alloc_host(hsmu_host_csr *m, int64_t M, int64_t N, int64_t nnz)
{
    m->M = M;
    m->N = N;
    m->nnz = nnz;
    m->rowptr = (int32_t *)malloc(sizeof(int32_t) * (size_t)(M + 1));
    m->col = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
    m->val = (double *)malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
    if (!m->rowptr || !m->col || !m->val) {
        hsmu_host_csr_free(m);
        return false;
    }
    return true;
}

// Laplacian stencil on a grid of `dims` dimensions with side n: diagonal 2*dims,
// -1 for each axis neighbour inside the grid; row index x + n*y (+ n*n*z).
int laplace(int dims, int32_t n, hsmu_host_csr *out)
{
    if (!out || n <= 0)
        return HSMU_ERR_INVALID;
    memset(out, 0, sizeof(*out));
    int64_t M = 1;
    for (int d = 0; d < dims; ++d)
        M *= n;
    if (M > INT32_MAX / 8)
        return HSMU_ERR_UNSUPPORTED;
    int64_t nnz = M;
    for (int d = 0; d < dims; ++d) {
        int64_t planes = M / n;
        nnz += 2 * planes * (n - 1);
    }
    if (!alloc_host(out, M, M, nnz))
        return HSMU_ERR_NOMEM;
    int64_t k = 0;
    const int64_t n2 = (int64_t)n * n;
    for (int64_t i = 0; i < M; ++i) {
        out->rowptr[i] = (int32_t)k;
        const int64_t x = i % n, y = (i / n) % n, z = dims == 3 ? i / n2 : 0;
        // neighbours in ascending column order
        if (dims == 3 && z > 0) {
            out->col[k] = (int32_t)(i - n2);
            out->val[k++] = -1.0;
        }
        if (y > 0) {
            out->col[k] = (int32_t)(i - n);
            out->val[k++] = -1.0;
        }
        if (x > 0) {
            out->col[k] = (int32_t)(i - 1);
            out->val[k++] = -1.0;
        }
        out->col[k] = (int32_t)i;
        out->val[k++] = 2.0 * dims;
        if (x < n - 1) {
            out->col[k] = (int32_t)(i + 1);
            out->val[k++] = -1.0;
        }
        if (y < n - 1) {
            out->col[k] = (int32_t)(i + n);
            out->val[k++] = -1.0;
        }
        if (dims == 3 && z < n - 1) {
            out->col[k] = (int32_t)(i + n2);
            out->val[k++] = -1.0;
        }
    }
    out->rowptr[M] = (int32_t)k;
    return k == nnz ? HSMU_OK : HSMU_ERR_INVALID;
}

// Build CSR from per-row column lists (sorted, distinct) and values.
int from_rows(int64_t M, int64_t N, const std::vector<std::vector<int32_t>> &cols,
              const std::vector<std::vector<double>> &vals, hsmu_host_csr *out)
{
    int64_t nnz = 0;
    for (auto &r : cols)
        nnz += (int64_t)r.size();
    if (nnz > INT32_MAX)
        return HSMU_ERR_UNSUPPORTED;
    if (!alloc_host(out, M, N, nnz))
        return HSMU_ERR_NOMEM;
    int64_t k = 0;
    for (int64_t i = 0; i < M; ++i) {
        out->rowptr[i] = (int32_t)k;
        for (size_t q = 0; q < cols[i].size(); ++q) {
            out->col[k] = cols[i][q];
            out->val[k++] = vals[i][q];
        }
    }
    out->rowptr[M] = (int32_t)k;
    return HSMU_OK;
}

} // namespace

extern "C" {

void hsmu_host_csr_free(hsmu_host_csr *m)
{
    if (!m)
        return;
    free(m->rowptr);
    free(m->col);
    free(m->val);
    m->rowptr = nullptr;
    m->col = nullptr;
    m->val = nullptr;
    m->M = m->N = m->nnz = 0;
}

int hsmu_gen_laplace2d(int32_t n, hsmu_host_csr *out) { return laplace(2, n, out); }
int hsmu_gen_laplace3d(int32_t n, hsmu_host_csr *out) { return laplace(3, n, out); }

int hsmu_gen_banded(int32_t M, int32_t lo, int32_t hi, int32_t halfwidth, uint64_t seed, hsmu_host_csr *out)
{
    if (!out || M <= 0 || lo < 0 || hi < lo || halfwidth < 0)
        return HSMU_ERR_INVALID;
    memset(out, 0, sizeof(*out));
    const int64_t chunk = 4096;
    const int64_t nchunks = (M + chunk - 1) / chunk;
    std::vector<std::vector<int32_t>> cols(M);
    std::vector<std::vector<double>> vals(M);
    bool bad = false;
#pragma omp parallel for schedule(dynamic)
    for (int64_t ch = 0; ch < nchunks; ++ch) {
        Rng rng(derive_seed(seed, (uint64_t)ch));
        std::vector<char> used;
        for (int64_t i = ch * chunk; i < std::min<int64_t>(M, (ch + 1) * chunk); ++i) {
            const int64_t c0 = std::max<int64_t>(0, i - halfwidth);
            const int64_t c1 = std::min<int64_t>(M - 1, i + halfwidth);
            const int64_t w = c1 - c0 + 1;
            int64_t cnt = lo + (int64_t)rng.below((uint64_t)(hi - lo + 1));
            if (cnt > w) {
                bad = true;
                cnt = w;
            }
            used.assign((size_t)w, 0);
            auto &rc = cols[i];
            rc.reserve((size_t)cnt);
            while ((int64_t)rc.size() < cnt) {
                const int64_t off = (int64_t)rng.below((uint64_t)w);
                if (!used[(size_t)off]) {
                    used[(size_t)off] = 1;
                    rc.push_back((int32_t)(c0 + off));
                }
            }
            std::sort(rc.begin(), rc.end());
            auto &rv = vals[i];
            rv.resize(rc.size());
            for (auto &v : rv)
                v = rng.normal();
        }
    }
    (void)bad; // rows narrower than lo are clipped to the window
    return from_rows(M, M, cols, vals, out);
}

int hsmu_gen_random(int32_t M, int32_t N, int32_t max_per_row, uint64_t seed, hsmu_host_csr *out)
{
    if (!out || M < 0 || N < 0 || max_per_row < 0)
        return HSMU_ERR_INVALID;
    memset(out, 0, sizeof(*out));
    std::vector<std::vector<int32_t>> cols(M);
    std::vector<std::vector<double>> vals(M);
    Rng rng(derive_seed(seed, 0));
    for (int64_t i = 0; i < M; ++i) {
        int64_t cnt = (int64_t)rng.below((uint64_t)max_per_row + 1);
        if (cnt > N)
            cnt = N;
        auto &rc = cols[i];
        while ((int64_t)rc.size() < cnt) {
            const int32_t c = (int32_t)rng.below((uint64_t)N);
            if (std::find(rc.begin(), rc.end(), c) == rc.end())
                rc.push_back(c);
        }
        std::sort(rc.begin(), rc.end());
        vals[i].resize(rc.size());
        for (auto &v : vals[i])
            v = 2.0 * rng.uniform() - 1.0;
    }
    return from_rows(M, N, cols, vals, out);
}

int hsmu_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c, uint64_t seed,
                  hsmu_host_csr *out)
{
    if (!out || scale < 1 || scale > 30 || edge_factor < 1 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        return HSMU_ERR_INVALID;
    memset(out, 0, sizeof(*out));
    const int64_t n = (int64_t)1 << scale;
    const int64_t E = (int64_t)edge_factor * n;
    std::vector<uint64_t> key((size_t)E);
    const int64_t chunk = 1 << 20;
    const int64_t nchunks = (E + chunk - 1) / chunk;
    const double ab = a + b, abc = a + b + c;
#pragma omp parallel for schedule(dynamic)
    for (int64_t ch = 0; ch < nchunks; ++ch) {
        Rng rng(derive_seed(seed, (uint64_t)ch));
        for (int64_t e = ch * chunk; e < std::min(E, (ch + 1) * chunk); ++e) {
            uint64_t r = 0, q = 0;
            for (int lev = 0; lev < scale; ++lev) {
                const double u = rng.uniform();
                r <<= 1;
                q <<= 1;
                if (u < a) {
                } else if (u < ab) {
                    q |= 1;
                } else if (u < abc) {
                    r |= 1;
                } else {
                    r |= 1;
                    q |= 1;
                }
            }
            key[(size_t)e] = (r << 32) | q;
        }
    }
    std::sort(key.begin(), key.end());
    key.erase(std::unique(key.begin(), key.end()), key.end());
    const int64_t nnz = (int64_t)key.size();
    if (nnz > INT32_MAX)
        return HSMU_ERR_UNSUPPORTED;
    if (!alloc_host(out, n, n, nnz))
        return HSMU_ERR_NOMEM;
    std::vector<int64_t> cnt((size_t)n + 1, 0);
    for (uint64_t k : key)
        ++cnt[(size_t)(k >> 32) + 1];
    for (int64_t i = 0; i < n; ++i)
        cnt[(size_t)i + 1] += cnt[(size_t)i];
    for (int64_t i = 0; i <= n; ++i)
        out->rowptr[i] = (int32_t)cnt[(size_t)i];
    // values after merging, in sorted order, chunked for determinism
    const int64_t vchunks = (nnz + chunk - 1) / chunk;
#pragma omp parallel for schedule(dynamic)
    for (int64_t ch = 0; ch < vchunks; ++ch) {
        Rng rng(derive_seed(seed ^ 0x5bd1e995ull, (uint64_t)ch));
        for (int64_t k = ch * chunk; k < std::min(nnz, (ch + 1) * chunk); ++k) {
            out->col[k] = (int32_t)(key[(size_t)k] & 0xffffffffull);
            out->val[k] = rng.uniform();
        }
    }
    return HSMU_OK;
}

int hsmu_gen_zipf(int32_t M, int32_t N, int32_t per_row, double s, uint64_t seed, hsmu_host_csr *out)
{
    if (!out || M < 0 || N <= 0 || per_row < 0 || per_row > N || s <= 0)
        return HSMU_ERR_INVALID;
    memset(out, 0, sizeof(*out));
    // cumulative distribution of P(k) ~ 1 / (k+1)^s over columns k = 0..N-1
    std::vector<double> cdf((size_t)N);
    double acc = 0;
    for (int64_t k = 0; k < N; ++k) {
        acc += std::pow((double)(k + 1), -s);
        cdf[(size_t)k] = acc;
    }
    for (auto &v : cdf)
        v /= acc;
    const int64_t nnz = (int64_t)M * per_row;
    if (nnz > INT32_MAX)
        return HSMU_ERR_UNSUPPORTED;
    if (!alloc_host(out, M, N, nnz))
        return HSMU_ERR_NOMEM;
    const int64_t chunk = 8192;
    const int64_t nchunks = (M + chunk - 1) / chunk;
#pragma omp parallel for schedule(dynamic)
    for (int64_t ch = 0; ch < nchunks; ++ch) {
        Rng rng(derive_seed(seed, (uint64_t)ch));
        std::vector<int32_t> rc;
        for (int64_t i = ch * chunk; i < std::min<int64_t>(M, (ch + 1) * chunk); ++i) {
            rc.clear();
            while ((int32_t)rc.size() < per_row) {
                const double u = rng.uniform();
                int32_t k = (int32_t)(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
                if (k >= N)
                    k = N - 1;
                if (std::find(rc.begin(), rc.end(), k) == rc.end())
                    rc.push_back(k);
            }
            std::sort(rc.begin(), rc.end());
            out->rowptr[i] = (int32_t)(i * per_row);
            for (int32_t q = 0; q < per_row; ++q) {
                out->col[i * per_row + q] = rc[(size_t)q];
                out->val[i * per_row + q] = 2.0 * rng.uniform() - 1.0;
            }
        }
    }
    out->rowptr[M] = (int32_t)nnz;
    return HSMU_OK;
}

// CSR of A^T: a histogram of columns, a scan and a stable scatter in row order, so the
// rows of A^T come out sorted (reference: csr2csc, src/mmio_highlevel.h:194-229).
int hsmu_transpose(const hsmu_host_csr *A, hsmu_host_csr *AT)
{
    if (!A || !AT || A->M < 0 || A->N < 0)
        return HSMU_ERR_INVALID;
    memset(AT, 0, sizeof(*AT));
    if (!alloc_host(AT, A->N, A->M, A->nnz))
        return HSMU_ERR_NOMEM;
    std::vector<int64_t> cnt((size_t)A->N + 1, 0);
    for (int64_t k = 0; k < A->nnz; ++k) {
        const int32_t c = A->col[k];
        if (c < 0 || c >= A->N) {
            hsmu_host_csr_free(AT);
            return HSMU_ERR_INDEX;
        }
        ++cnt[(size_t)c + 1];
    }
    for (int64_t c = 0; c < A->N; ++c)
        cnt[(size_t)c + 1] += cnt[(size_t)c];
    for (int64_t c = 0; c <= A->N; ++c)
        AT->rowptr[c] = (int32_t)cnt[(size_t)c];
    for (int64_t i = 0; i < A->M; ++i) {
        for (int32_t k = A->rowptr[i]; k < A->rowptr[i + 1]; ++k) {
            const int64_t pos = cnt[(size_t)A->col[k]]++;
            AT->col[pos] = (int32_t)i;
            AT->val[pos] = A->val[k];
        }
    }
    return HSMU_OK;
}

// Matrix Market coordinate reader with the semantics of the reference's mmio_allinone
// (src/mmio_highlevel.h:18-193): real / integer / pattern / complex fields (pattern
// entries get 1.0, complex keeps the real part), symmetric and hermitian files are
// expanded by mirroring every off-diagonal entry, duplicates are kept.  Each row is then
// stable-sorted by column, which the reference leaves to the file.
static int read_mtx(const char *path, hsmu_host_csr *out, int32_t *is_symmetric)
{
    FILE *f = fopen(path, "rb");
    if (!f)
        return HSMU_ERR_INVALID;
    std::string text;
    {
        char buf[1 << 16];
        size_t r;
        while ((r = fread(buf, 1, sizeof(buf), f)) > 0)
            text.append(buf, r);
        fclose(f);
    }
    const char *p = text.c_str();
    const char *end = p + text.size();
    // banner
    const char *nl = (const char *)memchr(p, '\n', (size_t)(end - p));
    std::string banner(p, nl ? nl : end);
    for (auto &ch : banner)
        ch = (char)std::tolower((unsigned char)ch);
    if (banner.rfind("%%matrixmarket", 0) != 0 || banner.find("matrix") == std::string::npos ||
        banner.find("coordinate") == std::string::npos)
        return HSMU_ERR_UNSUPPORTED;
    const bool pattern = banner.find("pattern") != std::string::npos;
    const bool complex_ = banner.find("complex") != std::string::npos;
    const bool symmetric = (banner.find(" symmetric") != std::string::npos &&
                            banner.find("skew-symmetric") == std::string::npos) ||
                           banner.find("hermitian") != std::string::npos;
    if (is_symmetric)
        *is_symmetric = symmetric ? 1 : 0;
    p = nl ? nl + 1 : end;
    // comments
    while (p < end && *p == '%') {
        nl = (const char *)memchr(p, '\n', (size_t)(end - p));
        p = nl ? nl + 1 : end;
    }
    char *q;
    const long long m = strtoll(p, &q, 10);
    const long long n = strtoll(q, &q, 10);
    const long long nz = strtoll(q, &q, 10);
    p = q;
    if (m < 0 || n < 0 || nz < 0 || m >= INT32_MAX || n >= INT32_MAX)
        return HSMU_ERR_UNSUPPORTED;
    // every entry takes at least four bytes of the file ("1 1\n"): a larger count is a
    // truncated or corrupt file, not a reason to reserve memory for it
    if (nz > (long long)(end - p) / 4 + 1)
        return HSMU_ERR_INVALID;
    // mirrored entries land in row c - 1 < n, so a symmetric file must be square
    if (symmetric && m != n)
        return HSMU_ERR_INVALID;
    std::vector<int32_t> ri((size_t)nz), ci((size_t)nz);
    std::vector<double> vi((size_t)nz);
    for (long long k = 0; k < nz; ++k) {
        const long long r = strtoll(p, &q, 10);
        const long long c = strtoll(q, &q, 10);
        double v = 1.0;
        if (!pattern) {
            v = strtod(q, &q);
            if (complex_)
                (void)strtod(q, &q);
        }
        p = q;
        if (r < 1 || r > m || c < 1 || c > n)
            return HSMU_ERR_INDEX;
        ri[(size_t)k] = (int32_t)(r - 1);
        ci[(size_t)k] = (int32_t)(c - 1);
        vi[(size_t)k] = v;
    }
    std::vector<int64_t> cnt((size_t)m + 1, 0);
    for (long long k = 0; k < nz; ++k) {
        ++cnt[(size_t)ri[(size_t)k] + 1];
        if (symmetric && ri[(size_t)k] != ci[(size_t)k])
            ++cnt[(size_t)ci[(size_t)k] + 1];
    }
    for (long long i = 0; i < m; ++i)
        cnt[(size_t)i + 1] += cnt[(size_t)i];
    const int64_t total = cnt[(size_t)m];
    if (total > INT32_MAX)
        return HSMU_ERR_UNSUPPORTED;
    if (!alloc_host(out, m, n, total))
        return HSMU_ERR_NOMEM;
    for (long long i = 0; i <= m; ++i)
        out->rowptr[i] = (int32_t)cnt[(size_t)i];
    std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
    for (long long k = 0; k < nz; ++k) {
        const int32_t r = ri[(size_t)k], c = ci[(size_t)k];
        int64_t pos = fill[(size_t)r]++;
        out->col[pos] = c;
        out->val[pos] = vi[(size_t)k];
        if (symmetric && r != c) {
            pos = fill[(size_t)c]++;
            out->col[pos] = r;
            out->val[pos] = vi[(size_t)k];
        }
    }
    // stable sort of each row by column
    std::vector<std::pair<int32_t, double>> row;
    for (long long i = 0; i < m; ++i) {
        const int32_t a = out->rowptr[i], b = out->rowptr[i + 1];
        bool sorted = true;
        for (int32_t k = a + 1; k < b && sorted; ++k)
            sorted = out->col[k - 1] <= out->col[k];
        if (sorted)
            continue;
        row.clear();
        for (int32_t k = a; k < b; ++k)
            row.emplace_back(out->col[k], out->val[k]);
        std::stable_sort(row.begin(), row.end(),
                         [](const std::pair<int32_t, double> &x, const std::pair<int32_t, double> &y) {
                             return x.first < y.first;
                         });
        for (int32_t k = a; k < b; ++k) {
            out->col[k] = row[(size_t)(k - a)].first;
            out->val[k] = row[(size_t)(k - a)].second;
        }
    }
    return HSMU_OK;
}

int hsmu_read_mtx(const char *path, hsmu_host_csr *out, int32_t *is_symmetric)
{
    if (!path || !out)
        return HSMU_ERR_INVALID;
    memset(out, 0, sizeof(*out));
    int rc;
    try {
        rc = read_mtx(path, out, is_symmetric);
    } catch (const std::bad_alloc &) {
        rc = HSMU_ERR_NOMEM;
    } catch (...) {
        rc = HSMU_ERR_INVALID;
    }
    if (rc != HSMU_OK)
        hsmu_host_csr_free(out);
    return rc;
}

} // extern "C"
