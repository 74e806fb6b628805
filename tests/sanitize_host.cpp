// Drives every host-side preprocessing routine of the library (the C ABI
// in capellinisptrsv_amd/csrc/host_prep.cpp) under AddressSanitizer and
// UndefinedBehaviorSanitizer: the Matrix Market reader on each fixture and
// on malformed and oversized inputs, change2tran, get_x_b, the level
// analysis, the mix chunk table, CSR->CSC, layout validation and the
// generators.  Built and run by tests/test_host.py; any sanitizer report
// ends the run with a non-zero status.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../include/capellini_sptrsv.h"

static int failures = 0;
#define EXPECT(c)                                                              \
    do {                                                                       \
        if (!(c)) {                                                            \
            std::fprintf(stderr, "FAILED %s at %s:%d\n", #c, __FILE__, __LINE__); \
            ++failures;                                                        \
        }                                                                      \
    } while (0)

static void write_file(const std::string &path, const char *text)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    std::fputs(text, f);
    std::fclose(f);
}

// change2tran -> get_x_b -> levels -> warp4 -> transpose -> validate on one matrix
template <typename T>
static void pipeline(int m, int n, int nnz, const int *rp, const int *ci, const T *v);

template <>
void pipeline<double>(int m, int n, int nnz, const int *rp, const int *ci, const double *v)
{
    const int mm = m < n ? m : n;
    for (int keep = 0; keep < 2; ++keep) {
        int nnzL = 0, *rpL = nullptr, *ciL = nullptr;
        double *vL = nullptr;
        EXPECT(cap_change2tran_d(mm, rp[mm], rp, ci, v, keep, &nnzL, &rpL, &ciL, &vL) == CAP_OK);
        std::vector<double> x((size_t)mm + 1), b((size_t)mm + 1);
        cap_get_x_b_d(mm, mm, rpL, ciL, vL, x.data(), b.data());
        int levels = 0;
        double par = 0;
        std::vector<int> lv((size_t)mm + 1);
        cap_matrix_layer(mm, rpL, ciL, &levels, &par, lv.data());
        std::vector<int> wn((size_t)mm + 2);
        int len = 0;
        double wo = 0, eo = 0;
        for (int chunk : {32, 64})
            for (int border : {1, 4, 10}) cap_matrix_warp4(mm, nnzL, rpL, border, chunk, &len, wn.data(), &wo, &eo);
        std::vector<int> cr((size_t)nnzL + 1), cp((size_t)mm + 2);
        std::vector<double> cv((size_t)nnzL + 1);
        cap_transpose_d(mm, mm, nnzL, rpL, ciL, vL, cr.data(), cp.data(), cv.data());
        int bad = -1;
        EXPECT(cap_validate_csr_d(mm, nnzL, rpL, ciL, vL, &bad) == CAP_OK);
        // serial solves: forward on the CSC, backward on the CSR read as CSC, 1 and 3 RHS
        std::vector<int> deg((size_t)mm + 1);
        EXPECT(cap_serial_analyser(cr.data(), mm, mm, nnzL, deg.data()) == CAP_OK);
        for (int rhs : {1, 3}) {
            std::vector<double> bk((size_t)mm * rhs + 1, 1.0), xk((size_t)mm * rhs + 1), xs((size_t)mm * rhs + 1);
            EXPECT(cap_serial_executor_d(cp.data(), cr.data(), cv.data(), deg.data(), mm, mm,
                                         CAP_SUBSTITUTION_FORWARD, rhs, bk.data(), xk.data()) == CAP_OK);
            EXPECT(cap_serial_csr_d(mm, nnzL, rpL, ciL, vL, rhs, bk.data(), xs.data(), CAP_ARITH_REFERENCE) == CAP_OK);
            EXPECT(cap_serial_csr_d(mm, nnzL, rpL, ciL, vL, rhs, bk.data(), xs.data(), CAP_ARITH_FUSED) == CAP_OK);
            std::vector<int> dr((size_t)mm + 1);
            EXPECT(cap_serial_analyser(ciL, mm, mm, nnzL, dr.data()) == CAP_OK);
            EXPECT(cap_serial_executor_d(rpL, ciL, vL, dr.data(), mm, mm, CAP_SUBSTITUTION_BACKWARD, rhs, bk.data(),
                                         xk.data()) == CAP_OK);
        }
        EXPECT(sptrsv_syncfree_serialref_d(cp.data(), cr.data(), cv.data(), mm, mm, nnzL, CAP_SUBSTITUTION_FORWARD,
                                           1, x.data(), b.data(), x.data()) == CAP_OK);
        cap_free(rpL);
        cap_free(ciL);
        cap_free(vL);
    }
    (void)nnz;
}

template <>
void pipeline<float>(int m, int n, int nnz, const int *rp, const int *ci, const float *v)
{
    const int mm = m < n ? m : n;
    int nnzL = 0, *rpL = nullptr, *ciL = nullptr;
    float *vL = nullptr;
    EXPECT(cap_change2tran_s(mm, rp[mm], rp, ci, v, 1, &nnzL, &rpL, &ciL, &vL) == CAP_OK);
    std::vector<float> x((size_t)mm + 1), b((size_t)mm + 1);
    cap_get_x_b_s(mm, mm, rpL, ciL, vL, x.data(), b.data());
    std::vector<int> cr((size_t)nnzL + 1), cp((size_t)mm + 2);
    std::vector<float> cv((size_t)nnzL + 1);
    cap_transpose_s(mm, mm, nnzL, rpL, ciL, vL, cr.data(), cp.data(), cv.data());
    int bad = -1;
    EXPECT(cap_validate_csr_s(mm, nnzL, rpL, ciL, vL, &bad) == CAP_OK);
    cap_free(rpL);
    cap_free(ciL);
    cap_free(vL);
    (void)nnz;
}

int main(int argc, char **argv)
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s FIXTURE_DIR SCRATCH_DIR\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1], tmp = argv[2];
    // every fixture, both precisions
    const char *names[] = {"MixedCase_banner.mtx", "err_banner.mtx", "err_complex.mtx", "err_size_blank.mtx",
                           "general_real.mtx", "integer_symmetric.mtx", "pattern_general.mtx", "rect_wide.mtx",
                           "skew.mtx", "symmetric_real.mtx"};
    for (const char *nm : names) {
        const std::string path = dir + "/" + nm;
        int m, n, nnz, *rp = nullptr, *ci = nullptr;
        double *v = nullptr;
        if (cap_read_mtx_d(path.c_str(), &m, &n, &nnz, &rp, &ci, &v) == 0) {
            pipeline<double>(m, n, nnz, rp, ci, v);
            cap_free(rp);
            cap_free(ci);
            cap_free(v);
        }
        float *vf = nullptr;
        if (cap_read_mtx_s(path.c_str(), &m, &n, &nnz, &rp, &ci, &vf) == 0) {
            pipeline<float>(m, n, nnz, rp, ci, vf);
            cap_free(rp);
            cap_free(ci);
            cap_free(vf);
        }
    }
    // malformed and hostile inputs: each must be refused without a bad access
    const char *bad[] = {
        "",
        "%%MatrixMarket matrix coordinate real general\n",
        "%%MatrixMarket matrix coordinate real general\n3 3 2000000000\n1 1 1.0\n",
        "%%MatrixMarket matrix coordinate pattern symmetric\n3 3 2000000000\n2 1\n3 1\n",
        "%%MatrixMarket matrix coordinate real general\n3 3 1\n1 4294967297 1.0\n",
        "%%MatrixMarket matrix coordinate real general\n3 3 1\n-1 1 1.0\n",
        "%%MatrixMarket matrix coordinate real general\n3 3 2\n1 1 1.0\n",
        "%%MatrixMarket matrix coordinate real general\n-3 3 1\n1 1 1.0\n",
        "%%MatrixMarket matrix coordinate integer general\n2 2 1\n1 1 99999999999\n",
        "%%MatrixMarket matrix coordinate real symmetric\n2 3 1\n1 3 1.0\n",
        "%%MatrixMarket matrix coordinate real general\n3 3 1\n1 1 nan\n",
    };
    int idx = 0;
    for (const char *text : bad) {
        const std::string path = tmp + "/bad" + std::to_string(idx++) + ".mtx";
        write_file(path, text);
        int m, n, nnz, *rp = nullptr, *ci = nullptr;
        double *v = nullptr;
        const int rc = cap_read_mtx_d(path.c_str(), &m, &n, &nnz, &rp, &ci, &v);
        if (rc == 0) {  // "nan" is a value strtod accepts: read, then preprocessed
            pipeline<double>(m, n, nnz, rp, ci, v);
            cap_free(rp);
            cap_free(ci);
            cap_free(v);
        }
    }
    // layout violations through validation
    {
        const int rp[] = {0, 1, 3, 2}, ci[] = {0, 0, 1};
        const double v[] = {1, 1, 0};
        int badrow = -1;
        EXPECT(cap_validate_csr_d(3, 3, rp, ci, v, &badrow) != CAP_OK);
        const int rp2[] = {0, 1, 3, 5}, ci2[] = {0, 5, 1, -7, 2};
        const double v2[] = {1, 1, 1, 1, 1};
        EXPECT(cap_validate_csr_d(3, 5, rp2, ci2, v2, &badrow) != CAP_OK);
        EXPECT(cap_validate_csr_d(0, 0, rp, ci, v, &badrow) == CAP_OK);
        // serial executor on CSC factors that would make the reference spin or
        // write out of range: a row index past m, a column without its
        // diagonal first, an entry above the diagonal, a wrong in-degree table
        const int cp3[] = {0, 2, 3, 4};
        const int ri_far[] = {0, 9, 1, 2}, ri_nodiag[] = {1, 0, 1, 2}, ri_above[] = {0, 1, 0, 2};
        const double v3[] = {2, 1, 2, 2};
        const double b3[] = {1, 1, 1};
        double x3[3];
        const int deg3[] = {1, 2, 1};
        EXPECT(cap_serial_executor_d(cp3, ri_far, v3, deg3, 3, 3, CAP_SUBSTITUTION_FORWARD, 1, b3, x3) != CAP_OK);
        EXPECT(cap_serial_executor_d(cp3, ri_nodiag, v3, deg3, 3, 3, CAP_SUBSTITUTION_FORWARD, 1, b3, x3) != CAP_OK);
        EXPECT(cap_serial_executor_d(cp3, ri_above, v3, deg3, 3, 3, CAP_SUBSTITUTION_FORWARD, 1, b3, x3) != CAP_OK);
        const int ri_ok[] = {0, 1, 1, 2}, deg_bad[] = {1, 1, 1};
        EXPECT(cap_serial_executor_d(cp3, ri_ok, v3, deg_bad, 3, 3, CAP_SUBSTITUTION_FORWARD, 1, b3, x3) != CAP_OK);
        EXPECT(cap_serial_executor_d(cp3, ri_ok, v3, deg3, 3, 3, CAP_SUBSTITUTION_FORWARD, 1, b3, x3) == CAP_OK);
        int deg_out[3];
        EXPECT(cap_serial_analyser(ri_far, 3, 3, 4, deg_out) != CAP_OK);
        EXPECT(cap_serial_csr_d(3, 3, rp2, ci2, v2, 1, b3, x3, CAP_ARITH_FUSED) != CAP_OK);
    }
    // generators: every kind, small sizes, both passes, both precisions
    for (int kind = 0; kind < 4; ++kind) {
        cap_gen_params g;
        std::memset(&g, 0, sizeof g);
        g.kind = kind;
        g.offdiag = 5;
        g.band = 40;
        g.p = 0.3;
        g.seed = 7;
        if (kind == 1) {
            g.levels = 20;
            g.width = 50;
            g.depth = 3;
            g.n = g.levels * g.width;
        } else {
            g.n = 1000;
        }
        int nnz = 0;
        std::vector<int> rp((size_t)g.n + 1);
        EXPECT(cap_generate_d(&g, &nnz, rp.data(), nullptr, nullptr) == CAP_OK);
        std::vector<int> ci((size_t)nnz);
        std::vector<double> v((size_t)nnz);
        std::vector<float> vf((size_t)nnz);
        EXPECT(cap_generate_d(&g, &nnz, rp.data(), ci.data(), v.data()) == CAP_OK);
        EXPECT(cap_generate_s(&g, &nnz, rp.data(), ci.data(), vf.data()) == CAP_OK);
        int badrow = -1;
        EXPECT(cap_validate_csr_d(g.n, nnz, rp.data(), ci.data(), v.data(), &badrow) == CAP_OK);
        int levels = 0;
        double par = 0;
        cap_matrix_layer(g.n, rp.data(), ci.data(), &levels, &par, nullptr);
        std::vector<double> u(17);
        cap_uniform_d(g.seed, (long long)u.size(), -1, 1, u.data());
        (void)cap_datasize(g.n, g.n, nnz, 8, 3);
    }
    for (int code = 1; code >= -14; --code) EXPECT(cap_strerror(code) != nullptr);
    std::printf("host sanitizer run: %d failure(s)\n", failures);
    return failures ? 1 : 0;
}
