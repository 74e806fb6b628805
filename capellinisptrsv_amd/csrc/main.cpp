// Drop-in driver: ./sptrsv_main A.mtx [-a wrtfst|2part|mix|csc|auto|local|partition] [--values]
//
// Same flow and output as the reference's drivers (CUDA/ourWrtFst/main.cu:9-115,
// CUDA/mix/main.cu:104-106, CUDA/SyncFree_csc/main.cu:92-134): read the
// matrix, truncate to the leading square block, extract the unit lower
// triangle (change2tran), count levels, run the solve wrapper twice (the
// first run is a warm-up) and append one row to result.csv in the current
// directory:
//   filename, m, nnzL, nnzL/m, layer, parallelism, pre_time, solve_time, gflops, bandwith
// (mix adds warp_occupy, element_occupy).  Precision is chosen at compile
// time with -DVALUE_TYPE=float|double, as in the reference's common.h.
// --values keeps the file's values (diagonal 1 + sum|offdiag|) instead of
// the reference's all-ones L.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/capellini_sptrsv.h"

namespace {

int read_mtx(const char *f, int *m, int *n, int *nnz, int **rp, int **ci, VALUE_TYPE **v)
{
    if (sizeof(VALUE_TYPE) == 8) return cap_read_mtx_d(f, m, n, nnz, rp, ci, (double **)(void *)v);
    return cap_read_mtx_s(f, m, n, nnz, rp, ci, (float **)(void *)v);
}
int change2tran(int m, int nnzA, const int *rp, const int *ci, const VALUE_TYPE *v, int keep, int *nnzL, int **rpL,
                int **ciL, VALUE_TYPE **vL)
{
    if (sizeof(VALUE_TYPE) == 8)
        return cap_change2tran_d(m, nnzA, rp, ci, (const double *)(const void *)v, keep, nnzL, rpL, ciL,
                                 (double **)(void *)vL);
    return cap_change2tran_s(m, nnzA, rp, ci, (const float *)(const void *)v, keep, nnzL, rpL, ciL,
                             (float **)(void *)vL);
}

}  // namespace

int main(int argc, char **argv)
{
    FILE *fp = std::fopen("result.csv", "a");
    if (!fp) return -1;
    const char *filename = nullptr;
    std::string algo = "wrtfst";
    int keep_values = 0;
    for (int a = 1; a < argc; ++a) {
        if (!std::strcmp(argv[a], "-a") && a + 1 < argc) algo = argv[++a];
        else if (!std::strcmp(argv[a], "--values")) keep_values = 1;
        else filename = argv[a];
    }
    if (!filename) {
        std::fprintf(stderr, "usage: %s A.mtx [-a wrtfst|2part|mix|csc|auto|local|partition] [--values]\n", argv[0]);
        std::fclose(fp);
        return -1;
    }
    int m = 0, n = 0, nnzA = 0;
    int *rpA = nullptr, *ciA = nullptr;
    VALUE_TYPE *vA = nullptr;
    int rc = read_mtx(filename, &m, &n, &nnzA, &rpA, &ciA, &vA);
    if (rc != 0) {
        std::fprintf(stderr, "read_mtx(%s) failed: %d\n", filename, rc);
        std::fclose(fp);
        return rc;
    }
    // leading square block (main.cu:66-71)
    if (m <= n) n = m;
    else m = n;
    if (m <= 1) {
        std::fclose(fp);
        return 0;
    }
    int nnzL = 0, *rpL = nullptr, *ciL = nullptr;
    VALUE_TYPE *vL = nullptr;
    rc = change2tran(m, nnzA, rpA, ciA, vA, keep_values, &nnzL, &rpL, &ciL, &vL);
    cap_free(rpA);
    cap_free(ciA);
    cap_free(vA);
    if (rc != 0) {
        std::fclose(fp);
        return rc;
    }
    if (m == 0 || nnzL == 0) {
        std::fclose(fp);
        return -3;
    }
    int layer = 0;
    double parallelism = 0, solve_time = 0, gflops = 0, bandwith = 0, pre_time = 0, wocc = 0, eocc = 0;
    cap_matrix_layer(m, rpL, ciL, &layer, &parallelism, nullptr);

    if (algo == "mix") {
        for (int r = 0; r < 2 && rc == 0; ++r)
            rc = spts_syncfree_hip_mix(rpL, ciL, vL, m, n, nnzL, &solve_time, &gflops, &bandwith, &pre_time, &wocc,
                                       &eocc);
    } else if (algo == "2part") {
        for (int r = 0; r < 2 && rc == 0; ++r)
            rc = spts_syncfree_hip_csr(rpL, ciL, vL, m, n, nnzL, &solve_time, &gflops, &bandwith);
    } else if (algo == "csc") {
        std::vector<int> cr((size_t)nnzL), cp((size_t)n + 1);
        std::vector<VALUE_TYPE> cv((size_t)nnzL), x((size_t)n), xr((size_t)n), b((size_t)m);
        if (sizeof(VALUE_TYPE) == 8) {
            cap_transpose_d(m, n, nnzL, rpL, ciL, (const double *)(const void *)vL, cr.data(), cp.data(),
                            (double *)(void *)cv.data());
            cap_get_x_b_d(m, n, rpL, ciL, (const double *)(const void *)vL, (double *)(void *)xr.data(),
                          (double *)(void *)b.data());
        } else {
            cap_transpose_s(m, n, nnzL, rpL, ciL, (const float *)(const void *)vL, cr.data(), cp.data(),
                            (float *)(void *)cv.data());
            cap_get_x_b_s(m, n, rpL, ciL, (const float *)(const void *)vL, (float *)(void *)xr.data(),
                          (float *)(void *)b.data());
        }
        for (int r = 0; r < 2 && rc == 0; ++r)
            rc = sptrsv_syncfree_hip_csc(cp.data(), cr.data(), cv.data(), m, n, nnzL, x.data(), b.data(), xr.data(),
                                         &pre_time, &solve_time, &gflops, &bandwith);
    } else if (algo == "wrtfst") {
        for (int r = 0; r < 2 && rc == 0; ++r)
            rc = spts_syncfree_hip_csr_wrt(rpL, ciL, vL, m, n, nnzL, &solve_time, &gflops, &bandwith);
    } else if (algo == "auto" || algo == "local" || algo == "partition") {
        // this library's autotuner and deep-matrix solvers, same columns
        const int a = algo == "auto" ? CAP_ALGO_AUTO : algo == "local" ? CAP_ALGO_LOCAL : CAP_ALGO_PARTITION;
        for (int r = 0; r < 2 && rc == 0; ++r)
            rc = spts_syncfree_hip_csr_algo(rpL, ciL, vL, m, n, nnzL, a, &solve_time, &gflops, &bandwith);
    } else {
        std::fprintf(stderr, "unknown algorithm %s\n", algo.c_str());
        rc = -2;
    }
    if (rc != 0) {
        std::fprintf(stderr, "solve failed: %d (%s)\n", rc, cap_strerror(rc));
    } else if (algo == "mix") {
        std::fprintf(fp, "filename,m,nnzL,nnzL/m,layer,parallelism,pre_time,solve_time,gflops,bandwith,warp_occupy,"
                         "element_occupy\n");
        std::fprintf(fp, "%s, %d, %d, %4.2f, %d, %4.2f, %4.2f, %4.2f, %4.2f, %4.2f, %4.2f, %4.2f\n", filename, m,
                     nnzL, (double)nnzL / m, layer, parallelism, pre_time, solve_time, gflops, bandwith, wocc, eocc);
    } else {
        std::fprintf(fp, "filename,m,nnzL,nnzL/m,layer,parallelism,pre_time,solve_time,gflops,bandwith\n");
        std::fprintf(fp, "%s, %d, %d, %4.2f, %d, %4.2f, %4.2f, %4.2f, %4.2f, %4.2f\n", filename, m, nnzL,
                     (double)nnzL / m, layer, parallelism, pre_time, solve_time, gflops, bandwith);
    }
    cap_free(rpL);
    cap_free(ciL);
    cap_free(vL);
    std::fclose(fp);
    return rc;
}
