// Drop-in driver: ./sptrsv_main A.mtx [-a wrtfst|2part|mix|csc|auto|local|partition|level|serial] [--values]
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
// the reference's all-ones L.  -a serial times the reference's serial CPU
// executor on the CSC of L (no GPU call) into the same columns.  Each run
// also appends one JSON object to result.jsonl (FP64 serial-solve error,
// microseconds per level, roofline figures); --no-sidecar skips it.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/capellini_sptrsv.h"

#define CAP_TRY_MAIN(expr)             \
    do {                               \
        int rc_ = (expr);              \
        if (rc_ != 0) return rc_;      \
    } while (0)

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

int serial_executor(const int *cp, const int *cr, const VALUE_TYPE *cv, const int *deg, int m, int n,
                    const VALUE_TYPE *b, VALUE_TYPE *x)
{
    if (sizeof(VALUE_TYPE) == 8)
        return cap_serial_executor_d(cp, cr, (const double *)(const void *)cv, deg, m, n, CAP_SUBSTITUTION_FORWARD, 1,
                                     (const double *)(const void *)b, (double *)(void *)x);
    return cap_serial_executor_s(cp, cr, (const float *)(const void *)cv, deg, m, n, CAP_SUBSTITUTION_FORWARD, 1,
                                 (const float *)(const void *)b, (float *)(void *)x);
}

double ms_between(std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// The reference's serial CPU path (openCL/SynFree_csc/sptrsv_syncfree_serialref.h:110-177)
// with the GPU wrappers' outputs: pre_time = analyser ms, solve_time = executor
// ms (the second of two runs), gflops and bandwidth from the same formulas, and
// solve_time = -1 when x misses x_ref by the relative L1 bar of 1e-4.
int serial_solve(const int *cp, const int *cr, const VALUE_TYPE *cv, int m, int n, int nnz, VALUE_TYPE *x,
                 const VALUE_TYPE *b, const VALUE_TYPE *xr, double *pre_time, double *solve_time, double *gflops,
                 double *bandwith)
{
    std::vector<int> deg((size_t)m);
    auto t0 = std::chrono::steady_clock::now();
    int rc = cap_serial_analyser(cr, m, n, nnz, deg.data());
    *pre_time = ms_between(t0, std::chrono::steady_clock::now());
    for (int r = 0; r < 2 && rc == 0; ++r) {
        t0 = std::chrono::steady_clock::now();
        rc = serial_executor(cp, cr, cv, deg.data(), m, n, b, x);
        *solve_time = ms_between(t0, std::chrono::steady_clock::now());
    }
    if (rc != 0) return rc;
    *gflops = 2.0 * nnz / (*solve_time * 1e6);
    *bandwith = cap_datasize(m, n, nnz, (int)sizeof(VALUE_TYPE), 0) / (*solve_time * 1e6);
    double ref = 0, res = 0;
    for (int i = 0; i < n; ++i) {
        ref += std::fabs((double)xr[i]);
        res += std::fabs((double)x[i] - (double)xr[i]);
    }
    if (ref != 0) res /= ref;
    if (!(res < 1e-4)) {
        std::printf("SpTRSV Serial executor _NOT_ passed! |x-xref|/|xref| = %8.2e\n", res);
        *solve_time = -1;
        *gflops = *bandwith = 0;
    }
    return 0;
}

int algo_code(const std::string &a)
{
    if (a == "wrtfst") return CAP_ALGO_WRTFST;
    if (a == "2part") return CAP_ALGO_TWOPART;
    if (a == "mix") return CAP_ALGO_MIX;
    if (a == "auto") return CAP_ALGO_AUTO;
    if (a == "local") return CAP_ALGO_LOCAL;
    return CAP_ALGO_PARTITION;
}

// One solve of L x = b through the pure solve entry with the CSR wrappers'
// options (the algorithm, the value protocol, host inputs).
int solve_once(const std::string &algo, int n, int nnz, const int *rp, const int *ci, const VALUE_TYPE *v,
               const VALUE_TYPE *b, VALUE_TYPE *x)
{
    cap_opts o;
    cap_opts_default(&o);
    o.algo = algo_code(algo);
    o.sync = CAP_SYNC_VALUE;
    o.host_inputs = 1;
    if (sizeof(VALUE_TYPE) == 8)
        return sptrsv_csr_solve_d(n, nnz, rp, ci, (const double *)(const void *)v, (const double *)(const void *)b,
                                  (double *)(void *)x, &o);
    return sptrsv_csr_solve_s(n, nnz, rp, ci, (const float *)(const void *)v, (const float *)(const void *)b,
                              (float *)(void *)x, &o);
}

void json_string(FILE *f, const char *s)
{
    std::fputc('"', f);
    for (; *s; ++s) {
        const unsigned char c = (unsigned char)*s;
        if (c == '"' || c == '\\') std::fprintf(f, "\\%c", c);
        else if (c < 0x20) std::fprintf(f, "\\u%04x", c);
        else std::fputc(c, f);
    }
    std::fputc('"', f);
}

struct Run {
    const char *filename;
    std::string algo;
    int m, nnzL, layer;
    double parallelism, pre_time, solve_time, gflops, bandwith;
};

// One JSON object per run, appended to result.jsonl beside result.csv (whose
// columns stay the reference's): the CSV fields, the relative max-norm error
// of x against the FP64 serial solve of the same L and b (rounded product and
// sum per term, the reference executor's bits), the microseconds per level,
// and the roofline figures of the GPU wrappers (dataSize bytes, achieved
// GB/s against MI355X's 8 TB/s).  x is the timed solve's where the wrapper
// returns it (csc, serial), else one more solve with the wrapper's options.
int write_sidecar(const Run &r, const int *rp, const int *ci, const VALUE_TYPE *v, std::vector<VALUE_TYPE> *x,
                  bool timed_x)
{
    const int n = r.m;
    std::vector<VALUE_TYPE> xr((size_t)n), b((size_t)n);
    if (sizeof(VALUE_TYPE) == 8)
        cap_get_x_b_d(n, n, rp, ci, (const double *)(const void *)v, (double *)(void *)xr.data(),
                      (double *)(void *)b.data());
    else
        cap_get_x_b_s(n, n, rp, ci, (const float *)(const void *)v, (float *)(void *)xr.data(),
                      (float *)(void *)b.data());
    if (!timed_x) {
        x->assign((size_t)n, (VALUE_TYPE)0);
        CAP_TRY_MAIN(solve_once(r.algo, n, r.nnzL, rp, ci, v, b.data(), x->data()));
    }
    std::vector<double> v64(v, v + r.nnzL), b64(b.begin(), b.end()), x64((size_t)n);
    CAP_TRY_MAIN(cap_serial_csr_d(n, r.nnzL, rp, ci, v64.data(), 1, b64.data(), x64.data(), CAP_ARITH_REFERENCE));
    double num = 0, den = 0;
    bool same = sizeof(VALUE_TYPE) == 8;
    for (int i = 0; i < n; ++i) {
        const double xi = (double)(*x)[(size_t)i];
        num = std::max(num, std::fabs(xi - x64[(size_t)i]));
        den = std::max(den, std::fabs(x64[(size_t)i]));
        same = same && std::memcmp(&xi, &x64[(size_t)i], sizeof(double)) == 0;
    }
    FILE *f = std::fopen("result.jsonl", "a");
    if (!f) return CAP_ERR_IO;
    const bool gpu = r.algo != "serial";
    const bool passed = r.solve_time > 0;
    std::fprintf(f, "{\"filename\": ");
    json_string(f, r.filename);
    std::fprintf(f, ", \"algo\": ");
    json_string(f, r.algo.c_str());
    std::fprintf(f,
                 ", \"value_type\": \"%s\", \"device\": \"%s\", \"m\": %d, \"nnzL\": %d, \"layer\": %d, "
                 "\"parallelism\": %.17g, \"pre_time_ms\": %.17g, \"solve_time_ms\": %.17g, \"gflops\": %.17g, "
                 "\"bandwidth_gbs\": %.17g, \"passed\": %s, \"us_per_level\": %.17g, "
                 "\"rel_err_vs_fp64_serial\": %.17g, \"bit_identical_to_fp64_serial\": %s, \"x_from\": \"%s\"",
                 sizeof(VALUE_TYPE) == 8 ? "double" : "float", gpu ? "gpu" : "cpu", r.m, r.nnzL, r.layer,
                 r.parallelism, r.pre_time, r.solve_time, r.gflops, r.bandwith, passed ? "true" : "false",
                 passed ? 1e3 * r.solve_time / r.layer : -1.0, den > 0 ? num / den : num, same ? "true" : "false",
                 timed_x ? "timed solve" : "separate solve, same options");
    if (gpu && passed) {
        const double bytes = r.bandwith * 1e6 * r.solve_time;
        std::fprintf(f,
                     ", \"roofline\": {\"bound\": \"hbm\", \"bytes\": %.17g, \"achieved_gbs\": %.17g, "
                     "\"peak_gbs\": 8000, \"frac\": %.17g}}\n",
                     bytes, r.bandwith, r.bandwith / 8000.0);
    } else {
        std::fprintf(f, ", \"roofline\": null}\n");
    }
    std::fclose(f);
    return 0;
}

}  // namespace

int main(int argc, char **argv)
{
    FILE *fp = std::fopen("result.csv", "a");
    if (!fp) return -1;
    const char *filename = nullptr;
    std::string algo = "wrtfst";
    int keep_values = 0;
    bool sidecar = true;
    for (int a = 1; a < argc; ++a) {
        if (!std::strcmp(argv[a], "-a") && a + 1 < argc) algo = argv[++a];
        else if (!std::strcmp(argv[a], "--values")) keep_values = 1;
        else if (!std::strcmp(argv[a], "--no-sidecar")) sidecar = false;
        else filename = argv[a];
    }
    if (!filename) {
        std::fprintf(stderr, "usage: %s A.mtx [-a wrtfst|2part|mix|csc|auto|local|partition|level|serial] [--values] [--no-sidecar]\n", argv[0]);
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
    std::vector<VALUE_TYPE> x_sol;  // x of the timed solve, where the wrapper returns it
    bool timed_x = false;

    if (algo == "mix") {
        for (int r = 0; r < 2 && rc == 0; ++r)
            rc = spts_syncfree_hip_mix(rpL, ciL, vL, m, n, nnzL, &solve_time, &gflops, &bandwith, &pre_time, &wocc,
                                       &eocc);
    } else if (algo == "2part") {
        for (int r = 0; r < 2 && rc == 0; ++r)
            rc = spts_syncfree_hip_csr(rpL, ciL, vL, m, n, nnzL, &solve_time, &gflops, &bandwith);
    } else if (algo == "csc" || algo == "serial") {
        std::vector<int> cr((size_t)nnzL), cp((size_t)n + 1);
        std::vector<VALUE_TYPE> cv((size_t)nnzL), x((size_t)n), xr((size_t)n), b((size_t)m);
        timed_x = true;
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
        if (algo == "csc") {
            for (int r = 0; r < 2 && rc == 0; ++r)
                rc = sptrsv_syncfree_hip_csc(cp.data(), cr.data(), cv.data(), m, n, nnzL, x.data(), b.data(),
                                             xr.data(), &pre_time, &solve_time, &gflops, &bandwith);
        } else {
            rc = serial_solve(cp.data(), cr.data(), cv.data(), m, n, nnzL, x.data(), b.data(), xr.data(), &pre_time,
                              &solve_time, &gflops, &bandwith);
        }
        x_sol.swap(x);
    } else if (algo == "wrtfst") {
        for (int r = 0; r < 2 && rc == 0; ++r)
            rc = spts_syncfree_hip_csr_wrt(rpL, ciL, vL, m, n, nnzL, &solve_time, &gflops, &bandwith);
    } else if (algo == "auto" || algo == "local" || algo == "partition" || algo == "level") {
        // this library's autotuner and deep-matrix solvers, same columns
        const int a = algo == "auto"    ? CAP_ALGO_AUTO
                      : algo == "local" ? CAP_ALGO_LOCAL
                      : algo == "level" ? CAP_ALGO_LEVEL
                                        : CAP_ALGO_PARTITION;
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
    if (rc == 0 && sidecar) {
        Run run{filename, algo, m, nnzL, layer, parallelism, pre_time, solve_time, gflops, bandwith};
        rc = write_sidecar(run, rpL, ciL, vL, &x_sol, timed_x);
        if (rc != 0) std::fprintf(stderr, "result.jsonl: %d (%s)\n", rc, cap_strerror(rc));
    }
    cap_free(rpL);
    cap_free(ciL);
    cap_free(vL);
    std::fclose(fp);
    return rc;
}
