/*
 * CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h).  Plain C99.
 *
 * Compiled with -O2 -ffp-contract=off so that "a += v*x" is a rounded product
 * followed by a rounded add, as the reference's serial executor computes it;
 * the fused variants call fma() explicitly.
 */
#include "oracle.h"

#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void orc_free(void *p) { free(p); }

/* ------------------------------------------------------------------------ */
/* Matrix Market banner + size line.  mmio.h:249-332 (banner: five tokens,  */
/* "%%MatrixMarket" prefix compared case-sensitively, the other four tokens */
/* lower-cased) and CUDA/ourWrtFst/mmio.h:334-365 (skip lines starting with */
/* '%', then the first other line must scan as "M N nz").                   */
/* ------------------------------------------------------------------------ */
typedef struct {
    int is_complex, is_real, is_pattern, is_integer, is_symmetric;
} orc_mm_type;

static void orc_lower(char *s) { for (; *s; ++s) *s = (char)tolower((unsigned char)*s); }

static int orc_read_banner(FILE *f, orc_mm_type *t)
{
    char line[1025], banner[64], mtx[64], crd[64], dtype[64], storage[64];
    memset(t, 0, sizeof(*t));
    if (!fgets(line, sizeof line, f)) return 12;
    if (sscanf(line, "%63s %63s %63s %63s %63s", banner, mtx, crd, dtype, storage) != 5) return 12;
    orc_lower(mtx); orc_lower(crd); orc_lower(dtype); orc_lower(storage);
    if (strncmp(banner, "%%MatrixMarket", 14) != 0) return 14;
    if (strcmp(mtx, "matrix") != 0) return 15;
    if (strcmp(crd, "coordinate") != 0 && strcmp(crd, "array") != 0) return 15;
    if (strcmp(dtype, "real") == 0) t->is_real = 1;
    else if (strcmp(dtype, "complex") == 0) t->is_complex = 1;
    else if (strcmp(dtype, "pattern") == 0) t->is_pattern = 1;
    else if (strcmp(dtype, "integer") == 0) t->is_integer = 1;
    else return 15;
    if (strcmp(storage, "general") == 0) { }
    else if (strcmp(storage, "symmetric") == 0 || strcmp(storage, "hermitian") == 0) t->is_symmetric = 1;
    else if (strcmp(storage, "skew-symmetric") == 0) { /* not mirrored: mm_is_symmetric is false */ }
    else return 15;
    return 0;
}

static int orc_read_size(FILE *f, int *m, int *n, int *nz)
{
    char line[1025];
    *m = *n = *nz = 0;
    do {
        if (!fgets(line, sizeof line, f)) return 12;
    } while (line[0] == '%');
    return sscanf(line, "%d %d %d", m, n, nz) == 3 ? 0 : -1;
}

/* read_mtx.h:6-191: COO entries in file order, 1-based -> 0-based (:98-99),
 * symmetric/hermitian off-diagonals mirrored (:110-117, :139-163), entries
 * placed into their row in file order (duplicates kept). */
#define ORC_DEFINE_READ(T, SFX)                                                      \
int orc_read_mtx_##SFX(const char *path, int *m_out, int *n_out, int *nnz_out,       \
                       int **rp_out, int **ci_out, T **v_out)                        \
{                                                                                    \
    orc_mm_type t;                                                                   \
    int m, n, nr, i;                                                                 \
    FILE *f = fopen(path, "r");                                                      \
    if (!f) return -1;                                                               \
    if (orc_read_banner(f, &t) != 0) { fclose(f); return -2; }                       \
    if (t.is_complex) { fclose(f); return -3; }                                      \
    if (orc_read_size(f, &m, &n, &nr) != 0) { fclose(f); return -4; }                \
    if (m < 0 || n < 0 || nr < 0) { fclose(f); return -4; }                          \
    int *cnt = (int *)calloc((size_t)m + 1, sizeof(int));                            \
    int *ri = (int *)malloc(sizeof(int) * (size_t)(nr ? nr : 1));                    \
    int *cj = (int *)malloc(sizeof(int) * (size_t)(nr ? nr : 1));                    \
    T *vv = (T *)malloc(sizeof(T) * (size_t)(nr ? nr : 1));                          \
    for (i = 0; i < nr; ++i) {                                                       \
        int a, b, iv, got; double fv = 0.0;                                          \
        if (t.is_real) got = fscanf(f, "%d %d %lg\n", &a, &b, &fv) == 3;             \
        else if (t.is_integer) { got = fscanf(f, "%d %d %d\n", &a, &b, &iv) == 3;   \
                                 fv = iv; }                                          \
        else { got = fscanf(f, "%d %d\n", &a, &b) == 2; fv = 1.0; }                  \
        if (!got || a < 1 || a > m || b < 1 || b > n) {                              \
            fclose(f); free(cnt); free(ri); free(cj); free(vv); return -5;           \
        }                                                                            \
        --a; --b;                                                                    \
        cnt[a]++;                                                                    \
        ri[i] = a; cj[i] = b; vv[i] = (T)fv;                                         \
    }                                                                                \
    fclose(f);                                                                       \
    if (t.is_symmetric) {                                                            \
        if (m != n) {                                                                \
            for (i = 0; i < nr; ++i)                                                 \
                if (ri[i] != cj[i] && cj[i] >= m) {                                  \
                    free(cnt); free(ri); free(cj); free(vv); return -5; }            \
        }                                                                            \
        for (i = 0; i < nr; ++i) if (ri[i] != cj[i]) cnt[cj[i]]++;                   \
    }                                                                                \
    int *rp = (int *)malloc(sizeof(int) * ((size_t)m + 1));                          \
    rp[0] = 0;                                                                       \
    for (i = 0; i < m; ++i) rp[i + 1] = rp[i] + cnt[i];                              \
    int nnz = rp[m];                                                                 \
    int *ci = (int *)malloc(sizeof(int) * (size_t)(nnz ? nnz : 1));                  \
    T *v = (T *)malloc(sizeof(T) * (size_t)(nnz ? nnz : 1));                         \
    memset(cnt, 0, sizeof(int) * ((size_t)m + 1));                                   \
    for (i = 0; i < nr; ++i) {                                                       \
        int r = ri[i], c = cj[i], o = rp[r] + cnt[r]++;                              \
        ci[o] = c; v[o] = vv[i];                                                     \
        if (t.is_symmetric && r != c) {                                              \
            o = rp[c] + cnt[c]++;                                                    \
            ci[o] = r; v[o] = vv[i];                                                 \
        }                                                                            \
    }                                                                                \
    free(cnt); free(ri); free(cj); free(vv);                                         \
    *m_out = m; *n_out = n; *nnz_out = nnz;                                          \
    *rp_out = rp; *ci_out = ci; *v_out = v;                                          \
    return 0;                                                                        \
}
ORC_DEFINE_READ(double, d)
ORC_DEFINE_READ(float, s)

/* ------------------------------------------------------------------------ */
/* change2tran: read_mtx.h:194-257.  Per row: for each position j, exchange */
/* with every later k holding a smaller column (:215-226); keep col < i     */
/* with value 1 (:228-232), stop at the first col >= i (:234-237), append   */
/* the unit diagonal (:240-242).                                            */
/* ------------------------------------------------------------------------ */
#define ORC_DEFINE_C2T(T, SFX)                                                       \
void orc_change2tran_##SFX(int m, int nnzA, int *rpA, int *ciA, T *vA,               \
                           int *nnzL, int **rpL_out, int **ciL_out, T **vL_out)      \
{                                                                                    \
    int *rpL = (int *)malloc(sizeof(int) * ((size_t)m + 1));                         \
    int *ciL = (int *)malloc(sizeof(int) * ((size_t)nnzA + (size_t)m + 1));          \
    T *vL = (T *)malloc(sizeof(T) * ((size_t)nnzA + (size_t)m + 1));                 \
    int p = 0, i, j, k;                                                              \
    rpL[0] = 0;                                                                      \
    for (i = 0; i < m; ++i) {                                                        \
        for (j = rpA[i]; j < rpA[i + 1]; ++j) {                                      \
            int c = ciA[j]; T cv = vA[j];                                            \
            for (k = j + 1; k < rpA[i + 1]; ++k) {                                   \
                if (ciA[k] < c) {                                                    \
                    ciA[j] = ciA[k]; vA[j] = vA[k];                                  \
                    ciA[k] = c; vA[k] = cv;                                          \
                    c = ciA[j]; cv = vA[j];                                          \
                }                                                                    \
            }                                                                        \
            if (ciA[j] < i) { ciL[p] = ciA[j]; vL[p] = (T)1; ++p; }                  \
            else break;                                                              \
        }                                                                            \
        ciL[p] = i; vL[p] = (T)1.0; ++p;                                             \
        rpL[i + 1] = p;                                                              \
    }                                                                                \
    *nnzL = p; *rpL_out = rpL; *ciL_out = ciL; *vL_out = vL;                         \
}
ORC_DEFINE_C2T(double, d)
ORC_DEFINE_C2T(float, s)

/* get_x_b: read_mtx.h:260-278 */
#define ORC_DEFINE_GXB(T, SFX)                                                       \
void orc_get_x_b_##SFX(int m, int n, const int *rp, const int *ci, const T *v,       \
                       T *x_ref, T *b)                                               \
{                                                                                    \
    int i, j;                                                                        \
    for (i = 0; i < n; ++i) x_ref[i] = (T)1;                                         \
    for (i = 0; i < m; ++i) {                                                        \
        b[i] = (T)0;                                                                 \
        for (j = rp[i]; j < rp[i + 1]; ++j) b[i] += v[j] * x_ref[ci[j]];             \
    }                                                                                \
}
ORC_DEFINE_GXB(double, d)
ORC_DEFINE_GXB(float, s)

/* matrix_layer: tranpose.h:78-142 */
void orc_row_levels(int m, const int *rp, const int *ci, int *level)
{
    int r, j;
    memset(level, 0, sizeof(int) * (size_t)m);
    for (r = 0; r < m; ++r) {
        int lv = 0;
        for (j = rp[r]; j < rp[r + 1]; ++j) {
            int c = ci[j];
            if (level[c] + 1 > lv) lv = level[c] + 1;
        }
        level[r] = lv;
    }
}

int orc_matrix_layer(int m, const int *rp, const int *ci, int *levels, double *parallelism)
{
    int *level = (int *)malloc(sizeof(int) * (size_t)(m ? m : 1));
    int r, mx = 0;
    orc_row_levels(m, rp, ci, level);
    for (r = 0; r < m; ++r) if (level[r] > mx) mx = level[r];
    free(level);
    *levels = mx;
    *parallelism = (double)m / mx;
    return mx;
}

/* matrix_transposition: tranpose.h:36-75 */
#define ORC_DEFINE_TR(T, SFX)                                                        \
void orc_transpose_##SFX(int m, int n, int nnz, const int *rp, const int *ci,        \
                         const T *v, int *csc_row, int *cp, T *cv)                   \
{                                                                                    \
    int i, r, j;                                                                     \
    memset(cp, 0, sizeof(int) * ((size_t)n + 1));                                    \
    for (i = 0; i < nnz; ++i) cp[ci[i] + 1]++;                                       \
    for (i = 0; i < n; ++i) cp[i + 1] += cp[i];                                      \
    int *fill = (int *)malloc(sizeof(int) * ((size_t)n + 1));                        \
    memcpy(fill, cp, sizeof(int) * ((size_t)n + 1));                                 \
    for (r = 0; r < m; ++r)                                                          \
        for (j = rp[r]; j < rp[r + 1]; ++j) {                                        \
            int o = fill[ci[j]]++;                                                   \
            csc_row[o] = r; cv[o] = v[j];                                            \
        }                                                                            \
    free(fill);                                                                      \
}
ORC_DEFINE_TR(double, d)
ORC_DEFINE_TR(float, s)

/* matrix_warp4: CUDA/mix/tranpose.h:1423-1477 */
void orc_matrix_warp4(int m, int nnz, const int *rp, int border, int chunk,
                      int *len, int *warp_num, double *warp_occupy, double *element_occupy)
{
    int i, j, k = 1, row = 0;
    double rows_in_wave_mode = 0, nnz_in_wave_mode = 0;
    warp_num[0] = 0;
    for (i = 0; i < m; i += chunk) {
        int end = i + chunk < m ? i + chunk : m;
        int e = rp[end] - rp[i];
        double avg = (double)e / (end - i);
        if (avg >= border) {
            nnz_in_wave_mode += e;
            rows_in_wave_mode += end - i;
            for (j = 0; j < end - i; ++j) warp_num[k++] = ++row;
        } else {
            row += end - i;
            warp_num[k++] = row;
        }
    }
    *len = k;
    *warp_occupy = rows_in_wave_mode / m;
    *element_occupy = nnz_in_wave_mode / nnz;
}

/* serial analyser: serialref.h:6-23 (histogram of row indices over all nnz) */
void orc_csc_analyser(const int *csc_row, int m, int nnz, int *indeg)
{
    int i;
    memset(indeg, 0, sizeof(int) * (size_t)m);
    for (i = 0; i < nnz; ++i) indeg[csc_row[i]]++;
}

/* serial executor: serialref.h:25-108.  The reference's wait loop (:51-55,
 * :80-84) never spins in a serial run because every column is visited after
 * all of its producers; the counter it compares is kept and checked here so
 * that an input violating that order is reported instead of silently solved. */
#define ORC_DEFINE_EXEC(T, SFX)                                                      \
void orc_csc_executor_##SFX(const int *cp, const int *ri, const T *v,                \
                            const int *indeg, int m, int n, int substitution,        \
                            int rhs, const T *b, T *x)                               \
{                                                                                    \
    int *seen = (int *)calloc((size_t)m + 1, sizeof(int));                           \
    T *ls = (T *)calloc((size_t)m * (size_t)rhs + 1, sizeof(T));                     \
    int i, j, k;                                                                     \
    if (substitution == ORC_SUBSTITUTION_FORWARD) {                                  \
        for (i = 0; i < n; ++i) {                                                    \
            if (indeg[i] - 1 != seen[i])                                             \
                fprintf(stderr, "oracle: column %d visited early\n", i);             \
            for (k = 0; k < rhs; ++k)                                                \
                x[(size_t)i * rhs + k] =                                             \
                    (b[(size_t)i * rhs + k] - ls[(size_t)i * rhs + k]) / v[cp[i]];   \
            for (j = cp[i] + 1; j < cp[i + 1]; ++j) {                                \
                int r = ri[j];                                                       \
                for (k = 0; k < rhs; ++k)                                            \
                    ls[(size_t)r * rhs + k] += x[(size_t)i * rhs + k] * v[j];        \
                seen[r]++;                                                           \
            }                                                                        \
        }                                                                            \
    } else {                                                                         \
        for (i = n - 1; i >= 0; --i) {                                               \
            if (indeg[i] - 1 != seen[i])                                             \
                fprintf(stderr, "oracle: column %d visited early\n", i);             \
            for (k = 0; k < rhs; ++k)                                                \
                x[(size_t)i * rhs + k] =                                             \
                    (b[(size_t)i * rhs + k] - ls[(size_t)i * rhs + k]) /             \
                    v[cp[i + 1] - 1];                                                \
            for (j = cp[i]; j < cp[i + 1] - 1; ++j) {                                \
                int r = ri[j];                                                       \
                for (k = 0; k < rhs; ++k)                                            \
                    ls[(size_t)r * rhs + k] += x[(size_t)i * rhs + k] * v[j];        \
                seen[r]++;                                                           \
            }                                                                        \
        }                                                                            \
    }                                                                                \
    free(seen); free(ls);                                                            \
}
ORC_DEFINE_EXEC(double, d)
ORC_DEFINE_EXEC(float, s)

/* Row-oriented restatement of the executor's FORWARD order (see oracle.h). */
#define ORC_DEFINE_FWD(T, SFX, FMA)                                                  \
void orc_csr_forward_##SFX(int n, const int *rp, const int *ci, const T *v,          \
                           int rhs, const T *b, T *x, int fused)                     \
{                                                                                    \
    int i, j, k;                                                                     \
    for (i = 0; i < n; ++i) {                                                        \
        int diag = rp[i + 1] - 1;                                                    \
        for (k = 0; k < rhs; ++k) {                                                  \
            T s = (T)0;                                                              \
            for (j = rp[i]; j < diag; ++j) {                                         \
                T xv = x[(size_t)ci[j] * rhs + k];                                   \
                if (fused) s = FMA(xv, v[j], s);                                     \
                else { T p = xv * v[j]; s = s + p; }                                 \
            }                                                                        \
            x[(size_t)i * rhs + k] = (b[(size_t)i * rhs + k] - s) / v[diag];         \
        }                                                                            \
    }                                                                                \
}
ORC_DEFINE_FWD(double, d, fma)
ORC_DEFINE_FWD(float, s, fmaf)

/* General CSR triangular solve in stored order (see oracle.h): lower rows
 * visited 0..n-1 with the diagonal last, upper rows n-1..0 with the diagonal
 * first; with unit = 1 no diagonal is stored and x = b - s. */
#define ORC_DEFINE_TRI(T, SFX, FMA)                                                  \
void orc_csr_solve_##SFX(int n, const int *rp, const int *ci, const T *v, int rhs,  \
                         const T *b, T *x, int upper, int unit, int fused)          \
{                                                                                    \
    int t, j, k;                                                                     \
    for (t = 0; t < n; ++t) {                                                        \
        int i = upper ? n - 1 - t : t;                                               \
        int lo = rp[i], hi = rp[i + 1], dg = -1;                                     \
        if (!unit) {                                                                 \
            if (upper) dg = lo++;                                                    \
            else dg = --hi;                                                          \
        }                                                                            \
        for (k = 0; k < rhs; ++k) {                                                  \
            T s = (T)0;                                                              \
            for (j = lo; j < hi; ++j) {                                              \
                T xv = x[(size_t)ci[j] * rhs + k];                                   \
                if (fused) s = FMA(xv, v[j], s);                                     \
                else { T p = xv * v[j]; s = s + p; }                                 \
            }                                                                        \
            T r = b[(size_t)i * rhs + k] - s;                                        \
            x[(size_t)i * rhs + k] = unit ? r : r / v[dg];                           \
        }                                                                            \
    }                                                                                \
}
ORC_DEFINE_TRI(double, d, fma)
ORC_DEFINE_TRI(float, s, fmaf)

/* Adjoint (see oracle.h): executor BACKWARD on the CSR of L read as CSC of L^T */
void orc_adjoint_d(int n, const int *rp, const int *ci, const double *v,
                   const double *g, double *y, int fused)
{
    double *ls = (double *)calloc((size_t)n + 1, sizeof(double));
    int i, j;
    for (i = n - 1; i >= 0; --i) {
        double yi = (g[i] - ls[i]) / v[rp[i + 1] - 1];
        y[i] = yi;
        for (j = rp[i]; j < rp[i + 1] - 1; ++j) {
            if (fused) ls[ci[j]] = fma(yi, v[j], ls[ci[j]]);
            else { double p = yi * v[j]; ls[ci[j]] = ls[ci[j]] + p; }
        }
    }
    free(ls);
}

void orc_grad_val_d(int n, const int *rp, const int *ci, const double *y,
                    const double *x, double *gv)
{
    int i, j;
    for (i = 0; i < n; ++i)
        for (j = rp[i]; j < rp[i + 1]; ++j) gv[j] = -y[i] * x[ci[j]];
}

void orc_residual_d(int n, const int *rp, const int *ci, const double *v, int rhs,
                    const double *b, const double *x, double *r)
{
    int i, j, k;
    for (i = 0; i < n; ++i)
        for (k = 0; k < rhs; ++k) {
            double s = 0.0;
            for (j = rp[i]; j < rp[i + 1]; ++j) s += v[j] * x[(size_t)ci[j] * rhs + k];
            r[(size_t)i * rhs + k] = b[(size_t)i * rhs + k] - s;
        }
}

/* FP32 fused forward solve in stored order whose rows end with a product by
 * the rounded reciprocal of the diagonal instead of a division. */
static void orc_forward_rcp_s(int n, const int *rp, const int *ci, const float *v, int rhs,
                              const float *b, float *x)
{
    int i, j, k;
    for (i = 0; i < n; ++i) {
        int diag = rp[i + 1] - 1;
        float r = 1.0f / v[diag];
        for (k = 0; k < rhs; ++k) {
            float s = 0.0f;
            for (j = rp[i]; j < diag; ++j) s = fmaf(x[(size_t)ci[j] * rhs + k], v[j], s);
            x[(size_t)i * rhs + k] = (b[(size_t)i * rhs + k] - s) * r;
        }
    }
}

void orc_refine2_d(int n, const int *rp, const int *ci, const double *v, int rhs,
                   const double *b, float *x0, float *d, double *x, int rcp)
{
    size_t nnz = (size_t)rp[n], len = (size_t)n * rhs, t;
    float *v32 = (float *)calloc(nnz + 1, sizeof(float));
    float *b32 = (float *)calloc(len + 1, sizeof(float));
    int i, j, k;
    for (t = 0; t < nnz; ++t) v32[t] = (float)v[t];
    for (t = 0; t < len; ++t) b32[t] = (float)b[t];
    if (rcp) orc_forward_rcp_s(n, rp, ci, v32, rhs, b32, x0);
    else orc_csr_forward_s(n, rp, ci, v32, rhs, b32, x0, 1);
    for (i = 0; i < n; ++i)
        for (k = 0; k < rhs; ++k) {
            double s = 0.0;
            for (j = rp[i]; j < rp[i + 1]; ++j) s = fma(v[j], (double)x0[(size_t)ci[j] * rhs + k], s);
            b32[(size_t)i * rhs + k] = (float)(b[(size_t)i * rhs + k] - s);
        }
    if (rcp) orc_forward_rcp_s(n, rp, ci, v32, rhs, b32, d);
    else orc_csr_forward_s(n, rp, ci, v32, rhs, b32, d, 1);
    for (t = 0; t < len; ++t) x[t] = (double)x0[t] + (double)d[t];
    free(v32);
    free(b32);
}

/* spts_syncfree_cuda_csr_wrt.h:208 (+ 4*Len, CUDA/mix/...:317) */
double orc_datasize(int m, int n, int nnzL, int vb, int len)
{
    return (double)(n + 1) * 4 + (double)(nnzL + m) * 4 + (double)nnzL * vb +
           2.0 * n * vb + (double)len * 4;
}

double orc_gflops(int nnzL, double ms) { return 2.0 * nnzL / (1e6 * ms); }
