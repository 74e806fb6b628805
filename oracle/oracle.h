/*
 * CPU oracle for the CapelliniSpTRSV hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C restatement of the reference's host-side algorithms and of
 * its serial solver.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it, and only as the checker; the library never links
 * against it and never falls back to it.
 *
 * Every function cites the reference lines it restates (paths relative to the
 * reference tree).  Each function is pinned by at least one of: the compiled
 * reference helpers (oracle/_ref), committed golden vectors (tests/golden),
 * closed forms (x = 1 on unit-valued L, level counts of diagonal / bidiagonal
 * matrices), dense triangular solves (scipy) and brute force on tiny inputs.
 * The pin of each function is listed in tests/test_oracle.py.
 */
#ifndef CAPELLINI_ORACLE_H
#define CAPELLINI_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_SUBSTITUTION_FORWARD 0
#define ORC_SUBSTITUTION_BACKWARD 1

/* ---- Matrix Market reader: CUDA/ourWrtFst/read_mtx.h:6-191 with the strict
 *      size line of CUDA/ourWrtFst/mmio.h:334-365 and the banner parser
 *      mmio.h:249-332.  Returns 0, or the reference's codes: -1 open failed,
 *      -2 banner, -3 complex, -4 size line.  -5 = malformed entry line (the
 *      reference leaves this undefined: fscanf's return is unchecked).
 *      Arrays are malloc'ed; free with orc_free. */
int orc_read_mtx_d(const char *path, int *m, int *n, int *nnz,
                   int **row_ptr, int **col_idx, double **val);
int orc_read_mtx_s(const char *path, int *m, int *n, int *nnz,
                   int **row_ptr, int **col_idx, float **val);

/* ---- change2tran: CUDA/ourWrtFst/read_mtx.h:194-257.  Sorts each row of A
 *      in place with the reference's exchange sort (stopping at the first
 *      column >= i), keeps the strictly lower entries with value 1 and
 *      appends a unit diagonal.  Output arrays are malloc'ed. */
void orc_change2tran_d(int m, int nnzA, int *rpA, int *ciA, double *vA,
                       int *nnzL, int **rpL, int **ciL, double **vL);
void orc_change2tran_s(int m, int nnzA, int *rpA, int *ciA, float *vA,
                       int *nnzL, int **rpL, int **ciL, float **vL);

/* ---- get_x_b: read_mtx.h:260-278.  x_ref = 1, b = A * x_ref in CSR order. */
void orc_get_x_b_d(int m, int n, const int *rp, const int *ci, const double *v,
                   double *x_ref, double *b);
void orc_get_x_b_s(int m, int n, const int *rp, const int *ci, const float *v,
                   float *x_ref, float *b);

/* ---- matrix_layer: CUDA/ourWrtFst/tranpose.h:78-142.  1-based levels (the
 *      diagonal entry contributes level[row]+1 with level[row] still 0).
 *      Returns the number of levels; *parallelism = m / levels. */
int orc_matrix_layer(int m, const int *rp, const int *ci, int *levels,
                     double *parallelism);
/* per-row level (same recurrence), for the level-set tests */
void orc_row_levels(int m, const int *rp, const int *ci, int *level);

/* ---- matrix_transposition: CUDA/ourWrtFst/tranpose.h:36-75 (CSR -> CSC,
 *      rows ascending inside every column). */
void orc_transpose_d(int m, int n, int nnz, const int *rp, const int *ci,
                     const double *v, int *csc_row, int *csc_colptr,
                     double *csc_val);
void orc_transpose_s(int m, int n, int nnz, const int *rp, const int *ci,
                     const float *v, int *csc_row, int *csc_colptr,
                     float *csc_val);

/* ---- matrix_warp4: CUDA/mix/tranpose.h:1423-1477 with the chunk width as a
 *      parameter (the reference's WARP_SIZE: 32 in CUDA/, 64 in openCL/).
 *      warp_num must hold m+1 ints. */
void orc_matrix_warp4(int m, int nnz, const int *rp, int border, int chunk,
                      int *len, int *warp_num, double *warp_occupy,
                      double *element_occupy);

/* ---- serial sync-free analyser + executor:
 *      openCL/SynFree_csc/sptrsv_syncfree_serialref.h:6-23 and :25-108.
 *      Column-driven; FORWARD takes the diagonal from the first entry of each
 *      column (:59), BACKWARD from the last (:88); multi-RHS interleaved as
 *      x[i*rhs+k] (:57-60); divides by the diagonal. */
void orc_csc_analyser(const int *csc_row, int m, int nnz, int *indeg);
void orc_csc_executor_d(const int *cp, const int *ri, const double *v,
                        const int *indeg, int m, int n, int substitution,
                        int rhs, const double *b, double *x);
void orc_csc_executor_s(const int *cp, const int *ri, const float *v,
                        const int *indeg, int m, int n, int substitution,
                        int rhs, const float *b, float *x);

/* ---- Row-oriented CSR forward substitution with the executor's accumulation
 *      order (serialref.h:63-68 adds column i's contribution to every row in
 *      ascending i, i.e. ascending column order per row) and its division by
 *      the diagonal (:59).  Diagonal = last entry of each row (the layout
 *      change2tran produces).  fused = 0: left_sum += v*x as a rounded product
 *      then a rounded add (the executor's arithmetic); fused = 1: a single
 *      fma(v, x, left_sum) per term.  Multi-RHS interleaved like the executor. */
void orc_csr_forward_d(int n, const int *rp, const int *ci, const double *v,
                       int rhs, const double *b, double *x, int fused);
void orc_csr_forward_s(int n, const int *rp, const int *ci, const float *v,
                       int rhs, const float *b, float *x, int fused);

/* Lower or upper CSR triangular solve in stored order, the upper-triangular
 * and implicit-unit-diagonal generalisation of orc_csr_forward (the
 * reference's SUBSTITUTION_BACKWARD mode, serialref.h:73-102, is its
 * column-driven form).  Lower: diagonal last, rows 0..n-1.  Upper: diagonal
 * first, rows n-1..0.  unit = 1: no diagonal stored, x_i = b_i - s. */
void orc_csr_solve_d(int n, const int *rp, const int *ci, const double *v, int rhs,
                     const double *b, double *x, int upper, int unit, int fused);
void orc_csr_solve_s(int n, const int *rp, const int *ci, const float *v, int rhs,
                     const float *b, float *x, int upper, int unit, int fused);

/* ---- Adjoint: y = L^{-T} g computed column-driven on the CSR of L, i.e. the
 *      executor's BACKWARD mode (serialref.h:73-102) reading row i of L as
 *      column i of L^T (diagonal last).  Then grad_val[k] = -y[row_k]*x[col_k]
 *      for every stored entry k. */
void orc_adjoint_d(int n, const int *rp, const int *ci, const double *v,
                   const double *g, double *y, int fused);
void orc_grad_val_d(int n, const int *rp, const int *ci, const double *y,
                    const double *x, double *grad_val);

/* ---- residual r = b - L x in FP64 (the refinement's SpMV), CSR order */
void orc_residual_d(int n, const int *rp, const int *ci, const double *v,
                    int rhs, const double *b, const double *x, double *r);

/* ---- two steps of mixed-precision refinement (no counterpart in the
 *      reference; the method of the library's refined solves):
 *      x0 = L32^{-1} b32 in FP32 (fused, stored order), r = b - L x0 in FP64
 *      (one fma chain over the row in stored order, diagonal included) rounded
 *      to FP32, d = L32^{-1} r32 in FP32, x = (double)x0 + (double)d.  L32 and
 *      b32 are the FP32 roundings of L and b.  Multi-RHS interleaved.
 *      rcp = 0: each FP32 row ends with a division by its diagonal;
 *      rcp = 1: with a product by the rounded reciprocal 1.0f / diag (the
 *      fused sweep's form; the correction absorbs the extra rounding). */
void orc_refine2_d(int n, const int *rp, const int *ci, const double *v, int rhs,
                   const double *b, float *x0, float *d, double *x, int rcp);

/* ---- metric formulas: CUDA/ourWrtFst/spts_syncfree_cuda_csr_wrt.h:206-215
 *      (dataSize; mix adds 4*Len, CUDA/mix/spts_syncfree_cuda_csr_wrt.h:317). */
double orc_datasize(int m, int n, int nnzL, int value_bytes, int len_warp_num);
double orc_gflops(int nnzL, double ms);

void orc_free(void *p);

#ifdef __cplusplus
}
#endif
#endif
