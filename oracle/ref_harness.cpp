/*
 * Harness around the reference's own plain-C host helpers -- TEST
 * INFRASTRUCTURE ONLY.  Nothing here is copied: the Makefile points the
 * include paths at the reference tree, and this file only adds extern "C"
 * entry points so the tests can run those helpers through ctypes and compare
 * them with the oracle's restatement.
 *
 * Sources compiled (relative to the reference tree):
 *   CUDA/ourWrtFst/mmio.h                      mm_read_banner, strict mm_read_mtx_crd_size
 *   openCL/mix/common.h                        VALUE_TYPE / WARP_SIZE macros (no CUDA headers)
 *   openCL/mix/read_mtx.h                      read_mtx, change2tran, get_x_b (identical in every CUDA dir)
 *   openCL/mix/tranpose.h                      matrix_transposition, matrix_layer, matrix_warp4
 *   openCL/SynFree_csc/sptrsv_syncfree_serialref.h   serial analyser + executor
 */
#include REF_MMIO
#include "common.h"
#include "read_mtx.h"
#include "tranpose.h"
#include REF_SERIALREF

extern "C" {

int ref_value_bytes(void) { return (int)sizeof(VALUE_TYPE); }
int ref_warp_size(void) { return WARP_SIZE; }

int ref_read_mtx(const char *path, int *m, int *n, int *nnz, int **rp, int **ci, VALUE_TYPE **v)
{
    return read_mtx((char *)path, m, n, nnz, rp, ci, v);
}

void ref_change2tran(int m, int nnzA, int *rpA, int *ciA, VALUE_TYPE *vA,
                     int *nnzL, int **rpL, int **ciL, VALUE_TYPE **vL)
{
    change2tran(m, nnzA, rpA, ciA, vA, nnzL, rpL, ciL, vL);
}

void ref_get_x_b(int m, int n, const int *rp, const int *ci, const VALUE_TYPE *v,
                 VALUE_TYPE *x_out, VALUE_TYPE *b_out)
{
    VALUE_TYPE *x = 0, *b = 0;
    get_x_b(m, n, rp, ci, v, &x, &b);
    memcpy(x_out, x, sizeof(VALUE_TYPE) * n);
    memcpy(b_out, b, sizeof(VALUE_TYPE) * m);
    free(x);
    free(b);
}

int ref_matrix_layer(int m, int n, int nnz, const int *rp, const int *ci, const VALUE_TYPE *v,
                     int *layer, double *parallelism)
{
    return matrix_layer(m, n, nnz, rp, ci, v, layer, parallelism);
}

void ref_matrix_transposition(int m, int n, int nnz, const int *rp, const int *ci,
                              const VALUE_TYPE *v, int *csc_row, int *csc_colptr, VALUE_TYPE *csc_val)
{
    matrix_transposition(m, n, nnz, rp, ci, v, csc_row, csc_colptr, csc_val);
}

void ref_matrix_warp4(int m, int n, int nnz, const int *rp, const int *ci, const VALUE_TYPE *v,
                      int border, int *len, int *warp_num, double *wocc, double *eocc)
{
    matrix_warp4(m, n, nnz, rp, ci, v, border, len, warp_num, wocc, eocc);
}

void ref_serial_analyser(const int *csc_row, int m, int n, int nnz, int *indeg)
{
    sptrsv_syncfree_analyser(csc_row, m, n, nnz, indeg);
}

void ref_serial_executor(const int *cp, const int *ri, const VALUE_TYPE *v, const int *indeg,
                         int m, int n, int substitution, int rhs, const VALUE_TYPE *b, VALUE_TYPE *x)
{
    sptrsv_syncfree_executor(cp, ri, v, indeg, m, n, substitution, rhs, b, x);
}

void ref_free(void *p) { free(p); }

}
