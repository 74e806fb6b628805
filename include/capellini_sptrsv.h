/*
 * capellini_sptrsv.h -- C ABI of the MI355X (gfx950) sparse triangular solve.
 *
 * Lower-triangular CSR solve L x = b with thread-level synchronization-free
 * kernels written for 64-lane wavefronts.  Every symbol is extern "C", takes
 * plain pointers and sizes, and carries no HIP or torch types: a stream is a
 * void* holding a hipStream_t (NULL = the null stream).
 *
 * Matrix layout expected by every solve entry (the layout the reference's
 * change2tran produces, CUDA/ourWrtFst/read_mtx.h:194-257): square n x n CSR
 * with int32 row_ptr[n+1] / col_idx[nnz]; every column of a row is < the
 * row except the LAST entry, which is the diagonal; the diagonal is nonzero.
 * Off-diagonals are accumulated in stored order (ascending after
 * change2tran); duplicates are allowed, as the reference keeps them.  Inputs violating this are
 * rejected with a negative code before anything is launched (the reference
 * would hang the GPU or return NaN).
 *
 * Precision follows the reference's compile-time switch
 * (CUDA/ourWrtFst/common.h:16-18): define VALUE_TYPE to float or double
 * (default double) before including this header and the unsuffixed drop-in
 * names resolve to the _s / _d entry points.  One shared library exports both.
 */
#ifndef CAPELLINI_SPTRSV_H
#define CAPELLINI_SPTRSV_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ codes */
#define CAP_OK 0
#define CAP_ERR_NOT_SQUARE (-1)   /* m != n: the reference's only error code */
#define CAP_ERR_ARG (-2)          /* NULL pointer, negative size, nnz mismatch */
#define CAP_ERR_ROWPTR (-3)       /* row_ptr not monotone or not ending at nnz */
#define CAP_ERR_UPPER (-4)        /* an off-diagonal entry with col >= row */
#define CAP_ERR_NO_DIAG (-5)      /* last entry of a row is not its diagonal */
#define CAP_ERR_ZERO_DIAG (-6)    /* diagonal value is zero */
#define CAP_ERR_HIP (-7)          /* a HIP runtime call failed */
#define CAP_ERR_NO_DEVICE (-8)    /* no usable gfx950 device */
#define CAP_ERR_ALLOC (-9)        /* host or device allocation failed */
#define CAP_ERR_UNSUPPORTED (-10) /* option combination not available */
#define CAP_ERR_IO (-11)          /* Matrix Market file could not be read */
#define CAP_ERR_NOT_CONVERGED (-12) /* refinement did not reach its tolerance */
#define CAP_ERR_WATCHDOG (-13)    /* a dependency wait exceeded spin_limit polls
                                     or wait_limit_ms; the rows involved hold NaN */

/* value types */
#define CAP_F64 0
#define CAP_F32 1

/* solve algorithms (cap_opts.algo) */
#define CAP_ALGO_AUTO 0    /* autotuner's choice (cached per matrix shape) */
#define CAP_ALGO_WRTFST 1  /* thread-per-row write-first (ourWrtFst) */
#define CAP_ALGO_TWOPART 2 /* thread-per-row, blocking for earlier waves (our2Part) */
#define CAP_ALGO_MIX 3     /* per-64-row chunk: thread or wavefront per row (mix) */
#define CAP_ALGO_CSC 4     /* wavefront-per-column in-degree scatter (SyncFree_csc) */
#define CAP_ALGO_LOCAL 5   /* one workgroup per matrix, wavefront per row, dependencies
                              resolved inside one compute unit: deep matrices
                              (few rows per level); value protocol on x itself */
#define CAP_ALGO_PARTITION 6 /* lower bandwidth <= 256: row blocks solved
                              independently, joined by a scan over the block
                              tails of W = 128 values (bandwidth <= 128) or
                              256 (FP64 tables built at plan creation;
                              FP32 plans widen b and round x once); sums are
                              reassociated (within rounding of the sequential
                              solve, not bit-exact).  Up to 64
                              right-hand sides per solve (its work vectors
                              grow on the first solve with a larger count,
                              which waits for the device); the refined solve
                              uses FP64 partitioned solves (no FP32 tables)
                              with one FP64 correction. */
#define CAP_ALGO_LEVEL 7     /* one workgroup walks the levels: a level's rows
                              in parallel (thread per row, steps of 64 to 256
                              rows, the workgroup size chosen by the plan),
                              a workgroup barrier between steps, values in an
                              LDS ring; no polling.  Needs every dependency
                              within a ring of <= 64 KiB of the newest row of
                              its level (CAP_ERR_UNSUPPORTED otherwise, also
                              when the padded steps exceed 4 slots per row);
                              1 right-hand side per solve, or 2 or 4 when
                              rows hold at most 4 off-diagonals and the wider
                              ring fits LDS; bit-exact (the
                              thread-per-row sum order).  Timed by algo AUTO
                              with the few-rows-per-level candidates; its
                              refined solves use the write-first kernels. */

/* adjoint method (cap_opts.adjoint) */
#define CAP_ADJOINT_AUTO 0
#define CAP_ADJOINT_SCATTER 1 /* column-driven on the CSR of L: in-degree counters
                                 and atomic left sums (SyncFree_csc BACKWARD) */
#define CAP_ADJOINT_GATHER 2  /* write-first solve on a reversed transpose of L
                                 built once per plan; deterministic */

/* readiness array of the value protocol (cap_opts.readiness).
 * CAP_READY_STAGED polls a plan-owned array in uncached device memory and
 * also stores each result into x; the array comes from a per-process pool
 * that is never freed (uncached memory read by a kernel stays allocated
 * until the process ends) and is reused by later plans.
 * CAP_READY_IN_PLACE polls x itself. */
#define CAP_READY_AUTO 0
#define CAP_READY_STAGED 1
#define CAP_READY_IN_PLACE 2

/* accumulation arithmetic (cap_opts.arith).  Both add a row's terms in its
 * stored order; they differ in the rounding of each term.
 * CAP_ARITH_FUSED: s = fma(v, x, s), one rounding per term (the default).
 * CAP_ARITH_REFERENCE: s = s + v*x with the product rounded, then the sum --
 * the arithmetic of the reference's serial CPU executor
 * (openCL/SynFree_csc/sptrsv_syncfree_serialref.h:67-68, built with g++ -O3
 * and no FMA contraction), so results are bit-identical to it.  Available
 * for the thread-per-row, two-part, mix and one-workgroup algorithms (and
 * the batch and gather adjoint built on them); CSC (atomic, order set by
 * scheduling) and PARTITION (reassociated) return CAP_ERR_UNSUPPORTED, and
 * the autotuner leaves them out. */
#define CAP_ARITH_FUSED 0
#define CAP_ARITH_REFERENCE 1

/* readiness protocol (cap_opts.sync) */
#define CAP_SYNC_AUTO 0
#define CAP_SYNC_VALUE 1 /* the value is the flag: x (or the staged readiness
                            array) starts as a signalling-NaN sentinel;
                            relaxed agent-scope 64/32-bit atomics */
#define CAP_SYNC_FLAG 2  /* separate int flags (the solve's epoch): the value is
                            stored first (a returning agent-scope exchange),
                            then the flag; readers load the value after the
                            flag */

typedef struct cap_opts {
    int algo;         /* CAP_ALGO_* (0 = auto) */
    int sync;         /* CAP_SYNC_* (0 = auto) */
    int border;       /* mix: nnz/row threshold for wavefront mode; 0 = from
                         the matrix's rows per level (4 where levels hold at
                         least 200 rows, else 32), timed among 4, 16 and 32
                         by algo AUTO */
    int validate;     /* 1 = check the CSR layout before the first solve */
    int host_inputs;  /* 1 = row_ptr/col_idx/val are host pointers (copied) */
    int spin_limit;   /* polls without progress before a wait gives up
                         (0 = no count bound, the default; the wall-time
                         bound below still applies) */
    int blocks_per_cu; /* persistent grid size in workgroups per compute unit
                          (0 = tuned / default 1; at most 4) */
    int adjoint;       /* CAP_ADJOINT_* (0 = gather) */
    int lds_window;    /* 1 = thread-per-row kernels resolve dependencies on
                          rows of their own workgroup in LDS (value protocol) */
    int readiness;     /* CAP_READY_* (value protocol): where readiness is
                          polled; 0 = tuned / default staged */
    int upper;         /* 1 = the matrix is upper triangular, diagonal first
                          in each row (backward substitution) */
    int unit_diag;     /* 1 = implicit unit diagonal: no diagonal is stored */
    int row_threads;   /* threads per workgroup of the thread-per-row solve
                          and the fused refinement sweep: 128 or 256 rows
                          polling per compute unit at one RHS
                          (0 = tuned / default 256) */
    int arith;         /* CAP_ARITH_* (0 = fused) */
    int count_polls;   /* 1 = the thread-per-row solve (value protocol) runs its
                          counting build: every poll load is counted, read
                          with cap_plan_poll_count (evidence, not speed) */
    int wait_limit_ms; /* wall time without progress after which a wait gives
                          up, publishes NaN and sets the watchdog code
                          (0 = default 10000 ms, < 0 = none, at most 40000):
                          every kernel drains in bounded time, whatever a
                          poll costs under memory contention */
    int rhs_per_lane;  /* algo WRTFST, value protocol, no LDS window, no poll
                          counting: solves of k > 1 right-hand sides run in
                          1024-thread workgroups (one per compute unit) with
                          this many right-hand sides per lane: 1, 2, 4, or 8
                          for FP32 (0 = the 128/256-thread solve above, or
                          the autotuned RHS block with algo AUTO).  Fewer
                          lanes per row put more rows in each workgroup
                          ticket, whose atomics serialize on one counter. */
    int row_tiles;     /* 1 = the thread-per-row solves (algo WRTFST at 128
                          or 256 threads, algo MIX; value protocol) first
                          copy each ticket's col_idx/val range (up to 2048
                          entries, 4096 for MIX) into LDS with coalesced
                          loads, and their rows read their entries there;
                          0 = off, or timed by algo AUTO.  Not with
                          count_polls, rhs_per_lane > 1 or batches
                          (CAP_ERR_UNSUPPORTED). */
} cap_opts;

/* Fills *o with the defaults (auto algorithm, auto sync, validate on). */
void cap_opts_default(cap_opts *o);

/* ------------------------------------------------------ plan (device) API
 * A plan binds one matrix resident in device memory.  With host_inputs = 0
 * the plan keeps the caller's device pointers (they must outlive the plan);
 * with host_inputs = 1 it copies them to the device.  Analysis (validation,
 * the mix chunk table, the CSC in-degree) runs once at creation.
 * cap_plan_solve is asynchronous on `stream`; b and x are device pointers of
 * n*rhs values interleaved as x[i*rhs + k] (serialref.h:57-60 layout). */
typedef struct cap_plan_s *cap_plan;

int cap_plan_create(cap_plan *plan, int n, int nnz, const int *row_ptr,
                    const int *col_idx, const void *val, int value_type,
                    const cap_opts *opts, void *stream);
int cap_plan_solve(cap_plan plan, const void *b, void *x, int rhs, void *stream);
/* cap_plan_solve = cap_plan_reset + cap_plan_launch.  The reset (tickets;
 * the sentinel fill of x in the value protocol, the in-degree restore for
 * CAP_ALGO_CSC) may be issued early to overlap other work; the launch is the
 * solve kernel alone.  Both asynchronous on `stream`, same b/x/rhs. */
int cap_plan_reset(cap_plan plan, void *x, int rhs, void *stream);
int cap_plan_launch(cap_plan plan, const void *b, void *x, int rhs, void *stream);
/* Synchronizes `stream`, then reads and clears the plan's watchdog word:
 * CAP_OK, or CAP_ERR_WATCHDOG if any wait gave up since the last call. */
int cap_plan_status(cap_plan plan, void *stream);
/* No counterpart in the reference, whose d_while_profiler
 * (CUDA/SyncFree_csc/spts_syncfree_cuda2.h:180-184) was allocated but never
 * written: synchronizes `stream`, returns the number of poll loads the
 * plan's thread-per-row solves issued since the last call (plans created
 * with cap_opts.count_polls = 1; CAP_ERR_UNSUPPORTED otherwise) and clears
 * it.  Divided by rows x right-hand sides it is the spin count per row. */
int cap_plan_poll_count(cap_plan plan, unsigned long long *polls, void *stream);
/* The RHS block an autotuned plan (algo AUTO, blocks_per_cu and row_threads
 * left 0) picked for k > 1 right-hand sides.  The first cap_plan_solve with
 * a new k times the candidates on its own b and x and waits for the device:
 * thread-per-row plans 1/2/4 workgroups per CU at 128/256 threads, and
 * 1024-thread workgroups with 1, 2, 4 or (FP32) 8 right-hand sides per
 * lane (*rhs_per_lane; 0 for the others), partitioned plans 1/2/4 right-hand
 * sides per wavefront of the block solve (*rhs_per_wave; 0 for the others).
 * The thread-per-row pick is then also timed with readiness in x itself
 * (*readiness: CAP_READY_STAGED or CAP_READY_IN_PLACE), which saves the
 * staged copy's stores; the 1024-thread shape with one right-hand side per
 * lane is timed with and without row tiles (*row_tiles, see cap_opts).
 * Every candidate gives the same bits.
 * CAP_ERR_UNSUPPORTED until a solve with that k has run.  (The reference's
 * OPT_WARP_RHS / OPT_WARP_AUTO switches, openCL/WrtFst/common.h:28-30, set
 * this by hand at compile time.) */
int cap_plan_rhs_block(cap_plan plan, int k, int *blocks_per_cu,
                       int *row_threads, int *rhs_per_wave, int *rhs_per_lane,
                       int *readiness, int *row_tiles);
/* Which algorithm / sync protocol the plan resolved to, and its mix table.
 * (An autotuned plan that resolved to mix, two-part, CSC or the level
 * steps, which solve one right-hand side, solves several with the
 * write-first kernel in the value protocol; a plan created with one of
 * those algorithms refuses k > 1 with CAP_ERR_UNSUPPORTED.) */
int cap_plan_info(cap_plan plan, int *algo, int *sync, int *border,
                  int *mix_len, double *warp_occupy, double *element_occupy);
/* Grid, window, readiness and row-tile choices the plan uses (after
 * autotuning); *readiness is CAP_READY_STAGED or CAP_READY_IN_PLACE;
 * *row_threads of a CAP_ALGO_LEVEL plan is its one workgroup's size. */
int cap_plan_tuning(cap_plan plan, int *blocks_per_cu, int *lds_window, int *adjoint,
                    int *readiness, int *row_threads, int *row_tiles);
int cap_plan_destroy(cap_plan plan);

/* Adjoint of x = L^{-1} b: grad_b = L^{-T} g and grad_val[k] =
 * -grad_b[row_k] * x[col_k] on the pattern (grad_val may be NULL).  FP64
 * plans only.  Device pointers, asynchronous on `stream`.  cap_opts.adjoint
 * selects the method: SCATTER is the column-driven sync-free transposed
 * solve on the same CSR (in-degree counters + atomic left sums; the order of
 * the atomic additions depends on scheduling), GATHER solves a reversed
 * transpose of L built on first use (bit-reproducible, the executor's
 * backward accumulation order). */
int cap_plan_adjoint(cap_plan plan, const void *g, const void *x, void *grad_b,
                     void *grad_val, void *stream);

/* FP32 solve inside FP64 iterative refinement.  `plan64` is an FP64 plan of
 * L (used for the FP64 residual r = b - L x); the FP32 copy of the values is
 * made once and cached inside it.  Iterates until max|x_k - x_{k-1}| <=
 * tol * max|x_k| or max_iter; writes the iteration count and the final
 * relative correction.  Device pointers, synchronous (returns after the last
 * check). */
int cap_plan_solve_refined(cap_plan plan64, const double *b, double *x, int rhs,
                           double tol, int max_iter, int *iters,
                           double *last_rel_correction, void *stream);

/* ------------------------------------------------------------ batched API
 * Many independent lower-triangular matrices with k interleaved right-hand
 * sides each (1 <= k <= 64), solved in one launch with per-matrix slices of
 * the readiness state.  Matrices are concatenated: matrix m owns rows
 * [row_off[m], row_off[m+1]); row_ptr (n_total+1 entries) indexes the
 * concatenated col_idx/val; column indices are LOCAL to their matrix.  b/x
 * are concatenated the same way, k interleaved values per row.
 * opts->algo: CAP_ALGO_WRTFST (thread-per-row over all matrices, row blocks
 * handed out by relative level -- level / depth of their matrix -- so that
 * every matrix advances at the same pace), CAP_ALGO_LOCAL (one workgroup per
 * matrix, largest first) or CAP_ALGO_AUTO (both timed at creation, the
 * faster kept).  opts->blocks_per_cu = 0 lets creation time 1, 2 and 4
 * workgroups per compute unit for the thread-per-row schedule, and, with
 * k > 1 (value protocol, no LDS window), 1024-thread workgroups with 1, 2
 * or 4 right-hand sides per lane, each with a schedule of its block size.
 * With CAP_ALGO_AUTO a batch may also go hybrid: its deepest matrices
 * (at least 1024 levels, at most 256 rows per level, taken deepest first
 * while their partitioned plans build: lower bandwidth <= 256, stable
 * blocks) are solved by partitioned plans of their own, and the rest by the
 * thread-per-row schedule without them; each prefix of that list is timed
 * at creation against the plain batch and kept only when faster.  Those
 * matrices' results are then within rounding of the sequential solve
 * (reassociated sums, <= 1e-12), not bit-exact.  Not with the reference
 * arithmetic or readiness in place. */
typedef struct cap_batch_s *cap_batch;
int cap_batch_create(cap_batch *batch, int n_mat, const int *row_off,
                     const int *row_ptr, const int *col_idx, const void *val,
                     int value_type, int k, const cap_opts *opts, void *stream);
/* The algorithm the batch resolved to (CAP_ALGO_WRTFST or CAP_ALGO_LOCAL)
 * and its persistent grid size in workgroups per compute unit. */
int cap_batch_info(cap_batch batch, int *algo, int *blocks_per_cu);
/* No counterpart in the reference: the thread-per-row solve's workgroup size
 * (128, 256 or 1024), right-hand sides per lane and row tiles (1: each
 * block's col_idx/val range staged in LDS) the batch resolved to, and how
 * many of its matrices it solves with partitioned plans of their own (the
 * hybrid batch of CAP_ALGO_AUTO: see cap_batch_create). */
int cap_batch_tuning(cap_batch batch, int *row_threads, int *rhs_per_lane, int *partitioned,
                     int *row_tiles);
/* No counterpart in the reference: the first rows of the matrices a hybrid
 * batch solves with partitioned plans (up to `capacity` of them); returns
 * how many there are, or a negative CAP_ERR_* code. */
int cap_batch_partitioned(cap_batch batch, int *first_rows, int capacity);
int cap_batch_solve(cap_batch batch, const void *b, void *x, void *stream);
/* FP32 solves inside FP64 refinement for every thread-per-row or
 * one-workgroup matrix.  A hybrid batch's partitioned matrices (FP64 tables,
 * checked at creation against a sequential FP64 solve to 1e-13) get one FP64
 * partitioned solve when tol >= 1e-13, FP64 refinement otherwise.  *iters
 * and *last_rel_correction are the largest over the batch's parts. */
int cap_batch_solve_refined(cap_batch batch64, const double *b, double *x,
                            double tol, int max_iter, int *iters,
                            double *last_rel_correction, void *stream);
int cap_batch_destroy(cap_batch batch);

/* ------------------------------------------------- one-shot solves (host)
 * sptrsv_csr_solve_{d,s}: validate, copy to the device, solve, copy x back.
 * All pointers are host pointers.  `opts` may be NULL. */
int sptrsv_csr_solve_d(int n, int nnz, const int *row_ptr, const int *col_idx,
                       const double *val, const double *b, double *x,
                       const cap_opts *opts);
int sptrsv_csr_solve_s(int n, int nnz, const int *row_ptr, const int *col_idx,
                       const float *val, const float *b, float *x,
                       const cap_opts *opts);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc: the one-shot solve above (one right-hand side, FP64,
 * thread-per-row write-first kernel, fused arithmetic) on the caller's
 * DEVICE arrays, with nothing allocated.  x itself is the readiness array;
 * the tickets and the watchdog word live in a static 256-byte array of the
 * library on each device (held until the process ends) that every call on
 * that device shares: calls that overlap on one device race, so serialize
 * them (one stream, or events).  Limits: the kernel's own (n < 2^31 rows);
 * the scratch does not grow with the input.  Asynchronous on `stream`; the
 * CSR layout is not validated (no host copy is made) -- a wait that exceeds
 * the default 10 s wall-time bound gives up and leaves NaN in the rows
 * involved.  Returns CAP_OK, CAP_ERR_ARG, CAP_ERR_NO_DEVICE or CAP_ERR_HIP.
 * The first call on a device does its set-up (device check, the scratch
 * address), which is not an allocation either. */
int sptrsv_csr_solve_noalloc_d(int n, int nnz, const int *row_ptr,
                               const int *col_idx, const double *val,
                               const double *b, double *x, void *stream);

/* -------------------------------------------- benchmark-compatible entries
 * Same parameter lists and semantics as the reference's wrappers: build
 * x_ref = 1 and b = L*x_ref internally (get_x_b), run a warm-up plus
 * CAP_BENCH_REPEAT timed solves (flags/x reset outside the timed region, as
 * the reference does), validate every repetition with the reference's test
 * (|x - x_ref| > 0.01|x_ref| counts as an error), and report
 *   solve_time = mean ms, gflops = 2*nnzL/(1e6*ms),
 *   bandwith  = dataSize/(1e6*ms), dataSize = 4(n+1) + 4(nnzL+m)
 *               + V*nnzL + 2V*n (+4*Len for mix).
 * A failed validation sets solve_time = -1, gflops = bandwith = 0 and still
 * returns 0; a non-square input returns -1.  Other negative codes are this
 * library's input checks (the reference does not check). */
#define CAP_BENCH_REPEAT 100

/* replaces spts_syncfree_cuda_csr_wrt, CUDA/ourWrtFst/spts_syncfree_cuda_csr_wrt.h:66-75 */
int spts_syncfree_hip_csr_wrt_d(const int *csrRowPtrL, const int *csrColIdxL,
                                const double *csrValL, int m, int n, int nnzL,
                                double *solve_time, double *gflops, double *bandwith);
int spts_syncfree_hip_csr_wrt_s(const int *csrRowPtrL, const int *csrColIdxL,
                                const float *csrValL, int m, int n, int nnzL,
                                double *solve_time, double *gflops, double *bandwith);

/* replaces spts_syncfree_cuda_csr (our2Part), CUDA/our2Part/spts_syncfree_cuda_csr.h:133-142 */
int spts_syncfree_hip_csr_d(const int *csrRowPtrL, const int *csrColIdxL,
                            const double *csrValL, int m, int n, int nnzL,
                            double *solve_time, double *gflops, double *bandwith);
int spts_syncfree_hip_csr_s(const int *csrRowPtrL, const int *csrColIdxL,
                            const float *csrValL, int m, int n, int nnzL,
                            double *solve_time, double *gflops, double *bandwith);

/* replaces the mix spts_syncfree_cuda_csr_wrt, CUDA/mix/spts_syncfree_cuda_csr_wrt.h:146-158
 * (pre_time = mean ms of the chunk-table construction over CAP_BENCH_REPEAT) */
int spts_syncfree_hip_mix_d(const int *csrRowPtrL, const int *csrColIdxL,
                            const double *csrValL, int m, int n, int nnzL,
                            double *solve_time, double *gflops, double *bandwith,
                            double *pre_time, double *warp_occupy,
                            double *element_occupy);
int spts_syncfree_hip_mix_s(const int *csrRowPtrL, const int *csrColIdxL,
                            const float *csrValL, int m, int n, int nnzL,
                            double *solve_time, double *gflops, double *bandwith,
                            double *pre_time, double *warp_occupy,
                            double *element_occupy);

/* No counterpart in the reference: the same benchmark-compatible wrapper
 * (get_x_b, validation, CAP_BENCH_REPEAT timed solves, the same outputs)
 * with the algorithm chosen by the caller: CAP_ALGO_AUTO, WRTFST, TWOPART,
 * CSC, LOCAL, PARTITION or LEVEL (for MIX use the mix wrapper, which also reports
 * the chunk table).  It lets the drop-in driver run the autotuned and the
 * deep-matrix solvers with the same result.csv columns.  A matrix the
 * algorithm refuses returns its CAP_ERR_* code. */
int spts_syncfree_hip_csr_algo_d(const int *csrRowPtrL, const int *csrColIdxL,
                                 const double *csrValL, int m, int n, int nnzL, int algo,
                                 double *solve_time, double *gflops, double *bandwith);
int spts_syncfree_hip_csr_algo_s(const int *csrRowPtrL, const int *csrColIdxL,
                                 const float *csrValL, int m, int n, int nnzL, int algo,
                                 double *solve_time, double *gflops, double *bandwith);

/* replaces sptrsv_syncfree_cuda, CUDA/SyncFree_csc/spts_syncfree_cuda2.h:89-102
 * (CSC input, x output; validation = relative L1 error < 1e-4, :245-264) */
int sptrsv_syncfree_hip_csc_d(const int *cscColPtr, const int *cscRowIdx,
                              const double *cscVal, int m, int n, int nnz,
                              double *x, const double *b, const double *x_ref,
                              double *pre_time, double *solve_time,
                              double *gflops, double *bandwidth);
int sptrsv_syncfree_hip_csc_s(const int *cscColPtr, const int *cscRowIdx,
                              const float *cscVal, int m, int n, int nnz,
                              float *x, const float *b, const float *x_ref,
                              double *pre_time, double *solve_time,
                              double *gflops, double *bandwidth);

/* ------------------------------------------------------- host helpers
 * MI355X-build restatements of the reference's host preprocessing.  Arrays
 * returned through T** are malloc'ed; release them with cap_free. */

/* read_mtx, CUDA/ourWrtFst/read_mtx.h:6-191 (strict size line of
 * CUDA/ourWrtFst/mmio.h:334-365).  Returns 0, -1 open, -2 banner,
 * -3 complex, -4 size line, CAP_ERR_IO malformed entry. */
int cap_read_mtx_d(const char *path, int *m, int *n, int *nnzA, int **row_ptr,
                   int **col_idx, double **val);
int cap_read_mtx_s(const char *path, int *m, int *n, int *nnzA, int **row_ptr,
                   int **col_idx, float **val);

/* change2tran, CUDA/ourWrtFst/read_mtx.h:194-257 (unit-valued strictly lower
 * part of the leading m x m block + unit diagonal).  keep_values = 1 keeps
 * A's values instead of 1 and sets the diagonal to 1 + sum|offdiag| (the
 * opt-in real-valued mode; the reference always uses 1). */
int cap_change2tran_d(int m, int nnzA, const int *rpA, const int *ciA,
                      const double *vA, int keep_values, int *nnzL, int **rpL,
                      int **ciL, double **vL);
int cap_change2tran_s(int m, int nnzA, const int *rpA, const int *ciA,
                      const float *vA, int keep_values, int *nnzL, int **rpL,
                      int **ciL, float **vL);

/* get_x_b, read_mtx.h:260-278 (x_ref = 1, b = L x_ref in CSR order) */
void cap_get_x_b_d(int m, int n, const int *rp, const int *ci, const double *v,
                   double *x_ref, double *b);
void cap_get_x_b_s(int m, int n, const int *rp, const int *ci, const float *v,
                   float *x_ref, float *b);

/* matrix_layer, CUDA/ourWrtFst/tranpose.h:78-142 (1-based level count,
 * parallelism = m / levels); row_level may be NULL. */
int cap_matrix_layer(int m, const int *rp, const int *ci, int *levels,
                     double *parallelism, int *row_level);

/* matrix_warp4, CUDA/mix/tranpose.h:1423-1477 with chunk = rows per
 * selector chunk (64 = one wavefront on gfx950; the reference's WARP_SIZE).
 * warp_num must hold m+1 ints. */
void cap_matrix_warp4(int m, int nnz, const int *rp, int border, int chunk,
                      int *len, int *warp_num, double *warp_occupy,
                      double *element_occupy);

/* matrix_transposition, CUDA/ourWrtFst/tranpose.h:36-75 */
void cap_transpose_d(int m, int n, int nnz, const int *rp, const int *ci,
                     const double *v, int *csc_row, int *csc_colptr,
                     double *csc_val);
void cap_transpose_s(int m, int n, int nnz, const int *rp, const int *ci,
                     const float *v, int *csc_row, int *csc_colptr,
                     float *csc_val);

/* Layout check of the solve contract above; returns CAP_OK or the first
 * violation's code and (if bad_row != NULL) its row. */
int cap_validate_csr_d(int n, int nnz, const int *rp, const int *ci,
                       const double *v, int *bad_row);
int cap_validate_csr_s(int n, int nnz, const int *rp, const int *ci,
                       const float *v, int *bad_row);

/* Synthetic generators for the benchmark workloads (splitmix64 seeded; the
 * same seed gives the same matrix on every host).  kind:
 *   0  band: row i gets min(k_max, i) distinct columns uniform in
 *      [max(0, i-band), i) (k_max = offdiag), values U[-1,1],
 *      diagonal 1 + sum|offdiag|
 *   1  level: n = levels*width; a row of level l >= 1 gets 1..offdiag
 *      distinct columns uniform over the rows of levels [max(0,l-depth), l)
 *   2  banded Bernoulli: every column in [max(0,i-band), i) kept with
 *      probability p
 *   3  band with per-row nnz: row i gets min(U{0..offdiag-1}, i) distinct
 *      columns in [max(0,i-band), i)
 * Two passes: call with ci/v = NULL to get *nnz and row_ptr, then again
 * with arrays of that size. */
typedef struct cap_gen_params {
    int kind, n, offdiag, band, levels, width, depth;
    double p;
    unsigned long long seed;
} cap_gen_params;
int cap_generate_d(const cap_gen_params *g, int *nnz, int *rp, int *ci, double *v);
int cap_generate_s(const cap_gen_params *g, int *nnz, int *rp, int *ci, float *v);
/* fills out[0..count) with U[lo,hi) doubles from splitmix64(seed) */
void cap_uniform_d(unsigned long long seed, long long count, double lo, double hi, double *out);

/* ------------------------------------------------------- serial solves
 * The reference's CPU path, for callers that want the CPU result beside the
 * GPU one (no GPU entry of this library calls these).  Multi-RHS values are
 * interleaved, x[i*rhs + k] (serialref.h:57-60).  Where the reference would
 * spin forever or index out of range (a column without its diagonal in the
 * expected place, an entry on the wrong side of it, a row index outside
 * [0, m), an in-degree table that does not match the matrix) these return a
 * CAP_ERR_* code instead. */
#define CAP_SUBSTITUTION_FORWARD 0  /* SUBSTITUTION_FORWARD, openCL/SynFree_csc/common.h:25 */
#define CAP_SUBSTITUTION_BACKWARD 1 /* SUBSTITUTION_BACKWARD, common.h:26 */

/* replaces sptrsv_syncfree_analyser, openCL/SynFree_csc/sptrsv_syncfree_serialref.h:6-23:
 * in_degree[r] = entries of row r in the CSC matrix (m ints). */
int cap_serial_analyser(const int *csc_row_idx, int m, int n, int nnz, int *in_degree);

/* replaces sptrsv_syncfree_executor, serialref.h:25-108: column-driven
 * substitution.  FORWARD visits columns 0..n-1 with each column's diagonal
 * first; BACKWARD visits n-1..0 with the diagonal last.  x = (b - left) / diag
 * and left += x * v with a rounded product and a rounded sum: the reference's
 * own bits. */
int cap_serial_executor_d(const int *cscColPtr, const int *cscRowIdx, const double *cscVal,
                          const int *in_degree, int m, int n, int substitution, int rhs,
                          const double *b, double *x);
int cap_serial_executor_s(const int *cscColPtr, const int *cscRowIdx, const float *cscVal,
                          const int *in_degree, int m, int n, int substitution, int rhs,
                          const float *b, float *x);

/* replaces sptrsv_syncfree_serialref, serialref.h:110-177: analyser +
 * executor, timed, with the relative L1 check against x_ref (< 1e-4) and the
 * reference's report lines on stdout.  Returns 0 (also when the check fails,
 * as the reference does), -1 for a non-square matrix, or the executor's code
 * for a malformed one. */
int sptrsv_syncfree_serialref_d(const int *cscColPtrTR, const int *cscRowIdxTR,
                                const double *cscValTR, int m, int n, int nnzTR,
                                int substitution, int rhs, double *x, const double *b,
                                const double *x_ref);
int sptrsv_syncfree_serialref_s(const int *cscColPtrTR, const int *cscRowIdxTR,
                                const float *cscValTR, int m, int n, int nnzTR,
                                int substitution, int rhs, float *x, const float *b,
                                const float *x_ref);

/* No counterpart in the reference (whose serial solve is column-driven only):
 * row-ordered forward substitution over the solve contract's CSR (host
 * arrays, checked as cap_validate_csr), off-diagonal terms in stored order,
 * arith = CAP_ARITH_FUSED or CAP_ARITH_REFERENCE.  It returns the bits of the
 * GPU thread and wavefront kernels with the same arith; with sorted rows and
 * CAP_ARITH_REFERENCE, the executor's bits too. */
int cap_serial_csr_d(int n, int nnz, const int *row_ptr, const int *col_idx, const double *val,
                     int rhs, const double *b, double *x, int arith);
int cap_serial_csr_s(int n, int nnz, const int *row_ptr, const int *col_idx, const float *val,
                     int rhs, const float *b, float *x, int arith);

/* reference metric formulas, spts_syncfree_cuda_csr_wrt.h:206-215 */
double cap_datasize(int m, int n, int nnzL, int value_bytes, int len_warp_num);

/* device queries */
int cap_device_count(void);
const char *cap_strerror(int code);
const char *cap_version(void);
void cap_free(void *p);

/* --------------------------------------------- VALUE_TYPE drop-in names */
#ifndef VALUE_TYPE
#define VALUE_TYPE double
#endif
static inline int spts_syncfree_hip_csr_wrt(const int *rp, const int *ci, const VALUE_TYPE *v,
                                            int m, int n, int nnzL, double *t, double *gf,
                                            double *bw)
{
    return sizeof(VALUE_TYPE) == 8
               ? spts_syncfree_hip_csr_wrt_d(rp, ci, (const double *)(const void *)v, m, n, nnzL, t, gf, bw)
               : spts_syncfree_hip_csr_wrt_s(rp, ci, (const float *)(const void *)v, m, n, nnzL, t, gf, bw);
}
static inline int spts_syncfree_hip_csr(const int *rp, const int *ci, const VALUE_TYPE *v,
                                        int m, int n, int nnzL, double *t, double *gf,
                                        double *bw)
{
    return sizeof(VALUE_TYPE) == 8
               ? spts_syncfree_hip_csr_d(rp, ci, (const double *)(const void *)v, m, n, nnzL, t, gf, bw)
               : spts_syncfree_hip_csr_s(rp, ci, (const float *)(const void *)v, m, n, nnzL, t, gf, bw);
}
static inline int spts_syncfree_hip_csr_algo(const int *rp, const int *ci, const VALUE_TYPE *v,
                                             int m, int n, int nnzL, int algo, double *t,
                                             double *gf, double *bw)
{
    return sizeof(VALUE_TYPE) == 8
               ? spts_syncfree_hip_csr_algo_d(rp, ci, (const double *)(const void *)v, m, n, nnzL, algo, t, gf, bw)
               : spts_syncfree_hip_csr_algo_s(rp, ci, (const float *)(const void *)v, m, n, nnzL, algo, t, gf, bw);
}
static inline int spts_syncfree_hip_mix(const int *rp, const int *ci, const VALUE_TYPE *v,
                                        int m, int n, int nnzL, double *t, double *gf,
                                        double *bw, double *pre, double *wocc, double *eocc)
{
    return sizeof(VALUE_TYPE) == 8
               ? spts_syncfree_hip_mix_d(rp, ci, (const double *)(const void *)v, m, n, nnzL, t, gf, bw, pre, wocc, eocc)
               : spts_syncfree_hip_mix_s(rp, ci, (const float *)(const void *)v, m, n, nnzL, t, gf, bw, pre, wocc, eocc);
}
static inline int sptrsv_syncfree_hip_csc(const int *cp, const int *ri, const VALUE_TYPE *v,
                                          int m, int n, int nnz, VALUE_TYPE *x,
                                          const VALUE_TYPE *b, const VALUE_TYPE *xr,
                                          double *pre, double *t, double *gf, double *bw)
{
    return sizeof(VALUE_TYPE) == 8
               ? sptrsv_syncfree_hip_csc_d(cp, ri, (const double *)(const void *)v, m, n, nnz,
                                           (double *)(void *)x, (const double *)(const void *)b,
                                           (const double *)(const void *)xr, pre, t, gf, bw)
               : sptrsv_syncfree_hip_csc_s(cp, ri, (const float *)(const void *)v, m, n, nnz,
                                           (float *)(void *)x, (const float *)(const void *)b,
                                           (const float *)(const void *)xr, pre, t, gf, bw);
}
static inline int sptrsv_syncfree_serialref(const int *cp, const int *ri, const VALUE_TYPE *v,
                                            int m, int n, int nnz, int substitution, int rhs,
                                            VALUE_TYPE *x, const VALUE_TYPE *b,
                                            const VALUE_TYPE *xr)
{
    return sizeof(VALUE_TYPE) == 8
               ? sptrsv_syncfree_serialref_d(cp, ri, (const double *)(const void *)v, m, n, nnz,
                                             substitution, rhs, (double *)(void *)x,
                                             (const double *)(const void *)b,
                                             (const double *)(const void *)xr)
               : sptrsv_syncfree_serialref_s(cp, ri, (const float *)(const void *)v, m, n, nnz,
                                             substitution, rhs, (float *)(void *)x,
                                             (const float *)(const void *)b,
                                             (const float *)(const void *)xr);
}

#ifdef __cplusplus
}
#endif
#endif
