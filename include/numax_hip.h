/*
 * numax_hip.h -- C ABI of libnumax_hip.so, the MI355X (gfx950) kernels behind
 * numax.linalg's dense factor-and-solve path.
 *
 * Conventions (all entry points):
 *   - matrices are row-major device pointers with a leading dimension
 *     (`lda` >= number of columns), as numax's `row_major` layouts;
 *   - pivots are int32, 0-based, LAPACK `ipiv` semantics: row j was
 *     interchanged with row ipiv[j], in order j = 0..n-1;
 *   - `info` is a device int32 written LAPACK-style (1-based index of the
 *     first failure, 0 if none), never raised inside a kernel;
 *   - every call is asynchronous on the context's stream;
 *   - return value: 0 on success, -i if argument i is invalid, a positive
 *     NMX_ERR_* code otherwise.  No exception crosses the ABI.
 *
 * The reference (numax 0.3.0, Mojo) has no FFI; each declaration names the
 * Mojo entry point whose device work it performs.
 */
#ifndef NUMAX_HIP_H
#define NUMAX_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NMX_OK 0
#define NMX_ERR_HIP 1         /* a HIP runtime call failed */
#define NMX_ERR_ALLOC 2       /* device allocation failed */
#define NMX_ERR_UNSUPPORTED 3 /* size or option outside what the kernels handle */
#define NMX_ERR_SYNC_TIMEOUT 4 /* a cross-workgroup wait hit its spin bound */
#define NMX_ERR_NOT_CONVERGED 5 /* refinement did not reach its target */
#define NMX_ERR_IO 6          /* a file could not be opened, read or written */
#define NMX_ERR_FORMAT 7      /* not a .npy / NMX1 file, an unsupported .npy version, or a truncated header */
#define NMX_ERR_DTYPE 8       /* the file's dtype is not the one requested, or not one numax stores */
#define NMX_ERR_SHAPE 9       /* the file's rank, shape or payload size disagrees with the request */
#define NMX_ERR_LAYOUT 10     /* a Fortran-ordered or big-endian .npy file (only row-major little-endian is read) */

typedef struct nmx_ctx nmx_ctx;

/* ---- context -------------------------------------------------------------
 * Owns the stream handle (borrowed when `stream` is non-null), a grow-only
 * device workspace and the cross-workgroup mailbox.  One context per host
 * thread; calls on one context are not thread-safe.
 * Replaces numax's DeviceContext plumbing (numax/linalg/panel.mojo:10-24). */
int nmx_ctx_create(int device, void* stream, nmx_ctx** out);
/* While graphs captured on the context exist, destroying it only marks it:
 * its buffers and stream stay valid for their replays and are released by
 * the last nmx_graph_destroy. */
int nmx_ctx_destroy(nmx_ctx* ctx);
/* Points the context at another stream (NULL: a new stream it owns).  Work
 * the new stream runs after this call waits for the work already queued on
 * the old one (an event), since both use the context's buffers. */
int nmx_ctx_set_stream(nmx_ctx* ctx, void* stream);
int nmx_sync(nmx_ctx* ctx);
/* Sticky device-side error word of the context (NMX_ERR_SYNC_TIMEOUT after a
 * bounded wait expired); reading it synchronizes the stream and clears it. */
int nmx_ctx_check(nmx_ctx* ctx);
const char* nmx_status_string(int status);
int nmx_version(void);

/* Optional per-kernel-class event timing on the context's stream (no
 * reference counterpart: it serves the roofline report).  Turning it on or
 * off clears the totals; nmx_ctx_profile synchronizes and returns the totals
 * of class `kind` (0 .. 13, names from nmx_kernel_name). */
int nmx_ctx_set_profiling(nmx_ctx* ctx, int on);
int nmx_ctx_profile(nmx_ctx* ctx, int kind, int64_t* launches, double* ms, double* flops, double* bytes);
const char* nmx_kernel_name(int kind);

/* ---- LU with partial pivoting --------------------------------------------
 * lu_factor[gpu=True] (numax/linalg/lu.mojo:454-701, runtime form :1200-1312):
 * in-place P A = L U, unit L below the diagonal, first-maximum pivots, full
 * row interchanges, an exact-zero pivot divided as 1e-30 and recorded in
 * `info` (not an error).  `block` is the panel width, `base` the recursive
 * panel's leaf width (0 picks the defaults 256 / 32).  n is at most 32768
 * (the rows the cooperative pivoting leaf holds); a larger n returns
 * NMX_ERR_UNSUPPORTED before any work is queued. */
int nmx_getrf_f64(nmx_ctx* ctx, int n, double* A, int lda, int32_t* ipiv, int32_t* info, int block, int base);
int nmx_getrf_f32(nmx_ctx* ctx, int n, float* A, int lda, int32_t* ipiv, int32_t* info, int block, int base);

/* LU.solve / DynamicLU.solve (lu.mojo:181-354, :1058-1124): laswp, then the
 * unit-lower and upper blocked substitutions, in place on the n x nrhs B. */
int nmx_getrs_f64(nmx_ctx* ctx, int n, int nrhs, const double* LU, int lda, const int32_t* ipiv, double* B, int ldb);
int nmx_getrs_f32(nmx_ctx* ctx, int n, int nrhs, const float* LU, int lda, const int32_t* ipiv, float* B, int ldb);

/* laswp / laswp_matrix (numax/linalg/panel.mojo:1824-1852, :2279-2313). */
int nmx_laswp_f64(nmx_ctx* ctx, int n, int nrhs, double* B, int ldb, const int32_t* ipiv);
int nmx_laswp_f32(nmx_ctx* ctx, int n, int nrhs, float* B, int ldb, const int32_t* ipiv);

/* inverse (numax/linalg/basic.mojo:149-195) after lu_factor: A^-1 written to
 * Ainv (n x n, ld ldi) by LU.solve against the identity -- n right-hand
 * sides, blocked substitutions.  A singular factor yields non-finite
 * entries, not an error. */
int nmx_getri_f64(nmx_ctx* ctx, int n, const double* LU, int lda, const int32_t* ipiv, double* Ainv, int ldi);
int nmx_getri_f32(nmx_ctx* ctx, int n, const float* LU, int lda, const int32_t* ipiv, float* Ainv, int ldi);

/* lu -> PLU (numax/linalg/lu.mojo:849-975): lu_factor's packed factor and
 * interchanges unpacked into A = P L U -- L unit lower, U upper, perm[i] the
 * row of A that row i of L U came from (the interchanges replayed in order
 * on the identity), P[perm[c]][c] = 1.  Every output is n x n row-major
 * (perm n entries), device memory. */
int nmx_lu_unpack_f64(nmx_ctx* ctx, int n, const double* LU, int lda, const int32_t* ipiv, int32_t* perm, double* P,
                      int ldp, double* L, int ldl, double* U, int ldu);
int nmx_lu_unpack_f32(nmx_ctx* ctx, int n, const float* LU, int lda, const int32_t* ipiv, int32_t* perm, float* P,
                      int ldp, float* L, int ldl, float* U, int ldu);

/* LU._diagonal (lu.mojo:355-379): gather diag(A) into an n-vector for
 * det / slogdet. */
int nmx_diag_f64(nmx_ctx* ctx, int n, const double* A, int lda, double* diag);
int nmx_diag_f32(nmx_ctx* ctx, int n, const float* A, int lda, float* diag);

/* solve_triangular / _trsm (numax/linalg/triangular.mojo:114-225, :228-425):
 * op(T) X = B in place, T n x n; `upper` is the triangle the substitution
 * walks, `unit` ignores T's diagonal, `trans` reads T transposed. */
int nmx_trsm_f64(nmx_ctx* ctx, int upper, int unit, int trans, int n, int nrhs, const double* T, int ldt, double* B,
                 int ldb);
int nmx_trsm_f32(nmx_ctx* ctx, int upper, int unit, int trans, int n, int nrhs, const float* T, int ldt, float* B,
                 int ldb);

/* ---- symmetric indefinite LDL^T ---------------------------------------------
 * ldl (numax/linalg/ldl.mojo:59-148): Bunch-Kaufman sytf2 on the lower
 * triangle (numax/linalg/panel.mojo:2737-2894) in place, one workgroup;
 * ipiv 1-based as LAPACK's (kp + 1 for a 1 x 1 block swapped with row kp,
 * -(kp + 1) on both rows of a 2 x 2 block); info = first k + 1 whose column
 * is exactly zero (the factorization goes on and yields non-finite entries,
 * as numax's does).  The strict upper triangle is not read. */
int nmx_sytrf_f64(nmx_ctx* ctx, int n, double* A, int lda, int32_t* ipiv, int32_t* info);
int nmx_sytrf_f32(nmx_ctx* ctx, int n, float* A, int lda, int32_t* ipiv, int32_t* info);
/* SciPy's (lu, d, perm) from the packed factor (panel.mojo:2897-3003):
 * A = lu d lu^T, lu[perm] unit lower triangular, d block diagonal. */
int nmx_ldl_unpack_f64(nmx_ctx* ctx, int n, const double* LDU, int lda, const int32_t* ipiv, double* L, int ldl,
                       double* D, int ldd, int64_t* perm);
int nmx_ldl_unpack_f32(nmx_ctx* ctx, int n, const float* LDU, int lda, const int32_t* ipiv, float* L, int ldl,
                       float* D, int ldd, int64_t* perm);

/* ---- Cholesky ------------------------------------------------------------
 * cholesky[gpu=True] (numax/linalg/cholesky.mojo:59-396): lower L with
 * L L^T = A reading only A's lower triangle; a non-positive pivot is floored
 * to 1e-30 and its 1-based index recorded in `info` (the caller raises);
 * the strict upper triangle of the result is zeroed (tril). */
int nmx_potrf_f64(nmx_ctx* ctx, int n, double* A, int lda, int32_t* info, int block);
int nmx_potrf_f32(nmx_ctx* ctx, int n, float* A, int lda, int32_t* info, int block);
/* cholesky_solve (cholesky.mojo:471-580): L y = b, then L^T x = y. */
int nmx_potrs_f64(nmx_ctx* ctx, int n, int nrhs, const double* L, int lda, double* B, int ldb);
int nmx_potrs_f32(nmx_ctx* ctx, int n, int nrhs, const float* L, int lda, float* B, int ldb);

/* ---- GEMM update used by the drivers (exported for tests) ----------------
 * C = beta C + alpha op(A) op(B), op(A) m x k, op(B) k x n.  The trailing
 * update `A22 -= L21 U12` of lu.mojo:629-673 is alpha = -1, beta = 1. */
int nmx_gemm_f64(nmx_ctx* ctx, int transa, int transb, int m, int n, int k, double alpha, const double* A, int lda,
                 const double* B, int ldb, double beta, double* C, int ldc);
int nmx_gemm_f32(nmx_ctx* ctx, int transa, int transb, int m, int n, int k, float alpha, const float* A, int lda,
                 const float* B, int ldb, float beta, float* C, int ldc);

/* ---- mixed-precision LU with iterative refinement --------------------------
 * No reference counterpart (numax has no mixed precision and no refinement);
 * it serves `solve(a, b)` (numax/linalg/basic.mojo:55-146) when a caller asks
 * for FP64 accuracy at low-precision factorization cost.  Blocked LU of A
 * in FP32 (P A = L U; pivoting per cfg.pivot) with the trailing updates on
 * FP16 / BF16 MFMA (FP32 accumulate, a power-of-two scale per 32-row block
 * of L21 and per column of U12); then
 * r = b - A x in FP64, x += U^-1 L^-1 P r (or GMRES on that correction
 * equation) until the normwise backward error
 * ||r||_inf / (||A||_inf ||x||_inf) <= tol.
 * A is read, never written; x receives the solution. */
typedef struct nmx_ir_cfg {
  int precision;     /* trailing update: 0 fp16, 1 bf16, 2 fp16 split (hi + 2^-11 lo, 3 products), 3 fp32 */
  int block;         /* panel width, a multiple of 64 (0: 256) */
  int max_iter;      /* refinement steps at most */
  int solver;        /* 0 auto, 1 classical refinement, 2 GMRES-IR */
  int gmres_restart; /* GMRES restart length */
  double tol;        /* target backward error */
  uint64_t seed;     /* butterfly seed (pivot = 0) */
  int pivot;         /* 2: tournament pivoting (default: all pivots of a 32-column leaf from one
                        reduction tree, Grigori-Demmel-Xiang CALU); 1: partial pivoting, column by
                        column (numax's pivots; n <= 32768); 0: random butterflies, no pivoting */
} nmx_ir_cfg;

typedef struct nmx_ir_report {
  int iterations;       /* refinement steps taken */
  int gmres_iterations; /* inner GMRES iterations (GMRES-IR) */
  int converged;        /* backward_error <= tol */
  int info;             /* first exactly-zero pivot (1-based) of the low-precision factor, 0 if none */
  double backward_error;
  double factor_ms;
  double refine_ms;
  double history[32];   /* backward error after each step */
} nmx_ir_report;

void nmx_ir_cfg_default(nmx_ir_cfg* cfg);
/* Classical refinement (solver 1, and solver 0 until a step stops halving
 * the backward error) runs as a fixed launch sequence whose stop test is
 * made on the device, so the call can be captured between nmx_graph_begin
 * and nmx_graph_end once it has run eagerly on the context (buffers sized,
 * second stream created); a captured call fills no report (it returns
 * NMX_OK) and holds max_iter + 1 steps, the ones past the stop returning at
 * once.  Under capture solver 0 means classical refinement only; solver 2
 * (GMRES-IR, decided on the host) and pivot 1 (the cooperative
 * partial-pivoting panel, host-side epochs) return NMX_ERR_UNSUPPORTED. */
int nmx_gesv_mpir(nmx_ctx* ctx, int n, const double* A, int lda, const double* b, double* x, const nmx_ir_cfg* cfg,
                  nmx_ir_report* report);
/* The report of the last nmx_gesv_mpir run on the context's stream, eager
 * or replayed from a graph (synchronizes; the timings and GMRES count are
 * not kept on the device and read 0).  No reference counterpart. */
int nmx_ir_report_read(nmx_ctx* ctx, nmx_ir_report* report);
/* Autotuner of the panel width (no reference counterpart): times the
 * factorization of a synthetic N(0,1) matrix of order n at block 128, 256
 * and 512 and writes the fastest into cfg->block (other fields as given). */
int nmx_ir_autotune(nmx_ctx* ctx, int n, nmx_ir_cfg* cfg);
/* The same solve for an n x nrhs block of right-hand sides (row-major B, X
 * with leading dimensions ldb, ldx >= nrhs): one factorization, then every
 * column refined together -- the FP64 residual R = B - A X is one MFMA GEMM
 * per step, the corrections one pair of blocked FP32 triangular solves --
 * until each column meets cfg->tol (converged columns get no further
 * correction).  Classical refinement; `reps` (nrhs entries, may be NULL)
 * receives a report per column.  No reference counterpart: numax's
 * solve(a, b) with a matrix b is the FP64 path. */
int nmx_gesv_mpir_nrhs(nmx_ctx* ctx, int n, int nrhs, const double* A, int lda, const double* B, int ldb, double* X,
                       int ldx, const nmx_ir_cfg* cfg, nmx_ir_report* reps);
/* Test hook of the factor stage: stage 0 returns W^T A V, stage 1 its FP32
 * no-pivot L\U; `diags` receives the 4n butterfly diagonals (W's r0, r1, s0,
 * s1, t0, t1 then V's), F is n x n row-major (device), info the first zero
 * pivot. */
int nmx_mp_factor_debug(nmx_ctx* ctx, int n, const double* A, int lda, int precision, int block, uint64_t seed,
                        int stage, int pivot, float* F, float* diags, int32_t* info, int32_t* ipiv);

/* ---- batched uneven-size solves --------------------------------------------
 * map_blocks over the register-tier solvers (numax/core/functional.mojo:691-834
 * with numax/linalg/_array/lu.mojo:165-216 lu_factor + PivotedLU.solve, and
 * numax/linalg/_array/cholesky.mojo:26-101 cholesky + cholesky_solve), past
 * the reference's one-n-per-launch, n <= 64 limit: `count` independent
 * systems A_p x_p = b_p of order n[p] in 1..256, each A_p row-major n x n at
 * element offset offA[p] of A (a_elems elements in all), b_p and x_p at
 * offset offB[p] of B and X; kind[p] 0 = LU with partial pivoting, 1 =
 * Cholesky (lower triangle read); B and X hold b_elems elements.  Every
 * array is device memory.  The
 * factorization is FP32, the residual and the solution FP64, refined until
 * the backward error ||b - A x||_inf / (||A||_inf ||x||_inf) <= tol.
 * Per problem: info (first zero / non-positive pivot, 1-based; -1 for a
 * problem outside n 1..256 or kind 0/1, which is skipped), iters (refinement
 * steps) and berr (final backward error).  A problem whose matrix or
 * vectors would fall outside a_elems / b_elems is skipped the same way.
 * Problems may share a matrix (equal or overlapping offA): each is factored
 * in its own FP32 scratch, taken in index order from a pool of a_elems
 * elements, and a problem whose scratch would not fit the pool (possible
 * only when matrices overlap) is skipped with info -1.  Solutions are
 * written at offB, so problems that share b also share x.  The Cholesky
 * problems run on the context's second stream beside the LU problems, forked
 * from and joined back into the context's stream within the call (under
 * graph capture the second stream must already exist from an earlier call,
 * else the call stays on one stream). */
typedef struct nmx_batch_cfg {
  int max_iter;         /* refinement steps at most (default 10) */
  double tol;           /* target backward error (default 1e-15) */
  int teams_per_cu[3];  /* resident teams per CU for n > 128, 64 < n <= 128, n <= 64 */
} nmx_batch_cfg;

void nmx_batch_cfg_default(nmx_batch_cfg* cfg);
int nmx_batched_gesv(nmx_ctx* ctx, int count, const int32_t* n, const int64_t* offA, const int64_t* offB,
                     const uint8_t* kind, const double* A, int64_t a_elems, const double* B, int64_t b_elems, double* X,
                     int32_t* info, int32_t* iters, double* berr, const nmx_batch_cfg* cfg);

/* nmx_batched_gesv without allocating: no reference counterpart; for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The same kernels and the
 * same results as nmx_batched_gesv, with the call's scratch -- 4 a_elems
 * bytes of FP32 working copies, 8 count bytes of offsets and about
 * 4 (count + 1100) bytes of queue, each rounded up to 256 -- in a 64 MiB
 * static device array each device holds for the life of the process.  A
 * call whose scratch exceeds 64 MiB (roughly a_elems + 3 count > 16.7e6)
 * returns NMX_ERR_UNSUPPORTED and queues nothing.  Every call on one device
 * uses that device's one array: calls that overlap on a device (different
 * streams or contexts, not ordered) race on it; each call clears what it
 * needs cleared. */
int nmx_batched_gesv_noalloc(nmx_ctx* ctx, int count, const int32_t* n, const int64_t* offA, const int64_t* offB,
                             const uint8_t* kind, const double* A, int64_t a_elems, const double* B, int64_t b_elems,
                             double* X, int32_t* info, int32_t* iters, double* berr, const nmx_batch_cfg* cfg);

/* Autotuner of the launch shape (no reference counterpart): for each class
 * (n > 128, 64 < n <= 128, n <= 64) times a synthetic batch of `samples`
 * problems of that class at 1, 2, 3, 4, 6 and 8 resident workgroups per CU
 * and writes the fastest into cfg->teams_per_cu; cfg's other fields are
 * used as given.  Allocates its own scratch; a few seconds at samples ~ 2000. */
int nmx_batch_autotune(nmx_ctx* ctx, int samples, nmx_batch_cfg* cfg);

/* ---- batched small symmetric eigensolver -----------------------------------
 * The register-tier eigh (numax/linalg/_array/eigen.mojo:316-392) that
 * map_blocks runs per matrix: cyclic Jacobi with numax's rotation
 * (:34-75) and a fixed number of sweeps (numax's default is 12), FP32.
 * `count` symmetric matrices of order n[p] in 1..64, row-major at element
 * offset offA[p] of A; eigenvalues (unsorted, as numax) to W at offset
 * offW[p], eigenvectors (columns, nullable V) at offA[p] of V.  The pairs
 * are rotated in a parallel (odd-even) order, not numax's row-cyclic one:
 * the converged values agree to rounding, not bit for bit.  A and V hold
 * a_elems elements, W w_elems; problems with n outside 1..64, or whose
 * matrix or eigenvalues would fall outside those sizes, are skipped. */
int nmx_batched_syevj_f32(nmx_ctx* ctx, int count, const int32_t* n, const int64_t* offA, const int64_t* offW,
                          const float* A, int64_t a_elems, float* W, int64_t w_elems, float* V, int sweeps);

/* ---- register-tier operations over batches of small matrices --------------
 * numax's Array-tier functions (numax/linalg/_array/) for `count` row-major
 * FP64 matrices of order n[p] in 1..64 (numax's register tier stops at 64,
 * numax/linalg/common.mojo:30) at element offset offA[p] of A (a_elems
 * elements), one wavefront per matrix.  Loop orders are numax's and no
 * multiply-add is fused, so results equal numax's float64 CPU results bit
 * for bit (slogdet's logarithms to rounding).  A problem whose n is out of
 * range, or whose matrix or vectors would fall outside the arrays, is
 * skipped and its outputs left untouched.
 *   lu_nopiv     lu (_array/lu.mojo:25-55): unpivoted Doolittle, packed L\U
 *                to LU at offA[p] (LU may be A).
 *   solve_nopiv  solve (_array/basic.mojo:20-31): unpivoted LU, forward and
 *                back substitution; b_p at offB[p] of B (b_elems), x_p at
 *                offB[p] of X.
 *   inverse      inverse (_array/basic.mojo:34-60): unpivoted, column by
 *                column of the identity, to Ainv at offA[p].
 *   det          det (_array/lu.mojo:219-233; pivoted = 0, the unpivoted
 *                product) or PivotedLU.det (:112-123; pivoted = 1, sign of
 *                the row swaps times U's diagonal), to det[p].
 *   slogdet      slogdet / PivotedLU.slogdet (_array/lu.mojo:125-162,
 *                :236-258): sign[p] in {-1, 0, +1}, logabsdet[p] (-inf when
 *                U has a zero on its diagonal).
 *   slogdet_cholesky  (_array/cholesky.mojo:95-116) 2 sum ln L_ii of the
 *                Cholesky factors L_p (row-major, lower triangle read) at
 *                offL[p], to logdet[p]. */
int nmx_batched_lu_nopiv_f64(nmx_ctx* ctx, int count, const int32_t* n, const int64_t* offA, const double* A,
                             int64_t a_elems, double* LU);
int nmx_batched_solve_nopiv_f64(nmx_ctx* ctx, int count, const int32_t* n, const int64_t* offA, const int64_t* offB,
                                const double* A, int64_t a_elems, const double* B, int64_t b_elems, double* X);
int nmx_batched_inverse_nopiv_f64(nmx_ctx* ctx, int count, const int32_t* n, const int64_t* offA, const double* A,
                                  int64_t a_elems, double* Ainv);
int nmx_batched_det_f64(nmx_ctx* ctx, int count, const int32_t* n, const int64_t* offA, const double* A,
                        int64_t a_elems, int pivoted, double* det);
int nmx_batched_slogdet_f64(nmx_ctx* ctx, int count, const int32_t* n, const int64_t* offA, const double* A,
                            int64_t a_elems, double* sign, double* logabsdet);
int nmx_batched_slogdet_cholesky_f64(nmx_ctx* ctx, int count, const int32_t* n, const int64_t* offL, const double* L,
                                     int64_t l_elems, double* logdet);

/* map_blocks' structure-of-arrays batch (numax/core/functional.mojo:691-834,
 * layout :729-735): xs is (n*n + n, batch) row-major -- row j < n*n holds
 * entry j of every problem's row-major A, row n*n + i entry i of every b --
 * and ys (n, batch) receives every x, one problem per column, as the
 * batched-solve step bodies compute it (examples/advanced/batched_solve.mojo
 * :87-105): kind 0 lu_factor + PivotedLU.solve, 1 cholesky + cholesky_solve
 * (lower triangle read), 2 unpivoted solve.  n in 1..64; up to n = 8 one GPU
 * thread holds one problem in registers, as map_blocks' lane does. */
int nmx_map_blocks_solve_f64(nmx_ctx* ctx, int kind, int n, int64_t batch, const double* xs, double* ys);
int nmx_map_blocks_solve_f32(nmx_ctx* ctx, int kind, int n, int64_t batch, const float* xs, float* ys);

/* ---- batched tall-skinny least squares ---------------------------------------
 * lstsq(a, b) with method "qr" (numax/linalg/qr.mojo:746-821: Householder
 * QR with LAPACK dlarfg reflectors, Q^T b, back substitution with R) for
 * `count` problems A_p (m x n row-major, contiguous, 1 <= n <= 32, m >= n),
 * b_p (m) and x_p (n), FP32, packed back to back. */
int nmx_batched_gels_f32(nmx_ctx* ctx, int count, int m, int n, const float* A, const float* B, float* X);

/* ---- Householder QR (FP64 and FP32, m >= n, row-major) -----------------------
 * numax's qr_factor[gpu=True] / lstsq[gpu=True] run at float32 only
 * (numax's GPU factorizations are float32); the _f32 entry points are those,
 * the _f64 ones serve the host-precision callers.
 * qr_factor (numax/linalg/qr.mojo:615-743): blocked right-looking geqrf in
 * place -- R on and above the diagonal, the reflectors' V below it (unit
 * diagonal implicit), tau[n] their scales -- with geqr2_panel's dlarfg
 * convention (numax/linalg/panel.mojo:1975-2073) per `block`-wide panel
 * (1..64; numax's default is 16) and the panel's block reflector applied
 * to the trailing columns as C -= V (T^T (V^T C)) (qr.mojo:123-327). */
int nmx_geqrf_f64(nmx_ctx* ctx, int m, int n, double* A, int lda, double* tau, int block);
int nmx_geqrf_f32(nmx_ctx* ctx, int m, int n, float* A, int lda, float* tau, int block);
/* QR.apply_q_transpose (qr.mojo:470-534) when trans != 0: B := Q^T B for an
 * m x nrhs B; trans == 0 applies Q. */
int nmx_ormqr_f64(nmx_ctx* ctx, int trans, int m, int n, int nrhs, const double* QR, int lda, const double* tau,
                  double* B, int ldb, int block);
int nmx_ormqr_f32(nmx_ctx* ctx, int trans, int m, int n, int nrhs, const float* QR, int lda, const float* tau,
                  float* B, int ldb, int block);
/* QR.q (qr.mojo:405-468): the thin m x n Q, the panels applied to the
 * identity last first. */
int nmx_orgqr_f64(nmx_ctx* ctx, int m, int n, const double* QR, int lda, const double* tau, double* Q, int ldq,
                  int block);
int nmx_orgqr_f32(nmx_ctx* ctx, int m, int n, const float* QR, int lda, const float* tau, float* Q, int ldq,
                  int block);
/* QR.r (qr.mojo:371-403): the n x n upper triangle, zeros below. */
int nmx_qr_r_f64(nmx_ctx* ctx, int n, const double* QR, int lda, double* R, int ldr);
int nmx_qr_r_f32(nmx_ctx* ctx, int n, const float* QR, int lda, float* R, int ldr);
/* lstsq(method "qr") and QR.solve (qr.mojo:746-821, :536-612): factors A in
 * place (tau[n]), then B (m x nrhs) := Q^T B and R X = B[0:n] solved in
 * place -- X is B's first n rows. */
int nmx_gels_f64(nmx_ctx* ctx, int m, int n, int nrhs, double* A, int lda, double* tau, double* B, int ldb, int block);
int nmx_gels_f32(nmx_ctx* ctx, int m, int n, int nrhs, float* A, int lda, float* tau, float* B, int ldb, int block);

/* ---- tensor files: .npy and NMX1 ---------------------------------------------
 * numax.io.numpy.save / load / load_dyn (numax/io/npy.mojo:262-453) and
 * numax.io.nmx.save / load (numax/io/io.mojo:59-194): `format` 0 = .npy
 * (read: format 1.x / 2.x / 3.x headers; written: 1.0, padded as
 * numpy.save pads, so the file is byte-identical to numpy.save's), 1 = NMX1
 * ("NMX1", a length-prefixed dtype name, a rank byte, little-endian int64
 * dims, the row-major payload).  dtype codes: 0 float16, 1 float32,
 * 2 float64, 3 int8, 4 int16, 5 int32, 6 int64, 7 uint8, 8 uint16,
 * 9 uint32, 10 uint64, 11 bool; rank <= 8.  The payload streams between
 * the file and device memory through two pinned staging buffers (the file
 * read into one while the other is copied on the context's stream); load
 * and save return when the data has landed (they synchronize the stream).
 * Errors: NMX_ERR_IO, _FORMAT (bad magic, e.g. an .npz archive; unsupported
 * version; truncated header), _DTYPE, _SHAPE (rank, dims, or a payload
 * longer or shorter than the header says), _LAYOUT (Fortran order,
 * big-endian). */
/* The file's dtype, rank and dims (dims: room for 8), no device work:
 * the shape load_dyn takes from the file. */
int nmx_tensor_file_info(const char* path, int format, int* dtype, int* rank, int64_t* dims);
/* Reads the file into device memory dst, which holds prod(dims) elements;
 * dtype, rank and dims must be the file's exactly (numax's typed load). */
int nmx_tensor_file_load(nmx_ctx* ctx, const char* path, int format, int dtype, int rank, const int64_t* dims,
                         void* dst);
/* Writes device memory src (prod(dims) elements, row-major) to the file. */
int nmx_tensor_file_save(nmx_ctx* ctx, const char* path, int format, int dtype, int rank, const int64_t* dims,
                         const void* src);

/* ---- graph capture ----------------------------------------------------------
 * Records the library calls made on the context's stream between
 * nmx_graph_begin and nmx_graph_end into one HIP graph; nmx_graph_launch
 * replays them on that stream with one launch.  Only calls that neither
 * synchronize nor allocate nor advance host-side state can be captured:
 * the batched solvers (whose refinement loops run inside their kernels),
 * the batched eigensolver and least squares, QR, GEMM, TRSM.  A call that
 * would need more workspace than an earlier call left (run it once first),
 * the pivoting LU drivers (their cooperative panels take host-side epochs),
 * the block-of-right-hand-sides and GMRES-IR mixed-precision solves (their
 * loops read norms on the host) return NMX_ERR_UNSUPPORTED while capturing;
 * nmx_gesv_mpir with classical refinement is captured whole.  A live graph keeps every
 * context buffer it reads: a workspace that grows afterwards gets a new
 * buffer and the old one is freed with the context's last graph.  No
 * reference counterpart: for callers that replay the same batch shape
 * (serving loops). */
typedef struct nmx_graph nmx_graph;
int nmx_graph_begin(nmx_ctx* ctx);
int nmx_graph_end(nmx_ctx* ctx, nmx_graph** out);
int nmx_graph_launch(nmx_graph* graph);
int nmx_graph_destroy(nmx_graph* graph);

/* ---- synthetic inputs ------------------------------------------------------
 * Philox4x32-10 stream (the same words as numax_amd.random): kind 0 fills
 * lo + (hi - lo) U[0,1), kind 1 fills N(0,1) by Box--Muller.  Element e of
 * `out` is stream element offset + e. */
int nmx_fill_random_f64(nmx_ctx* ctx, int kind, uint64_t seed, int64_t offset, int64_t count, double lo, double hi,
                        double* out);

#ifdef __cplusplus
}
#endif
#endif /* NUMAX_HIP_H */
