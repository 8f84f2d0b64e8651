/*
 * nmx_oracle.c -- CPU restatement of numax's float64 host algorithms for the
 * dense factor-and-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, and
 * nothing in the numax_amd product links, loads or calls it.
 *
 * Every function restates one routine of the reference (numax 0.3.0, Mojo)
 * and cites the file:line it follows.  Storage is row-major with a leading
 * dimension, pivots are int32 0-based LAPACK ipiv, and a zero pivot is
 * divided as 1e-30 (numax/linalg/common.mojo:27-35).
 *
 * What the reference leaves to third-party code (MAX linalg.matmul, max-core
 * 26.6.0) is restated as "dot product in ascending k, then subtract from the
 * destination", which is what the matmul-with-subtracting-epilogue call sites
 * compute (numax/linalg/lu.mojo:646-673).  MAX's internal accumulation order
 * is not observable from the reference sources, so those GEMMs match numax to
 * rounding, not bit for bit.
 *
 * Parallel loops follow numax's `parallelize` split (rows or columns are
 * independent), so the thread count changes speed, never results.
 *
 * Parity pinning (see tests/test_oracle.py): LAPACK dgetrf/dpotrf/dgeqrf via
 * SciPy goldens in tests/golden/, the reference's own known-answer tests,
 * exact rational brute force on small integer matrices, and invariants
 * (P A = L U, L L^T = A, Q^T Q = I, trace / Frobenius identities for eigh).
 */
#include <math.h>
#include <immintrin.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define PIVOT_FLOOR 1e-30          /* numax/linalg/common.mojo:27-29 */
#define PARALLEL_MIN_WORK (1 << 15) /* numax/linalg/panel.mojo:86-101 */
#define SIMD_LANES 4               /* float64 width of an AVX2 host (trsm_right_lower_t) */

#define AT(a, lda, i, j) ((a)[(int64_t)(i) * (lda) + (j)])

/* ------------------------------------------------------------------------ */
/* A tiny parallel-for: numax's `parallelize` over independent rows/cols.   */
/* ------------------------------------------------------------------------ */
static int g_threads = 1;

int orc_set_threads(int t) {
  g_threads = t < 1 ? 1 : (t > 256 ? 256 : t);
  return g_threads;
}

typedef void (*range_fn)(void* ctx, int64_t lo, int64_t hi);
typedef struct { range_fn fn; void* ctx; int64_t lo, hi; } par_task;

static void* par_entry(void* p) {
  par_task* t = (par_task*)p;
  t->fn(t->ctx, t->lo, t->hi);
  return NULL;
}

static void parallel_for(int64_t count, int64_t work, range_fn fn, void* ctx) {
  int nt = g_threads;
  if (count <= 0) return;
  if (nt <= 1 || work < PARALLEL_MIN_WORK || count < 2) { fn(ctx, 0, count); return; }
  if (nt > count) nt = (int)count;
  pthread_t th[256];
  par_task tasks[256];
  int64_t chunk = (count + nt - 1) / nt;
  int launched = 0;
  for (int t = 0; t < nt; t++) {
    int64_t lo = t * chunk, hi = lo + chunk < count ? lo + chunk : count;
    if (lo >= hi) break;
    tasks[t] = (par_task){fn, ctx, lo, hi};
    if (t == 0) continue;
    if (pthread_create(&th[t], NULL, par_entry, &tasks[t]) != 0) { fn(ctx, lo, hi); tasks[t].fn = NULL; }
    launched = t + 1;
  }
  fn(ctx, tasks[0].lo, tasks[0].hi);
  for (int t = 1; t < launched; t++)
    if (tasks[t].fn) pthread_join(th[t], NULL);
}

/* guard_nonzero: numax/core/numeric.mojo:341-364 -- max(|x|, floor) with x's
 * sign, +floor at +0 (copysign's convention), NaN propagates. */
static inline double guard_nonzero(double x, double floor_) {
  if (isnan(x)) return x;
  double m = fabs(x) >= floor_ ? fabs(x) : floor_;
  return copysign(m, x);
}

/* ------------------------------------------------------------------------ */
/* C -= A * B (or A * B^T): the matmul + subtracting epilogue of the        */
/* trailing updates (lu.mojo:646-673, panel.mojo:557-582,                   */
/* triangular.mojo:193-218, cholesky.mojo:126-225).  Each element is        */
/* c - sum_k a[i,k] b[k,j], the sum ascending in k.                         */
/* ------------------------------------------------------------------------ */
typedef struct {
  int m, n, k;
  const double* a; int lda;
  const double* b; int ldb; int trans_b;
  double* c; int ldc;
} gemm_ctx;

/* One element: c - sum_k a[i,k] b[k,j], products added in ascending k. */
static inline double gemm_elem(const gemm_ctx* g, int64_t i, int j) {
  double acc = 0.0;
  for (int p = 0; p < g->k; p++)
    acc += AT(g->a, g->lda, i, p) * (g->trans_b ? AT(g->b, g->ldb, j, p) : AT(g->b, g->ldb, p, j));
  return AT(g->c, g->ldc, i, j) - acc;
}

/* Rows lo..hi in blocks of 4 rows x 8 columns held in AVX2 registers; every
 * element keeps gemm_elem's operation order (a separate multiply and add per
 * k, ascending -- the file is built with -ffp-contract=off), so the results
 * are the same bits as the scalar loop.  B^T is read through a packed copy
 * of the k x 8 panel. */
static void gemm_sub_rows(void* p, int64_t lo, int64_t hi) {
  gemm_ctx* g = (gemm_ctx*)p;
  const int k = g->k;
  double* bp = g->trans_b ? (double*)malloc(sizeof(double) * 8 * (size_t)(k > 0 ? k : 1)) : NULL;
  const int n8 = g->n & ~7;
  for (int j0 = 0; j0 < n8; j0 += 8) {
    const double* bcol;
    int ldb8;
    if (g->trans_b) {
      for (int q = 0; q < k; q++)
        for (int j = 0; j < 8; j++) bp[(size_t)q * 8 + j] = AT(g->b, g->ldb, j0 + j, q);
      bcol = bp;
      ldb8 = 8;
    } else {
      bcol = &AT(g->b, g->ldb, 0, j0);
      ldb8 = g->ldb;
    }
    int64_t i = lo;
    for (; i + 4 <= hi; i += 4) {
      __m256d acc[4][2];
      for (int r = 0; r < 4; r++) acc[r][0] = acc[r][1] = _mm256_setzero_pd();
      for (int q = 0; q < k; q++) {
        const __m256d b0 = _mm256_loadu_pd(bcol + (size_t)q * ldb8);
        const __m256d b1 = _mm256_loadu_pd(bcol + (size_t)q * ldb8 + 4);
        for (int r = 0; r < 4; r++) {
          const __m256d av = _mm256_set1_pd(AT(g->a, g->lda, i + r, q));
          acc[r][0] = _mm256_add_pd(acc[r][0], _mm256_mul_pd(av, b0));
          acc[r][1] = _mm256_add_pd(acc[r][1], _mm256_mul_pd(av, b1));
        }
      }
      for (int r = 0; r < 4; r++) {
        double* crow = &AT(g->c, g->ldc, i + r, j0);
        _mm256_storeu_pd(crow, _mm256_sub_pd(_mm256_loadu_pd(crow), acc[r][0]));
        _mm256_storeu_pd(crow + 4, _mm256_sub_pd(_mm256_loadu_pd(crow + 4), acc[r][1]));
      }
    }
    for (; i < hi; i++)
      for (int j = j0; j < j0 + 8; j++) AT(g->c, g->ldc, i, j) = gemm_elem(g, i, j);
  }
  for (int64_t i = lo; i < hi; i++)
    for (int j = n8; j < g->n; j++) AT(g->c, g->ldc, i, j) = gemm_elem(g, i, j);
  free(bp);
}

void orc_gemm_sub(int m, int n, int k, const double* a, int lda, const double* b, int ldb, int trans_b,
                  double* c, int ldc) {
  if (m <= 0 || n <= 0) return;
  gemm_ctx g = {m, n, k, a, lda, b, ldb, trans_b, c, ldc};
  parallel_for(m, (int64_t)m * n * (k > 0 ? k : 1) * 2, gemm_sub_rows, &g);
}

/* ------------------------------------------------------------------------ */
/* LU with partial pivoting                                                 */
/* ------------------------------------------------------------------------ */

/* getrf_panel, host path (panel.mojo:289-431 with _lanes = 1): first-max
 * scan of |a[i,col]| for i >= col with strict '>' starting from magnitude 0,
 * full-row swap across all n columns, 1e-30 divisor on an exact-zero pivot,
 * true division for the multipliers and the rank-1 update inside the panel. */
static void getrf_panel(double* a, int lda, int32_t* ipiv, int32_t* info, int k, int nb, int n) {
  for (int j = 0; j < nb; j++) {
    int col = k + j;
    int best = col;
    double best_mag = 0.0;
    for (int i = col; i < n; i++) {
      double mag = fabs(AT(a, lda, i, col));
      if (mag > best_mag) { best_mag = mag; best = i; }
    }
    ipiv[col] = best;
    if (fabs(AT(a, lda, best, col)) == 0.0 && *info == 0) *info = col + 1;
    if (best != col) {
      for (int c = 0; c < n; c++) {
        double keep = AT(a, lda, col, c);
        AT(a, lda, col, c) = AT(a, lda, best, c);
        AT(a, lda, best, c) = keep;
      }
    }
    double pivot = AT(a, lda, col, col);
    if (pivot == 0.0) pivot = PIVOT_FLOOR;
    for (int i = col + 1; i < n; i++) {
      double mult = AT(a, lda, i, col) / pivot;
      AT(a, lda, i, col) = mult;
      for (int c = j + 1; c < nb; c++) AT(a, lda, i, k + c) = AT(a, lda, i, k + c) - mult * AT(a, lda, col, k + c);
    }
  }
}

/* trsm_left_lower_unit (panel.mojo:2411-2463): Y := L^-1 Y for
 * Y = a[k:k+nb, k+nb:cols], one column at a time, sequential subtraction. */
typedef struct { double* a; int lda, k, nb, cols; } trsm_llu_ctx;
static void trsm_llu_cols(void* p, int64_t lo, int64_t hi) {
  trsm_llu_ctx* t = (trsm_llu_ctx*)p;
  for (int64_t idx = lo; idx < hi; idx++) {
    int col = t->k + t->nb + (int)idx;
    for (int i = 0; i < t->nb; i++) {
      double total = AT(t->a, t->lda, t->k + i, col);
      for (int q = 0; q < i; q++) total = total - AT(t->a, t->lda, t->k + i, t->k + q) * AT(t->a, t->lda, t->k + q, col);
      AT(t->a, t->lda, t->k + i, col) = total;
    }
  }
}
static void trsm_left_lower_unit(double* a, int lda, int k, int nb, int cols) {
  int width = cols - k - nb;
  if (width <= 0) return;
  trsm_llu_ctx t = {a, lda, k, nb, cols};
  parallel_for(width, (int64_t)width * nb * nb, trsm_llu_cols, &t);
}

/* getrf2 (panel.mojo:445-586): split the panel until it is `base` wide (or
 * narrower than 4), solve the left half's effect on the right half with a
 * trsm over the panel's own right half and a GEMM below it. */
static void getrf2(double* a, int lda, int32_t* ipiv, int32_t* info, int k, int nb, int n, int base) {
  if (nb <= base || nb < 4) { getrf_panel(a, lda, ipiv, info, k, nb, n); return; }
  int half = nb / 2;
  getrf2(a, lda, ipiv, info, k, half, n, base);
  trsm_left_lower_unit(a, lda, k, half, k + nb);
  int rows = n - k - half, cols = nb - half;
  if (rows > 0 && cols > 0) {
    int r0 = k + half;
    orc_gemm_sub(rows, cols, half, &AT(a, lda, r0, k), lda, &AT(a, lda, k, r0), lda, 0, &AT(a, lda, r0, r0), lda);
  }
  getrf2(a, lda, ipiv, info, k + half, nb - half, n, base);
}

/* lu_factor (lu.mojo:454-701): right-looking blocked GEPP.  Returns the swap
 * parity sign (lu.mojo:691-699). */
int orc_getrf(int n, double* a, int lda, int32_t* ipiv, int32_t* info, int block, int base) {
  *info = 0;
  if (block < 1) block = 32;
  if (base < 1) base = 16;
  for (int k = 0; k < n; k += block) {
    int nb = block < n - k ? block : n - k;
    getrf2(a, lda, ipiv, info, k, nb, n, base);
    trsm_left_lower_unit(a, lda, k, nb, n);
    int m = n - k - nb;
    if (m > 0) {
      int b0 = k + nb;
      orc_gemm_sub(m, m, nb, &AT(a, lda, b0, k), lda, &AT(a, lda, k, b0), lda, 0, &AT(a, lda, b0, b0), lda);
    }
  }
  int sign = 1;
  for (int j = 0; j < n; j++)
    if (ipiv[j] != j) sign = -sign;
  return sign;
}

/* laswp (panel.mojo:1824-1852) and laswp_matrix (:2279-2313): replay the
 * interchanges in order j = 0..n-1. */
void orc_laswp(int n, int nrhs, double* b, int ldb, const int32_t* ipiv) {
  for (int j = 0; j < n; j++) {
    int other = ipiv[j];
    if (other != j) {
      for (int c = 0; c < nrhs; c++) {
        double keep = AT(b, ldb, j, c);
        AT(b, ldb, j, c) = AT(b, ldb, other, c);
        AT(b, ldb, other, c) = keep;
      }
    }
  }
}

static inline double at_t(const double* a, int lda, int trans, int i, int j) {
  return trans ? AT(a, lda, j, i) : AT(a, lda, i, j);
}

/* trsv_diag (panel.mojo:1855-1907) + gemv_sub (:1910-1972) inside _trsv
 * (triangular.mojo:43-111): blocked vector triangular solve, the upper walk
 * from the bottom with a ragged top block. */
void orc_trsv(int n, const double* a, int lda, double* x, int upper, int unit, int trans, int block) {
  if (block < 1) block = 16;
  int steps = (n + block - 1) / block;
  for (int step = 0; step < steps; step++) {
    int k, nb = block;
    if (upper) {
      k = n - (step + 1) * block;
      if (k < 0) { nb = block + k; k = 0; }
    } else {
      k = step * block;
      nb = block < n - k ? block : n - k;
    }
    for (int s = 0; s < nb; s++) {
      int i = upper ? nb - 1 - s : s;
      double total = x[k + i];
      if (upper) {
        for (int p = i + 1; p < nb; p++) total = total - at_t(a, lda, trans, k + i, k + p) * x[k + p];
      } else {
        for (int p = 0; p < i; p++) total = total - at_t(a, lda, trans, k + i, k + p) * x[k + p];
      }
      x[k + i] = unit ? total : total / AT(a, lda, k + i, k + i);
    }
    int row0 = upper ? 0 : k + nb, rows = upper ? k : n - k - nb;
    for (int r = 0; r < rows; r++) {
      int row = row0 + r;
      double total = x[row];
      for (int j = 0; j < nb; j++) total = total - at_t(a, lda, trans, row, k + j) * x[k + j];
      x[row] = total;
    }
  }
}

/* _trsm (triangular.mojo:114-225) with trsm_diag (panel.mojo:2221-2276):
 * per block a column-wise substitution, then the rows still unsolved get
 * `b -= A_block * b[k:k+nb]` through the GEMM. */
void orc_trsm(int n, int nrhs, const double* a, int lda, double* b, int ldb, int upper, int unit, int trans,
              int block) {
  if (n <= 0 || nrhs <= 0) return;
  if (block < 1) block = 16;
  double* op = (double*)malloc(sizeof(double) * (size_t)n * block);
  int steps = (n + block - 1) / block;
  for (int step = 0; step < steps; step++) {
    int k, nb = block;
    if (upper) {
      k = n - (step + 1) * block;
      if (k < 0) { nb = block + k; k = 0; }
    } else {
      k = step * block;
      nb = block < n - k ? block : n - k;
    }
    for (int c = 0; c < nrhs; c++) {
      for (int s = 0; s < nb; s++) {
        int i = upper ? nb - 1 - s : s;
        double total = AT(b, ldb, k + i, c);
        if (upper) {
          for (int p = i + 1; p < nb; p++) total = total - at_t(a, lda, trans, k + i, k + p) * AT(b, ldb, k + p, c);
        } else {
          for (int p = 0; p < i; p++) total = total - at_t(a, lda, trans, k + i, k + p) * AT(b, ldb, k + p, c);
        }
        AT(b, ldb, k + i, c) = unit ? total : total / AT(a, lda, k + i, k + i);
      }
    }
    int row0 = upper ? 0 : k + nb, rows = upper ? k : n - k - nb;
    if (rows <= 0) continue;
    for (int r = 0; r < rows; r++)
      for (int j = 0; j < nb; j++) op[(size_t)r * nb + j] = at_t(a, lda, trans, row0 + r, k + j);
    orc_gemm_sub(rows, nrhs, nb, op, nb, &AT(b, ldb, k, 0), ldb, 0, &AT(b, ldb, row0, 0), ldb);
  }
  free(op);
}

/* LU.solve (lu.mojo:181-354): laswp then forward (unit) and back
 * substitution; the vector form for one right-hand side (_trsv), the matrix
 * form otherwise (_trsm). */
void orc_getrs(int n, int nrhs, const double* lu, int lda, const int32_t* ipiv, double* b, int ldb, int block) {
  orc_laswp(n, nrhs, b, ldb, ipiv);
  if (nrhs == 1 && ldb == 1) {
    orc_trsv(n, lu, lda, b, 0, 1, 0, block);
    orc_trsv(n, lu, lda, b, 1, 0, 0, block);
  } else {
    orc_trsm(n, nrhs, lu, lda, b, ldb, 0, 1, 0, block);
    orc_trsm(n, nrhs, lu, lda, b, ldb, 1, 0, 0, block);
  }
}

/* ------------------------------------------------------------------------ */
/* Cholesky                                                                 */
/* ------------------------------------------------------------------------ */

/* potrf_diag (panel.mojo:216-286): unblocked right-looking Cholesky of the
 * nb x nb block at (k,k), lower triangle only, d <= 0 floored to 1e-30 with
 * the first failure recorded 1-based in info. */
static void potrf_diag(double* a, int lda, int32_t* info, int k0, int nb) {
  for (int j = 0; j < nb; j++) {
    double d = AT(a, lda, k0 + j, k0 + j);
    if (d <= 0) {
      if (*info == 0) *info = k0 + j + 1;
      d = PIVOT_FLOOR;
    }
    AT(a, lda, k0 + j, k0 + j) = sqrt(d);
    double root = AT(a, lda, k0 + j, k0 + j);
    for (int i = j + 1; i < nb; i++) AT(a, lda, k0 + i, k0 + j) = AT(a, lda, k0 + i, k0 + j) / root;
    for (int i = j + 1; i < nb; i++) {
      double lij = AT(a, lda, k0 + i, k0 + j);
      for (int c = j + 1; c <= i; c++)
        AT(a, lda, k0 + i, k0 + c) = AT(a, lda, k0 + i, k0 + c) - lij * AT(a, lda, k0 + c, k0 + j);
    }
  }
}

/* trsm_right_lower_t (panel.mojo:2316-2408): X := X L^-T row by row; the
 * prefix dot is accumulated in SIMD_LANES partial sums, reduced pairwise,
 * subtracted, then the remainder subtracted sequentially. */
typedef struct { double* a; int lda, k, nb; } trsm_rlt_ctx;
static void trsm_rlt_rows(void* p, int64_t lo, int64_t hi) {
  trsm_rlt_ctx* t = (trsm_rlt_ctx*)p;
  for (int64_t idx = lo; idx < hi; idx++) {
    int row = t->k + t->nb + (int)idx;
    for (int j = 0; j < t->nb; j++) {
      double acc[SIMD_LANES] = {0};
      int q = 0;
      while (q + SIMD_LANES <= j) {
        for (int l = 0; l < SIMD_LANES; l++)
          acc[l] += AT(t->a, t->lda, row, t->k + q + l) * AT(t->a, t->lda, t->k + j, t->k + q + l);
        q += SIMD_LANES;
      }
      double red = (acc[0] + acc[1]) + (acc[2] + acc[3]);
      double total = AT(t->a, t->lda, row, t->k + j) - red;
      while (q < j) {
        total = total - AT(t->a, t->lda, row, t->k + q) * AT(t->a, t->lda, t->k + j, t->k + q);
        q++;
      }
      AT(t->a, t->lda, row, t->k + j) = total / AT(t->a, t->lda, t->k + j, t->k + j);
    }
  }
}

/* _potrf_blocked (cholesky.mojo:59-230): potrf_diag, the panel solve, and the
 * lower block triangle of A22 -= L21 L21^T in tile x tile pieces; a ragged
 * last column tile narrower than 2 takes the elementwise `narrow` form. */
int orc_potrf(int n, double* a, int lda, int32_t* info, int block, int tile) {
  *info = 0;
  if (block < 1) block = 64;
  if (tile < 1) tile = 128;
  double* panel = (double*)malloc(sizeof(double) * (size_t)n * block);
  for (int k = 0; k < n; k += block) {
    int nb = block < n - k ? block : n - k;
    potrf_diag(a, lda, info, k, nb);
    int height = n - k - nb;
    if (height > 0) {
      trsm_rlt_ctx t = {a, lda, k, nb};
      parallel_for(height, (int64_t)height * nb * nb, trsm_rlt_rows, &t);
    }
    int m = n - k - nb;
    if (m <= 0) continue;
    int base = k + nb;
    for (int r = 0; r < m; r++)
      for (int c = 0; c < nb; c++) panel[(size_t)r * nb + c] = AT(a, lda, base + r, k + c);
    int tiles = (m + tile - 1) / tile;
    for (int ti = 0; ti < tiles; ti++) {
      int row0 = ti * tile, rows = tile < m - row0 ? tile : m - row0;
      for (int tj = 0; tj <= ti; tj++) {
        int col0 = tj * tile, cols = tile < m - col0 ? tile : m - col0;
        if (cols < 2) {
          for (int i = 0; i < rows; i++)
            for (int j = 0; j < cols; j++) {
              double total = AT(a, lda, base + row0 + i, base + col0 + j);
              for (int q = 0; q < nb; q++) total -= panel[(size_t)(row0 + i) * nb + q] * panel[(size_t)(col0 + j) * nb + q];
              AT(a, lda, base + row0 + i, base + col0 + j) = total;
            }
          continue;
        }
        orc_gemm_sub(rows, cols, nb, panel + (size_t)row0 * nb, nb, panel + (size_t)col0 * nb, nb, 1,
                     &AT(a, lda, base + row0, base + col0), lda);
      }
    }
  }
  free(panel);
  /* tril: the returned factor has zeros above the diagonal (cholesky.mojo:385-396). */
  for (int i = 0; i < n; i++)
    for (int j = i + 1; j < n; j++) AT(a, lda, i, j) = 0.0;
  return *info;
}

/* cholesky_solve (cholesky.mojo:471-580): L y = b then L^T x = y, the second
 * reading L transposed in place. */
void orc_potrs(int n, int nrhs, const double* l, int lda, double* b, int ldb, int block) {
  if (nrhs == 1 && ldb == 1) {
    orc_trsv(n, l, lda, b, 0, 0, 0, block);
    orc_trsv(n, l, lda, b, 1, 0, 1, block);
  } else {
    orc_trsm(n, nrhs, l, lda, b, ldb, 0, 0, 0, block);
    orc_trsm(n, nrhs, l, lda, b, ldb, 1, 0, 1, block);
  }
}

/* ------------------------------------------------------------------------ */
/* Householder QR                                                           */
/* ------------------------------------------------------------------------ */

/* geqr2_panel (panel.mojo:1975-2073): LAPACK dlarfg convention. */
static void geqr2_panel(double* a, int lda, double* tau, int k0, int nb, int m) {
  for (int j = 0; j < nb; j++) {
    int col = k0 + j;
    double below = 0.0;
    for (int i = col + 1; i < m; i++) below += AT(a, lda, i, col) * AT(a, lda, i, col);
    double alpha = AT(a, lda, col, col);
    if (below == 0.0) { tau[col] = 0.0; continue; }
    double beta = sqrt(alpha * alpha + below);
    if (alpha > 0) beta = -beta;
    double this_tau = (beta - alpha) / beta;
    double scale = 1.0 / (alpha - beta);
    for (int i = col + 1; i < m; i++) AT(a, lda, i, col) = AT(a, lda, i, col) * scale;
    AT(a, lda, col, col) = beta;
    tau[col] = this_tau;
    for (int c = j + 1; c < nb; c++) {
      int other = k0 + c;
      double dot = AT(a, lda, col, other);
      for (int r = col + 1; r < m; r++) dot += AT(a, lda, r, col) * AT(a, lda, r, other);
      double factor = this_tau * dot;
      AT(a, lda, col, other) = AT(a, lda, col, other) - factor;
      for (int r = col + 1; r < m; r++) AT(a, lda, r, other) = AT(a, lda, r, other) - factor * AT(a, lda, r, col);
    }
  }
}

/* larft_panel (panel.mojo:2076-2169): forward columnwise T, built as T^T
 * (lower) when `transposed`. t is nb x nb dense. */
static void larft_panel(const double* a, int lda, const double* tau, double* t, int k0, int nb, int m, int transposed) {
#define TAT(r, c) (*(transposed ? &t[(size_t)(c) * nb + (r)] : &t[(size_t)(r) * nb + (c)]))
  for (int e = 0; e < nb * nb; e++) t[e] = 0.0;
  TAT(0, 0) = tau[k0];
  for (int i = 1; i < nb; i++) {
    double ti = tau[k0 + i];
    for (int p = 0; p < i; p++) {
      double total = AT(a, lda, k0 + i, k0 + p);
      for (int r = k0 + i + 1; r < m; r++) total += AT(a, lda, r, k0 + p) * AT(a, lda, r, k0 + i);
      TAT(p, i) = -ti * total;
    }
    for (int q = 0; q < i; q++) {
      double total = 0.0;
      for (int r = q; r < i; r++) total += TAT(q, r) * TAT(r, i);
      TAT(q, i) = total;
    }
    TAT(i, i) = ti;
  }
#undef TAT
}

/* _apply_block_reflector (qr.mojo:123-327): C := (I - V T^T V^T) C with
 * W = V^T C, Y = T^T W, C -= V Y; sums ascending. */
static void apply_block_reflector_t(const double* a, int lda, const double* tau, double* c, int ldc, int k, int nb,
                                    int m, int rows, int width, int row0, int col0) {
  if (rows <= 0 || width <= 0 || nb <= 0) return;
  double* t = (double*)malloc(sizeof(double) * nb * nb);
  double* v = (double*)malloc(sizeof(double) * (size_t)rows * nb);
  double* w = (double*)malloc(sizeof(double) * (size_t)nb * width);
  double* y = (double*)malloc(sizeof(double) * (size_t)nb * width);
  larft_panel(a, lda, tau, t, k, nb, m, 1);
  for (int i = 0; i < rows; i++)
    for (int j = 0; j < nb; j++) v[(size_t)i * nb + j] = i == j ? 1.0 : (i > j ? AT(a, lda, k + i, k + j) : 0.0);
  for (int p = 0; p < nb; p++)
    for (int cc = 0; cc < width; cc++) {
      double s = 0.0;
      for (int r = 0; r < rows; r++) s += v[(size_t)r * nb + p] * AT(c, ldc, row0 + r, col0 + cc);
      w[(size_t)p * width + cc] = s;
    }
  for (int p = 0; p < nb; p++)
    for (int cc = 0; cc < width; cc++) {
      double s = 0.0;
      for (int q = 0; q < nb; q++) s += t[(size_t)p * nb + q] * w[(size_t)q * width + cc];
      y[(size_t)p * width + cc] = s;
    }
  orc_gemm_sub(rows, width, nb, v, nb, y, width, 0, &AT(c, ldc, row0, col0), ldc);
  free(t); free(v); free(w); free(y);
}

/* qr_factor (qr.mojo:615-743): blocked geqrf, m >= n. */
void orc_geqrf(int m, int n, double* a, int lda, double* tau, int block) {
  if (block < 1) block = 16;
  for (int k = 0; k < n; k += block) {
    int nb = block < n - k ? block : n - k;
    geqr2_panel(a, lda, tau, k, nb, m);
    apply_block_reflector_t(a, lda, tau, a, lda, k, nb, m, m - k, n - k - nb, k, k + nb);
  }
}

/* QR.apply_q_transpose (qr.mojo:470-534): Q^T B for an m x nrhs B. */
void orc_ormqr_t(int m, int n, int nrhs, const double* qr, int lda, const double* tau, double* b, int ldb, int block) {
  if (block < 1) block = 16;
  for (int k = 0; k < n; k += block) {
    int nb = block < n - k ? block : n - k;
    apply_block_reflector_t(qr, lda, tau, b, ldb, k, nb, m, m - k, nrhs, k, 0);
  }
}

/* QR.solve / lstsq (qr.mojo:536-612, :746-821): Q^T b, first n entries,
 * then the upper blocked substitution of R (block 16).  b is length m and
 * is overwritten; x (length n) receives the solution. */
void orc_gels(int m, int n, double* a, int lda, double* tau, double* b, double* x, int block) {
  orc_geqrf(m, n, a, lda, tau, block);
  orc_ormqr_t(m, n, 1, a, lda, tau, b, 1, block);
  double* r = (double*)calloc((size_t)n * n, sizeof(double));
  for (int i = 0; i < n; i++)
    for (int j = i; j < n; j++) r[(size_t)i * n + j] = AT(a, lda, i, j);
  for (int i = 0; i < n; i++) x[i] = b[i];
  orc_trsv(n, r, n, x, 1, 0, 0, 16);
  free(r);
}

/* ------------------------------------------------------------------------ */
/* Register (Array) tier: one small problem, the algorithms map_blocks runs */
/* per lane.                                                                */
/* ------------------------------------------------------------------------ */

/* lu_factor / PivotedLU (_array/lu.mojo:132-216): first-max from the
 * diagonal's own magnitude, full-row swaps, permutation[i] = source row,
 * pivot = guard_nonzero(a_kk, 1e-30). Returns the sign. */
int orc_array_lu_factor(int n, double* a, int32_t* perm) {
  int sign = 1;
  for (int i = 0; i < n; i++) perm[i] = i;
  for (int k = 0; k < n; k++) {
    int best = k;
    double best_mag = fabs(a[k * n + k]);
    for (int i = k + 1; i < n; i++) {
      double mag = fabs(a[i * n + k]);
      if (mag > best_mag) { best = i; best_mag = mag; }
    }
    if (best != k) {
      for (int j = 0; j < n; j++) { double s = a[k * n + j]; a[k * n + j] = a[best * n + j]; a[best * n + j] = s; }
      int s = perm[k]; perm[k] = perm[best]; perm[best] = s;
      sign = -sign;
    }
    double pivot = guard_nonzero(a[k * n + k], PIVOT_FLOOR);
    for (int i = k + 1; i < n; i++) {
      double f = a[i * n + k] / pivot;
      a[i * n + k] = f;
      for (int j = k + 1; j < n; j++) a[i * n + j] = a[i * n + j] - f * a[k * n + j];
    }
  }
  return sign;
}

/* forward_substitution / back_substitution (_array/triangular.mojo:24-63). */
static void array_forward(int n, const double* l, const double* b, double* x, int unit) {
  for (int i = 0; i < n; i++) {
    double total = b[i];
    for (int j = 0; j < i; j++) total = total - l[i * n + j] * x[j];
    x[i] = unit ? total : total / guard_nonzero(l[i * n + i], PIVOT_FLOOR);
  }
}
static void array_back(int n, const double* u, const double* b, double* x) {
  for (int s = 0; s < n; s++) {
    int i = n - 1 - s;
    double total = b[i];
    for (int j = i + 1; j < n; j++) total = total - u[i * n + j] * x[j];
    x[i] = total / guard_nonzero(u[i * n + i], PIVOT_FLOOR);
  }
}

/* PivotedLU.solve (_array/lu.mojo:85-103). */
void orc_array_lu_solve(int n, const double* lu, const int32_t* perm, const double* b, double* x) {
  double* pb = (double*)malloc(sizeof(double) * n);
  double* y = (double*)malloc(sizeof(double) * n);
  for (int i = 0; i < n; i++) pb[i] = b[perm[i]];
  array_forward(n, lu, pb, y, 1);
  array_back(n, lu, y, x);
  free(pb); free(y);
}

/* cholesky (_array/cholesky.mojo:26-58): left-looking, guard then sqrt (a
 * negative diagonal gives NaN because the guard keeps the sign). */
void orc_array_cholesky(int n, const double* a, double* l) {
  for (int i = 0; i < n * n; i++) l[i] = 0.0;
  for (int j = 0; j < n; j++) {
    double diag = a[j * n + j];
    for (int k = 0; k < j; k++) diag = diag - l[j * n + k] * l[j * n + k];
    double ljj = sqrt(guard_nonzero(diag, PIVOT_FLOOR));
    l[j * n + j] = ljj;
    for (int i = j + 1; i < n; i++) {
      double total = a[i * n + j];
      for (int k = 0; k < j; k++) total = total - l[i * n + k] * l[j * n + k];
      l[i * n + j] = total / ljj;
    }
  }
}

/* cholesky_solve (_array/cholesky.mojo:61-80). */
void orc_array_cholesky_solve(int n, const double* l, const double* b, double* x) {
  double* y = (double*)malloc(sizeof(double) * n);
  double* lt = (double*)malloc(sizeof(double) * n * n);
  array_forward(n, l, b, y, 0);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) lt[i * n + j] = l[j * n + i];
  array_back(n, lt, y, x);
  free(y); free(lt);
}

/* solve (_array/basic.mojo:20-31): unpivoted LU and two substitutions. */
void orc_array_solve_nopivot(int n, const double* a, const double* b, double* x) {
  double* f = (double*)malloc(sizeof(double) * n * n);
  double* y = (double*)malloc(sizeof(double) * n);
  memcpy(f, a, sizeof(double) * n * n);
  for (int k = 0; k < n; k++) {
    double pivot = guard_nonzero(f[k * n + k], PIVOT_FLOOR);
    for (int i = k + 1; i < n; i++) {
      double fac = f[i * n + k] / pivot;
      f[i * n + k] = fac;
      for (int j = k + 1; j < n; j++) f[i * n + j] = f[i * n + j] - fac * f[k * n + j];
    }
  }
  array_forward(n, f, b, y, 1);
  array_back(n, f, y, x);
  free(f); free(y);
}

/* lu (_array/lu.mojo:25-55): unpivoted Doolittle in place, packed L\U. */
void orc_array_lu_nopivot(int n, double* f) {
  for (int k = 0; k < n; k++) {
    double pivot = guard_nonzero(f[k * n + k], PIVOT_FLOOR);
    for (int i = k + 1; i < n; i++) {
      double fac = f[i * n + k] / pivot;
      f[i * n + k] = fac;
      for (int j = k + 1; j < n; j++) f[i * n + j] = f[i * n + j] - fac * f[k * n + j];
    }
  }
}

/* inverse (_array/basic.mojo:34-60): unpivoted lu, then forward and back
 * substitution of each column of the identity. */
void orc_array_inverse_nopivot(int n, const double* a, double* out) {
  double* f = (double*)malloc(sizeof(double) * n * n);
  double* e = (double*)malloc(sizeof(double) * n);
  double* y = (double*)malloc(sizeof(double) * n);
  double* x = (double*)malloc(sizeof(double) * n);
  memcpy(f, a, sizeof(double) * n * n);
  orc_array_lu_nopivot(n, f);
  for (int c = 0; c < n; c++) {
    for (int i = 0; i < n; i++) e[i] = i == c ? 1.0 : 0.0;
    array_forward(n, f, e, y, 1);
    array_back(n, f, y, x);
    for (int i = 0; i < n; i++) out[i * n + c] = x[i];
  }
  free(f); free(e); free(y); free(x);
}

/* det (_array/lu.mojo:219-233, pivoted = 0): the unpivoted lu's diagonal
 * product; PivotedLU.det (:112-123, pivoted = 1): sign times the pivoted
 * factor's diagonal product, multiplied in row order. */
double orc_array_det(int n, const double* a, int pivoted) {
  double* f = (double*)malloc(sizeof(double) * n * n);
  int32_t* perm = (int32_t*)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
  memcpy(f, a, sizeof(double) * n * n);
  double product;
  if (pivoted) {
    product = (double)orc_array_lu_factor(n, f, perm);
  } else {
    orc_array_lu_nopivot(n, f);
    product = 1.0;
  }
  for (int i = 0; i < n; i++) product = product * f[i * n + i];
  free(f); free(perm);
  return product;
}

/* PivotedLU.slogdet (_array/lu.mojo:125-162) of lu_factor(a): sign 0 and
 * -inf at the first exact zero on the diagonal. */
void orc_array_slogdet(int n, const double* a, double* sign_out, double* logabs_out) {
  double* f = (double*)malloc(sizeof(double) * n * n);
  int32_t* perm = (int32_t*)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
  memcpy(f, a, sizeof(double) * n * n);
  double sign = (double)orc_array_lu_factor(n, f, perm), total = 0.0;
  for (int i = 0; i < n; i++) {
    double entry = f[i * n + i];
    if (entry == 0.0) { sign = 0.0; total = log(0.0); break; }
    if (entry < 0.0) sign = -sign;
    total += log(fabs(entry));
  }
  *sign_out = sign;
  *logabs_out = total;
  free(f); free(perm);
}

/* slogdet_cholesky (_array/cholesky.mojo:95-116): 2 * sum ln(L_ii). */
double orc_array_slogdet_cholesky(int n, const double* l) {
  double total = 0.0;
  for (int i = 0; i < n; i++) total = total + log(l[i * n + i]);
  return total * 2.0;
}

/* sytf2_lower (numax/linalg/panel.mojo:2737-2894): Bunch-Kaufman L D L^T of
 * the lower triangle in place, alpha = (1 + sqrt(17)) / 8, first-maximum
 * scans, ipiv 1-based (kp + 1, or -(kp + 1) on both rows of a 2 x 2 block).
 * Returns the first k + 1 whose column is exactly zero (0: none). */
int orc_sytf2_lower(int n, double* a, int32_t* ipiv) {
  const double alpha = 0.6403882032022076;
  int info = 0;
  int k = 0;
#define A_(i, j) a[(int64_t)(i) * n + (j)]
  while (k < n) {
    int kstep = 1, kp = k, imax = k;
    double absakk = fabs(A_(k, k)), colmax = 0.0;
    for (int i = k + 1; i < n; i++) {
      double v = fabs(A_(i, k));
      if (v > colmax) { colmax = v; imax = i; }
    }
    if (absakk == 0.0 && colmax == 0.0) {
      kp = k;
      if (info == 0) info = k + 1;
    } else if (absakk >= alpha * colmax) {
      kp = k;
    } else {
      double rowmax = 0.0;
      for (int j = k; j < imax; j++) { double v = fabs(A_(imax, j)); if (v > rowmax) rowmax = v; }
      for (int i = imax + 1; i < n; i++) { double v = fabs(A_(i, imax)); if (v > rowmax) rowmax = v; }
      if (absakk >= alpha * colmax * (colmax / rowmax)) kp = k;
      else if (fabs(A_(imax, imax)) >= alpha * rowmax) kp = imax;
      /* A 2 x 2 block needs k + 1 < n.  At the last column (imax == k) only
       * the NaN comparisons that follow a zero pivot reach this branch; the
       * reference then stores ipiv[n], one past its end (panel.mojo:2809-2812),
       * and here the step stays 1 x 1 instead. */
      else if (k + 1 < n) { kp = imax; kstep = 2; }
      else kp = k;
    }
    if (kstep == 1) ipiv[k] = kp + 1;
    else { ipiv[k] = -(kp + 1); ipiv[k + 1] = -(kp + 1); }
    int kk = k + kstep - 1;
    if (kp != kk) {
      for (int i = kp + 1; i < n; i++) { double h = A_(i, kk); A_(i, kk) = A_(i, kp); A_(i, kp) = h; }
      for (int j = kk + 1; j < kp; j++) { double h = A_(j, kk); A_(j, kk) = A_(kp, j); A_(kp, j) = h; }
      double h = A_(kk, kk); A_(kk, kk) = A_(kp, kp); A_(kp, kp) = h;
      if (kstep == 2) { double h2 = A_(k + 1, k); A_(k + 1, k) = A_(kp, k); A_(kp, k) = h2; }
    }
    if (kstep == 1) {
      if (k < n - 1) {
        double d11 = 1.0 / A_(k, k);
        for (int j = k + 1; j < n; j++) {
          double xj = A_(j, k);
          for (int i = j; i < n; i++) A_(i, j) = A_(i, j) - d11 * A_(i, k) * xj;
        }
        for (int i = k + 1; i < n; i++) A_(i, k) = A_(i, k) * d11;
      }
    } else if (k < n - 2) {
      double d21 = A_(k + 1, k);
      double d11 = A_(k + 1, k + 1) / d21;
      double d22 = A_(k, k) / d21;
      double tt = 1.0 / (d11 * d22 - 1.0);
      d21 = tt / d21;
      for (int j = k + 2; j < n; j++) {
        double wk = d21 * (d11 * A_(j, k) - A_(j, k + 1));
        double wkp1 = d21 * (d22 * A_(j, k + 1) - A_(j, k));
        for (int i = j; i < n; i++) A_(i, j) = A_(i, j) - A_(i, k) * wk - A_(i, k + 1) * wkp1;
      }
      for (int j = k + 2; j < n; j++) {
        double wk = d21 * (d11 * A_(j, k) - A_(j, k + 1));
        double wkp1 = d21 * (d22 * A_(j, k + 1) - A_(j, k));
        A_(j, k) = wk;
        A_(j, k + 1) = wkp1;
      }
    }
    k += kstep;
  }
#undef A_
  return info;
}

/* ldl_unpack_lower (numax/linalg/panel.mojo:2897-3003), SciPy's
 * _ldl_sanitize_ipiv / _ldl_get_d_and_l / _ldl_construct_tri_factor:
 * (lu, d, perm) from sytf2's packed factor. */
void orc_ldl_unpack(int n, const double* ldu, const int32_t* ipiv, double* lu, double* d, int64_t* perm) {
  int32_t* swap = (int32_t*)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
  int32_t* piv = (int32_t*)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
  int32_t* pr = (int32_t*)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
  for (int i = 0; i < n; i++) { swap[i] = i; piv[i] = 0; pr[i] = i; }
  int ind = 0;
  while (ind < n) {
    int cur = ipiv[ind];
    if (cur > 0) { if (cur != ind + 1) swap[ind] = swap[cur - 1]; piv[ind] = 1; ind += 1; }
    else { if (-cur != ind + 2) swap[ind + 1] = swap[-cur - 1]; piv[ind] = 2; ind += 2; }
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      d[i * n + j] = i == j ? ldu[i * n + i] : 0.0;
      lu[i * n + j] = i == j ? 1.0 : (i > j ? ldu[i * n + j] : 0.0);
    }
  for (int blk = 0; blk < n;) {
    if (piv[blk] == 2) {
      double v = ldu[(blk + 1) * n + blk];
      d[(blk + 1) * n + blk] = v; d[blk * n + blk + 1] = v; lu[(blk + 1) * n + blk] = 0.0;
      blk += 2;
    } else {
      blk += 1;
    }
  }
  for (ind = n - 1; ind >= 0; ind--) {
    int s = swap[ind];
    if (s != ind) {
      int col_s = piv[ind] == 0 ? ind - 1 : ind;
      for (int c = col_s; c < n; c++) { double h = lu[s * n + c]; lu[s * n + c] = lu[ind * n + c]; lu[ind * n + c] = h; }
      int32_t ps = pr[s]; pr[s] = pr[ind]; pr[ind] = ps;
    }
  }
  for (int i = 0; i < n; i++) perm[pr[i]] = i;
  free(swap); free(piv); free(pr);
}

/* _jacobi_rotation (_array/eigen.mojo:34-75). */
static void jacobi_rotation(double num, double den, double* c, double* s) {
  double twice = 2.0 * num;
  double radius = sqrt(den * den + twice * twice);
  double safe = guard_nonzero(fabs(den) + radius, PIVOT_FLOOR);
  double tangent = copysign(1.0, den) * twice / safe;
  *c = 1.0 / sqrt(1.0 + tangent * tangent);
  *s = tangent * *c;
}

/* eigh (_array/eigen.mojo:316-392): cyclic-by-row Jacobi, a fixed number of
 * sweeps, pairs p < q in order; rows, then columns, then the vectors.
 * Eigenvalues are the diagonal, unsorted; vectors are the columns of v. */
void orc_array_eigh(int n, const double* a, double* w, double* v, int sweeps) {
  double* work = (double*)malloc(sizeof(double) * n * n);
  memcpy(work, a, sizeof(double) * n * n);
  for (int i = 0; i < n * n; i++) v[i] = 0.0;
  for (int i = 0; i < n; i++) v[i * n + i] = 1.0;
  for (int sw = 0; sw < sweeps; sw++)
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        double c, s;
        jacobi_rotation(work[p * n + q], work[q * n + q] - work[p * n + p], &c, &s);
        for (int k = 0; k < n; k++) {
          double akp = work[p * n + k], akq = work[q * n + k];
          work[p * n + k] = c * akp - s * akq;
          work[q * n + k] = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {
          double apk = work[k * n + p], aqk = work[k * n + q];
          work[k * n + p] = c * apk - s * aqk;
          work[k * n + q] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; k++) {
          double vkp = v[k * n + p], vkq = v[k * n + q];
          v[k * n + p] = c * vkp - s * vkq;
          v[k * n + q] = s * vkp + c * vkq;
        }
      }
  for (int i = 0; i < n; i++) w[i] = work[i * n + i];
  free(work);
}

/* lstsq (_array/qr.mojo:133-204): Householder with alpha =
 * copysign(||x||, -r_kk), H = I - (2/v^T v) v v^T applied to R and to b,
 * then guarded back substitution. */
void orc_array_lstsq(int m, int n, const double* a, const double* b, double* x) {
  double* r = (double*)malloc(sizeof(double) * (size_t)m * n);
  double* y = (double*)malloc(sizeof(double) * m);
  double* v = (double*)malloc(sizeof(double) * m);
  memcpy(r, a, sizeof(double) * (size_t)m * n);
  memcpy(y, b, sizeof(double) * m);
  for (int k = 0; k < n; k++) {
    double norm_sq = 0.0;
    for (int i = k; i < m; i++) norm_sq = norm_sq + r[i * n + k] * r[i * n + k];
    double alpha = copysign(sqrt(norm_sq), -r[k * n + k]);
    for (int i = 0; i < m; i++) v[i] = 0.0;
    for (int i = k; i < m; i++) v[i] = r[i * n + k];
    v[k] = v[k] - alpha;
    double vv = 0.0;
    for (int i = k; i < m; i++) vv = vv + v[i] * v[i];
    double scale = 2.0 / guard_nonzero(vv, PIVOT_FLOOR);
    for (int j = k; j < n; j++) {
      double vr = 0.0;
      for (int i = k; i < m; i++) vr = vr + v[i] * r[i * n + j];
      double f = vr * scale;
      for (int i = k; i < m; i++) r[i * n + j] = r[i * n + j] - f * v[i];
    }
    double vy = 0.0;
    for (int i = k; i < m; i++) vy = vy + v[i] * y[i];
    double fy = vy * scale;
    for (int i = k; i < m; i++) y[i] = y[i] - fy * v[i];
  }
  for (int s = 0; s < n; s++) {
    int i = n - 1 - s;
    double total = y[i];
    for (int j = i + 1; j < n; j++) total = total - r[i * n + j] * x[j];
    x[i] = total / guard_nonzero(r[i * n + i], PIVOT_FLOOR);
  }
  free(r); free(y); free(v);
}

/* ------------------------------------------------------------------------ */
/* Batches: map_blocks' CPU form (core/functional.mojo:691-834) runs one    */
/* problem per lane over all cores; here one problem per loop index over    */
/* the threads of parallel_for.  Each problem is exactly the single-problem */
/* function above.                                                          */
/* ------------------------------------------------------------------------ */
typedef struct {
  const int32_t* n; const int64_t* offa; const int64_t* offb; const uint8_t* kind;
  const double* a; const double* b; double* x;
} gesv_batch_ctx;

static void gesv_batch_range(void* p, int64_t lo, int64_t hi) {
  gesv_batch_ctx* g = (gesv_batch_ctx*)p;
  for (int64_t q = lo; q < hi; q++) {
    const int n = g->n[q];
    double* f = (double*)malloc(sizeof(double) * (size_t)n * n);
    if (g->kind[q] == 0) {
      int32_t* perm = (int32_t*)malloc(sizeof(int32_t) * n);
      memcpy(f, g->a + g->offa[q], sizeof(double) * (size_t)n * n);
      orc_array_lu_factor(n, f, perm);
      orc_array_lu_solve(n, f, perm, g->b + g->offb[q], g->x + g->offb[q]);
      free(perm);
    } else {
      orc_array_cholesky(n, g->a + g->offa[q], f);
      orc_array_cholesky_solve(n, f, g->b + g->offb[q], g->x + g->offb[q]);
    }
    free(f);
  }
}

/* count uneven systems, kind 0 = lu_factor + PivotedLU.solve, 1 = cholesky +
 * cholesky_solve, packed back to back at element offsets offa / offb. */
void orc_array_gesv_batch(int count, const int32_t* n, const int64_t* offa, const int64_t* offb,
                          const uint8_t* kind, const double* a, const double* b, double* x) {
  gesv_batch_ctx g = {n, offa, offb, kind, a, b, x};
  parallel_for(count, (int64_t)count << 20, gesv_batch_range, &g);
}

typedef struct {
  const int32_t* n; const int64_t* offa; const int64_t* offw;
  const double* a; double* w; double* v; int sweeps;
} eigh_batch_ctx;

static void eigh_batch_range(void* p, int64_t lo, int64_t hi) {
  eigh_batch_ctx* g = (eigh_batch_ctx*)p;
  for (int64_t q = lo; q < hi; q++)
    orc_array_eigh(g->n[q], g->a + g->offa[q], g->w + g->offw[q], g->v + g->offa[q], g->sweeps);
}

void orc_array_eigh_batch(int count, const int32_t* n, const int64_t* offa, const int64_t* offw, const double* a,
                          double* w, double* v, int sweeps) {
  eigh_batch_ctx g = {n, offa, offw, a, w, v, sweeps};
  parallel_for(count, (int64_t)count << 20, eigh_batch_range, &g);
}

typedef struct { int m, n; const double* a; const double* b; double* x; } lstsq_batch_ctx;

static void lstsq_batch_range(void* p, int64_t lo, int64_t hi) {
  lstsq_batch_ctx* g = (lstsq_batch_ctx*)p;
  for (int64_t q = lo; q < hi; q++)
    orc_array_lstsq(g->m, g->n, g->a + q * (int64_t)g->m * g->n, g->b + q * (int64_t)g->m, g->x + q * (int64_t)g->n);
}

/* count problems m x n packed back to back (A), b of m, x of n each. */
void orc_array_lstsq_batch(int count, int m, int n, const double* a, const double* b, double* x) {
  lstsq_batch_ctx g = {m, n, a, b, x};
  parallel_for(count, (int64_t)count << 20, lstsq_batch_range, &g);
}

/* ------------------------------------------------------------------------ */
/* Checks                                                                   */
/* ------------------------------------------------------------------------ */

/* Normwise backward error max_c ||b_c - A x_c||_inf / (||A||_inf ||x_c||_inf),
 * residual accumulated in long double. */
double orc_backward_error(int n, int nrhs, const double* a, int lda, const double* x, int ldx, const double* b, int ldb) {
  double anorm = 0.0;
  for (int i = 0; i < n; i++) {
    double s = 0.0;
    for (int j = 0; j < n; j++) s += fabs(AT(a, lda, i, j));
    if (s > anorm) anorm = s;
  }
  double worst = 0.0;
  for (int c = 0; c < nrhs; c++) {
    double rmax = 0.0, xmax = 0.0;
    for (int i = 0; i < n; i++) {
      long double r = AT(b, ldb, i, c);
      for (int j = 0; j < n; j++) r -= (long double)AT(a, lda, i, j) * AT(x, ldx, j, c);
      double ar = fabs((double)r);
      if (ar > rmax) rmax = ar;
      double ax = fabs(AT(x, ldx, i, c));
      if (ax > xmax) xmax = ax;
    }
    double eta = rmax / (anorm * xmax);
    if (!(eta <= worst)) worst = eta;
  }
  return worst;
}
