// Register-tier dense operations over batches of small matrices: numax's
// Array-tier lu / solve / inverse / det (unpivoted), PivotedLU's det and
// slogdet, slogdet_cholesky, and the three solves a map_blocks step runs
// per lane (pivoted LU, Cholesky, unpivoted LU), in numax's own loop orders
// and without fused multiply-adds, so that FP64 results are the CPU
// restatement's bit for bit (log excepted).
//
// Two layouts:
//  * row-major problems back to back at element offsets (one wavefront per
//    problem, the matrix in LDS, n <= 64 -- numax's register tier stops at
//    64, numax/linalg/common.mojo:30);
//  * map_blocks' structure of arrays (numax/core/functional.mojo:729-735):
//    a (k_in, batch) array whose row j holds element j of every problem,
//    A row-major in rows 0 .. n^2 - 1 and b in rows n^2 .. n^2 + n - 1,
//    answers (n, batch).  Up to n = 8 one thread owns one problem in
//    registers, as map_blocks' lane does, and neighbouring threads read
//    neighbouring addresses; larger n use the wavefront kernel with the
//    batch as element stride.
#include <math.h>

#include "nmx_internal.h"

// every a - b * c below is a multiply, then a subtraction: numax's loops
// round twice
#pragma clang fp contract(off)

namespace {

constexpr int kSmallMax = 64;  // register tier's largest order
constexpr int kLd = kSmallMax + 1;
constexpr int kRegMax = 8;     // map_blocks orders solved in one thread's registers

enum SmallOp {
  OP_LU = 0,           // unpivoted Doolittle, packed (_array/lu.mojo:25-55)
  OP_SOLVE,            // unpivoted solve (_array/basic.mojo:20-31)
  OP_INVERSE,          // unpivoted inverse (_array/basic.mojo:34-60)
  OP_DET,              // unpivoted det (_array/lu.mojo:219-233)
  OP_DET_PIV,          // PivotedLU.det (_array/lu.mojo:112-123)
  OP_SLOGDET,          // slogdet through lu_factor (_array/lu.mojo:125-162, :236-258)
  OP_SLOGDET_CHOL,     // slogdet_cholesky of a factor (_array/cholesky.mojo:95-116)
  OP_SOLVE_PIV,        // lu_factor + PivotedLU.solve (_array/lu.mojo:92-110, :165-216)
  OP_SOLVE_CHOL,       // cholesky + cholesky_solve (_array/cholesky.mojo:26-92)
};

__host__ __device__ constexpr bool op_is_solve(int op) {
  return op == OP_SOLVE || op == OP_SOLVE_PIV || op == OP_SOLVE_CHOL;
}
__host__ __device__ constexpr bool op_is_matrix(int op) { return op == OP_LU || op == OP_INVERSE; }

template <typename T>
struct SmallArgs {
  int op, count;
  const int32_t* n;     // orders (row-major layout); nullptr: every problem has order n0 (map_blocks)
  int n0;
  const int64_t* offA;  // row-major layout: element offsets of each matrix / vector
  const int64_t* offB;
  int64_t batch;        // map_blocks layout: problems (the element stride)
  int64_t a_elems, b_elems;
  const T* A;           // matrices (map_blocks: the whole (k_in, batch) array)
  const T* B;           // right-hand sides (map_blocks: same as A)
  T* X;                 // matrix or vector results
  T* s1;                // scalar results
  T* s2;
};

template <typename T>
__device__ __forceinline__ T guard(T x) {  // guard_nonzero(x, 1e-30), numax/core/numeric.mojo:346-364
  const T f = (T)1e-30;
  const T m = fabs(x);
  return (m >= f || m != m) ? x : copysign(f, x);
}

// Where problem p lives: its order, and the first element and element
// stride of its matrix, its right-hand side and its result.  False: skip it.
template <typename T>
__device__ __forceinline__ bool locate(const SmallArgs<T>& a, int p, int& n, int64_t& ao, int64_t& bo, int64_t& xo,
                                       int64_t& es) {
  if (!a.n) {
    n = a.n0;
    es = a.batch;
    ao = p;
    bo = (int64_t)n * n * a.batch + p;
    xo = p;
    return true;
  }
  n = a.n[p];
  es = 1;
  if (n < 1 || n > kSmallMax) return false;
  ao = a.offA[p];
  if (ao < 0 || ao + (int64_t)n * n > a.a_elems) return false;
  bo = xo = 0;
  if (op_is_solve(a.op)) {
    bo = xo = a.offB[p];
    if (bo < 0 || bo + n > a.b_elems) return false;
  } else if (op_is_matrix(a.op)) {
    xo = ao;
  }
  return true;
}

template <typename T>
__device__ __forceinline__ T wave_max_t(T v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m));
  return v;
}

// ---- one wavefront per problem: lane i owns row i of M (LDS) ------------

// Unpivoted Doolittle in place (lu, _array/lu.mojo:25-55).
template <typename T>
__device__ void lds_lu(T* M, int n, int lane) {
  for (int k = 0; k < n; k++) {
    const T piv = guard(M[k * kLd + k]);
    if (lane > k && lane < n) {
      const T f = M[lane * kLd + k] / piv;
      M[lane * kLd + k] = f;
      for (int j = k + 1; j < n; j++) M[lane * kLd + j] = M[lane * kLd + j] - f * M[k * kLd + j];
    }
    __syncthreads();
  }
}

// lu_factor (_array/lu.mojo:165-216): the first largest |a_ik| from the
// diagonal down (a NaN on the diagonal stays; a NaN below never wins), full
// row swaps, then the unpivoted step.  The swaps are also applied to the
// vector v (nullable), which is PivotedLU.solve's b[permutation].  Returns
// the permutation's sign.
template <typename T>
__device__ int lds_lu_piv(T* M, T* v, int n, int lane) {
  int sign = 1;
  for (int k = 0; k < n; k++) {
    const T dk = fabs(M[k * kLd + k]);
    int best = k;
    if (dk == dk) {
      T mag = (lane >= k && lane < n) ? fabs(M[lane * kLd + k]) : (T)-1;
      if (mag != mag) mag = (T)-1;
      const T mx = wave_max_t(mag);
      const uint64_t hit = __ballot(lane >= k && lane < n && mag == mx);
      best = __ffsll((long long)hit) - 1;
    }
    if (best != k) {
      if (lane < n) {
        const T s = M[k * kLd + lane];
        M[k * kLd + lane] = M[best * kLd + lane];
        M[best * kLd + lane] = s;
      }
      if (v && lane == 0) {
        const T s = v[k];
        v[k] = v[best];
        v[best] = s;
      }
      sign = -sign;
      __syncthreads();
    }
    const T piv = guard(M[k * kLd + k]);
    if (lane > k && lane < n) {
      const T f = M[lane * kLd + k] / piv;
      M[lane * kLd + k] = f;
      for (int j = k + 1; j < n; j++) M[lane * kLd + j] = M[lane * kLd + j] - f * M[k * kLd + j];
    }
    __syncthreads();
  }
  return sign;
}

// cholesky (_array/cholesky.mojo:26-58), left-looking, in place in the lower
// triangle (the upper is left as read; nothing below reads it).
template <typename T>
__device__ void lds_cholesky(T* M, int n, int lane) {
  for (int j = 0; j < n; j++) {
    T acc = (lane >= j && lane < n) ? M[lane * kLd + j] : (T)0;
    if (lane >= j && lane < n)
      for (int k = 0; k < j; k++) acc = acc - M[lane * kLd + k] * M[j * kLd + k];
    if (lane == j) M[j * kLd + j] = sqrt(guard(acc));
    __syncthreads();
    const T ljj = M[j * kLd + j];
    if (lane > j && lane < n) M[lane * kLd + j] = acc / ljj;
    __syncthreads();
  }
}

// forward_substitution (_array/triangular.mojo:24-43) of the vector in v,
// in place: row i subtracts l_ij x_j for j = 0 .. i-1 in that order, which
// the column sweep keeps (lane i accumulates row i's total as x_j appear).
template <typename T>
__device__ void lds_forward(const T* M, T* v, int n, int lane, bool unit) {
  T total = lane < n ? v[lane] : (T)0;
  for (int j = 0; j < n; j++) {
    if (lane == j) v[j] = unit ? total : total / guard(M[j * kLd + j]);
    __syncthreads();
    const T xj = v[j];
    if (lane > j && lane < n) total = total - M[lane * kLd + j] * xj;
  }
  __syncthreads();
}

// back_substitution (_array/triangular.mojo:46-63), in place: row i
// subtracts u_ij x_j for j = i+1 .. n-1 in that order, so every row waits
// for the one below it -- one lane walks it.  TRANS reads U = M^T (the
// Cholesky solve's L^T).
template <typename T, bool TRANS>
__device__ void lds_back(const T* M, T* v, int n, int lane) {
  if (lane == 0) {
    for (int s = 0; s < n; s++) {
      const int i = n - 1 - s;
      T total = v[i];
      for (int j = i + 1; j < n; j++) total = total - (TRANS ? M[j * kLd + i] : M[i * kLd + j]) * v[j];
      v[i] = total / guard(M[i * kLd + i]);
    }
  }
  __syncthreads();
}

template <typename T, bool INV>
__global__ void __launch_bounds__(64) small_wave_kernel(SmallArgs<T> a) {
  __shared__ T M[kSmallMax * kLd];
  __shared__ T V[INV ? kSmallMax * kLd : kSmallMax];
  const int lane = threadIdx.x;
  for (int p = blockIdx.x; p < a.count; p += gridDim.x) {
    int n;
    int64_t ao, bo, xo, es;
    if (!locate(a, p, n, ao, bo, xo, es)) continue;
    for (int e = lane; e < n * n; e += 64) M[(e / n) * kLd + e % n] = a.A[ao + e * es];
    if (op_is_solve(a.op) && lane < n) V[lane] = a.B[bo + lane * es];
    __syncthreads();
    switch (a.op) {
      case OP_LU:
        lds_lu(M, n, lane);
        for (int e = lane; e < n * n; e += 64) a.X[xo + e * es] = M[(e / n) * kLd + e % n];
        break;
      case OP_SOLVE:
        lds_lu(M, n, lane);
        lds_forward(M, V, n, lane, true);
        lds_back<T, false>(M, V, n, lane);
        if (lane < n) a.X[xo + lane * es] = V[lane];
        break;
      case OP_SOLVE_PIV:
        lds_lu_piv(M, V, n, lane);
        lds_forward(M, V, n, lane, true);
        lds_back<T, false>(M, V, n, lane);
        if (lane < n) a.X[xo + lane * es] = V[lane];
        break;
      case OP_SOLVE_CHOL:
        lds_cholesky(M, n, lane);
        lds_forward(M, V, n, lane, false);
        lds_back<T, true>(M, V, n, lane);
        if (lane < n) a.X[xo + lane * es] = V[lane];
        break;
      case OP_INVERSE:
        if constexpr (INV) {
          // inverse (_array/basic.mojo:34-60): lane c substitutes column c
          // of the identity, forward then back, in numax's row order
          lds_lu(M, n, lane);
          if (lane < n) {
            T* x = V + lane * kLd;
            for (int i = 0; i < n; i++) {
              T total = i == lane ? (T)1 : (T)0;
              for (int j = 0; j < i; j++) total = total - M[i * kLd + j] * x[j];
              x[i] = total;
            }
            for (int s = 0; s < n; s++) {
              const int i = n - 1 - s;
              T total = x[i];
              for (int j = i + 1; j < n; j++) total = total - M[i * kLd + j] * x[j];
              x[i] = total / guard(M[i * kLd + i]);
            }
          }
          __syncthreads();
          for (int e = lane; e < n * n; e += 64) a.X[xo + e * es] = V[(e % n) * kLd + e / n];
        }
        break;
      case OP_DET:
        lds_lu(M, n, lane);
        if (lane == 0) {
          T prod = (T)1;
          for (int i = 0; i < n; i++) prod = prod * M[i * kLd + i];
          a.s1[p] = prod;
        }
        break;
      case OP_DET_PIV:
      case OP_SLOGDET: {
        const int sign = lds_lu_piv<T>(M, nullptr, n, lane);
        if (lane == 0) {
          if (a.op == OP_DET_PIV) {
            T prod = (T)sign;
            for (int i = 0; i < n; i++) prod = prod * M[i * kLd + i];
            a.s1[p] = prod;
          } else {
            double sg = (double)sign, total = 0.0;
            for (int i = 0; i < n; i++) {
              const double d = (double)M[i * kLd + i];
              if (d == 0.0) {
                sg = 0.0;
                total = -INFINITY;
                break;
              }
              if (d < 0.0) sg = -sg;
              total += log(fabs(d));
            }
            a.s1[p] = (T)sg;
            a.s2[p] = (T)total;
          }
        }
        break;
      }
      case OP_SLOGDET_CHOL:
        if (lane == 0) {
          T total = (T)0;
          for (int i = 0; i < n; i++) total = total + log(M[i * kLd + i]);
          a.s1[p] = total * (T)2;
        }
        break;
    }
    __syncthreads();  // M and V are free for the next problem
  }
}

// ---- map_blocks, n <= 8: one thread per problem, in registers ------------

template <typename T, int N>
__device__ __forceinline__ void reg_forward(const T (&m)[N][N], T (&v)[N], bool unit) {
#pragma unroll
  for (int i = 0; i < N; i++) {
    T total = v[i];
#pragma unroll
    for (int j = 0; j < i; j++) total = total - m[i][j] * v[j];
    v[i] = unit ? total : total / guard(m[i][i]);
  }
}

template <typename T, int N, bool TRANS>
__device__ __forceinline__ void reg_back(const T (&m)[N][N], T (&v)[N]) {
#pragma unroll
  for (int s = 0; s < N; s++) {
    const int i = N - 1 - s;
    T total = v[i];
#pragma unroll
    for (int j = i + 1; j < N; j++) total = total - (TRANS ? m[j][i] : m[i][j]) * v[j];
    v[i] = total / guard(m[i][i]);
  }
}

template <typename T, int N>
__device__ __forceinline__ void reg_eliminate(T (&m)[N][N], int k) {
  const T piv = guard(m[k][k]);
#pragma unroll
  for (int i = k + 1; i < N; i++) {
    const T f = m[i][k] / piv;
    m[i][k] = f;
#pragma unroll
    for (int j = k + 1; j < N; j++) m[i][j] = m[i][j] - f * m[k][j];
  }
}

// KIND: 0 lu_factor + PivotedLU.solve, 1 cholesky + cholesky_solve, 2
// unpivoted solve -- the step bodies of numax's map_blocks batched solve
// (examples/advanced/batched_solve.mojo:87-105).  Every index is static:
// the pivot row is swapped in by selects over the rows below.
template <typename T, int N, int KIND>
__global__ void __launch_bounds__(256) map_blocks_reg_kernel(int64_t batch, const T* __restrict__ xs,
                                                             T* __restrict__ ys) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= batch) return;
  T m[N][N], v[N];
#pragma unroll
  for (int i = 0; i < N; i++) {
#pragma unroll
    for (int j = 0; j < N; j++) m[i][j] = xs[(int64_t)(i * N + j) * batch + p];
    v[i] = xs[(int64_t)(N * N + i) * batch + p];
  }
  if constexpr (KIND == 0) {
#pragma unroll
    for (int k = 0; k < N; k++) {
      int best = k;
      T bm = fabs(m[k][k]);
#pragma unroll
      for (int i = k + 1; i < N; i++) {
        const T mag = fabs(m[i][k]);
        const bool c = mag > bm;
        best = c ? i : best;
        bm = c ? mag : bm;
      }
#pragma unroll
      for (int i = k + 1; i < N; i++) {
        const bool s = best == i;
#pragma unroll
        for (int j = 0; j < N; j++) {
          const T t = m[k][j];
          m[k][j] = s ? m[i][j] : t;
          m[i][j] = s ? t : m[i][j];
        }
        const T t = v[k];
        v[k] = s ? v[i] : t;
        v[i] = s ? t : v[i];
      }
      reg_eliminate<T, N>(m, k);
    }
    reg_forward<T, N>(m, v, true);
    reg_back<T, N, false>(m, v);
  } else if constexpr (KIND == 1) {
#pragma unroll
    for (int j = 0; j < N; j++) {
      T diag = m[j][j];
#pragma unroll
      for (int k = 0; k < j; k++) diag = diag - m[j][k] * m[j][k];
      const T ljj = sqrt(guard(diag));
#pragma unroll
      for (int i = j + 1; i < N; i++) {
        T total = m[i][j];
#pragma unroll
        for (int k = 0; k < j; k++) total = total - m[i][k] * m[j][k];
        m[i][j] = total / ljj;
      }
      m[j][j] = ljj;
    }
    reg_forward<T, N>(m, v, false);
    reg_back<T, N, true>(m, v);
  } else {
#pragma unroll
    for (int k = 0; k < N; k++) reg_eliminate<T, N>(m, k);
    reg_forward<T, N>(m, v, true);
    reg_back<T, N, false>(m, v);
  }
#pragma unroll
  for (int i = 0; i < N; i++) ys[(int64_t)i * batch + p] = v[i];
}

template <typename T, int N>
int map_blocks_reg_launch(nmx_ctx* ctx, int kind, int64_t batch, const T* xs, T* ys) {
  const unsigned grid = (unsigned)((batch + 255) / 256);
  if (kind == 0) map_blocks_reg_kernel<T, N, 0><<<grid, 256, 0, ctx->stream>>>(batch, xs, ys);
  if (kind == 1) map_blocks_reg_kernel<T, N, 1><<<grid, 256, 0, ctx->stream>>>(batch, xs, ys);
  if (kind == 2) map_blocks_reg_kernel<T, N, 2><<<grid, 256, 0, ctx->stream>>>(batch, xs, ys);
  NMX_HIP(ctx, hipGetLastError());
  return NMX_OK;
}

template <typename T>
int small_launch(nmx_ctx* ctx, const SmallArgs<T>& a) {
  ProfScope ps(ctx, K_BATCH, 0.0, 0.0);
  const int grid = (int)std::min<int64_t>(a.count, (int64_t)ctx->num_cus * 8);
  if (a.op == OP_INVERSE)
    small_wave_kernel<T, true><<<grid, 64, 0, ctx->stream>>>(a);
  else
    small_wave_kernel<T, false><<<grid, 64, 0, ctx->stream>>>(a);
  NMX_HIP(ctx, hipGetLastError());
  return NMX_OK;
}

// The row-major batch entry points share their argument checks.
int small_rowmajor(nmx_ctx* ctx, int op, int count, const int32_t* n, const int64_t* offA, const double* A,
                   int64_t a_elems, const int64_t* offB, const double* B, int64_t b_elems, double* X, double* s1,
                   double* s2) {
  if (!ctx) return -1;
  if (count < 0) return -2;
  if (count == 0) return NMX_OK;
  if (!n) return -3;
  if (!offA) return -4;
  if (!A || a_elems <= 0) return -5;
  if (op_is_solve(op) && (!offB || !B || b_elems <= 0)) return -6;
  if ((op_is_solve(op) || op_is_matrix(op)) && !X) return -7;
  if (!op_is_solve(op) && !op_is_matrix(op) && !s1) return -8;
  if (op == OP_SLOGDET && !s2) return -9;
  SmallArgs<double> a{op, count, n, 0, offA, offB, 0, a_elems, b_elems, A, B, X, s1, s2};
  return small_launch<double>(ctx, a);
}

template <typename T>
int map_blocks_api(nmx_ctx* ctx, int kind, int n, int64_t batch, const T* xs, T* ys) {
  if (!ctx) return -1;
  if (kind < 0 || kind > 2) return -2;
  if (n < 1 || n > kSmallMax) return -3;
  if (batch < 0) return -4;
  if (batch == 0) return NMX_OK;
  if (!xs) return -5;
  if (!ys) return -6;
  if (n <= kRegMax) {
    ProfScope ps(ctx, K_BATCH, 0.0, 0.0);
    switch (n) {
      case 1: return map_blocks_reg_launch<T, 1>(ctx, kind, batch, xs, ys);
      case 2: return map_blocks_reg_launch<T, 2>(ctx, kind, batch, xs, ys);
      case 3: return map_blocks_reg_launch<T, 3>(ctx, kind, batch, xs, ys);
      case 4: return map_blocks_reg_launch<T, 4>(ctx, kind, batch, xs, ys);
      case 5: return map_blocks_reg_launch<T, 5>(ctx, kind, batch, xs, ys);
      case 6: return map_blocks_reg_launch<T, 6>(ctx, kind, batch, xs, ys);
      case 7: return map_blocks_reg_launch<T, 7>(ctx, kind, batch, xs, ys);
      default: return map_blocks_reg_launch<T, 8>(ctx, kind, batch, xs, ys);
    }
  }
  if (batch > INT32_MAX) return -4;
  const int op = kind == 0 ? OP_SOLVE_PIV : kind == 1 ? OP_SOLVE_CHOL : OP_SOLVE;
  SmallArgs<T> a{op, (int)batch, nullptr, n, nullptr, nullptr, batch, 0, 0, xs, xs, ys, nullptr, nullptr};
  return small_launch<T>(ctx, a);
}

}  // namespace

NMX_API int nmx_batched_lu_nopiv_f64(nmx_ctx* ctx, int count, const int32_t* n, const int64_t* offA,
                                     const double* A, int64_t a_elems, double* LU) {
  return small_rowmajor(ctx, OP_LU, count, n, offA, A, a_elems, nullptr, nullptr, 0, LU, nullptr, nullptr);
}

NMX_API int nmx_batched_solve_nopiv_f64(nmx_ctx* ctx, int count, const int32_t* n, const int64_t* offA,
                                        const int64_t* offB, const double* A, int64_t a_elems, const double* B,
                                        int64_t b_elems, double* X) {
  return small_rowmajor(ctx, OP_SOLVE, count, n, offA, A, a_elems, offB, B, b_elems, X, nullptr, nullptr);
}

NMX_API int nmx_batched_inverse_nopiv_f64(nmx_ctx* ctx, int count, const int32_t* n, const int64_t* offA,
                                          const double* A, int64_t a_elems, double* Ainv) {
  return small_rowmajor(ctx, OP_INVERSE, count, n, offA, A, a_elems, nullptr, nullptr, 0, Ainv, nullptr, nullptr);
}

NMX_API int nmx_batched_det_f64(nmx_ctx* ctx, int count, const int32_t* n, const int64_t* offA, const double* A,
                                int64_t a_elems, int pivoted, double* det) {
  if (pivoted != 0 && pivoted != 1) return -7;
  return small_rowmajor(ctx, pivoted ? OP_DET_PIV : OP_DET, count, n, offA, A, a_elems, nullptr, nullptr, 0,
                        nullptr, det, nullptr);
}

NMX_API int nmx_batched_slogdet_f64(nmx_ctx* ctx, int count, const int32_t* n, const int64_t* offA,
                                    const double* A, int64_t a_elems, double* sign, double* logabsdet) {
  return small_rowmajor(ctx, OP_SLOGDET, count, n, offA, A, a_elems, nullptr, nullptr, 0, nullptr, sign,
                        logabsdet);
}

NMX_API int nmx_batched_slogdet_cholesky_f64(nmx_ctx* ctx, int count, const int32_t* n, const int64_t* offL,
                                             const double* L, int64_t l_elems, double* logdet) {
  return small_rowmajor(ctx, OP_SLOGDET_CHOL, count, n, offL, L, l_elems, nullptr, nullptr, 0, nullptr, logdet,
                        nullptr);
}

NMX_API int nmx_map_blocks_solve_f64(nmx_ctx* ctx, int kind, int n, int64_t batch, const double* xs, double* ys) {
  return map_blocks_api<double>(ctx, kind, n, batch, xs, ys);
}

NMX_API int nmx_map_blocks_solve_f32(nmx_ctx* ctx, int kind, int n, int64_t batch, const float* xs, float* ys) {
  return map_blocks_api<float>(ctx, kind, n, batch, xs, ys);
}
