// Blocked Householder QR, FP64 and FP32, device-resident: numax's Tensor-tier
// qr_factor / QR.apply_q_transpose / QR.q / QR.r / QR.solve / lstsq
// (numax/linalg/qr.mojo:330-821).
//
// Right-looking and blocked as numax's: a `block`-wide panel of reflectors
// is factored in place (geqr2_panel, numax/linalg/panel.mojo:1975-2073,
// LAPACK dlarfg's convention), then its block reflector reaches everything
// to its right at once as C -= V (T^T (V^T C)) (qr.mojo:123-327) with T
// the forward columnwise triangular factor (larft_panel, panel.mojo:
// 2076-2169).  The three products are the library's MFMA GEMM; the panel
// is one workgroup of 16 wavefronts streaming the panel's rows, one column
// at a time: the column norm and the panel's dot products are workgroup
// reductions (shuffles within a wavefront, LDS across them).  T is built
// from the Gram matrix V^T V, itself one GEMM of an explicit V (unit
// diagonal, zeros above), which the three block-reflector products use too.
#include <math.h>

#include "nmx_internal.h"

namespace {

constexpr int kQrMaxBlock = 64;  // panel width the kernels take
constexpr int kQrThreads = 1024;
constexpr int kQrWaves = kQrThreads / 64;
constexpr int kQrChunk = 16;     // panel columns whose dot products are reduced together

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m);
  return v;
}

// Sum over the workgroup of NV values per thread; every thread gets the totals.
template <typename T, int NV>
__device__ __forceinline__ void block_sum(T (&v)[NV], T (*red)[kQrChunk], T* out) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
#pragma unroll
  for (int u = 0; u < NV; u++) {
    const T s = wave_sum(v[u]);
    if (lane == 0) red[w][u] = s;
  }
  __syncthreads();
  if (t < NV) {
    T s = 0;
    for (int q = 0; q < kQrWaves; q++) s += red[q][t];
    out[t] = s;
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < NV; u++) v[u] = out[u];
  __syncthreads();
}

// geqr2_panel on A[k:m, k:k+nb] in place, tau[k..k+nb).  Per column: sigma =
// sum of squares below the diagonal; sigma == 0 leaves the column and gives
// tau = 0; otherwise beta = -sign(alpha) sqrt(alpha^2 + sigma), tau =
// (beta - alpha) / beta, v = x / (alpha - beta) below the diagonal, R_kk =
// beta, and each later panel column c gets c -= tau (c_k + v^T c_below) v.
template <typename T>
__global__ void __launch_bounds__(kQrThreads) geqr2_kernel(int m, T* __restrict__ A, int lda, T* __restrict__ tau,
                                                           int k, int nb) {
  __shared__ T red[kQrWaves][kQrChunk];
  __shared__ T tot[kQrChunk];
  const int t = threadIdx.x;
  for (int j = 0; j < nb; j++) {
    const int col = k + j;
    T s[1] = {T(0)};
    for (int i = col + 1 + t; i < m; i += kQrThreads) {
      const T a = A[(int64_t)i * lda + col];
      s[0] += a * a;
    }
    block_sum<T, 1>(s, red, tot);
    const T sigma = s[0];
    const T alpha = A[(int64_t)col * lda + col];
    if (sigma == T(0)) {
      if (t == 0) tau[col] = T(0);
      continue;
    }
    T beta = sqrt(alpha * alpha + sigma);
    if (alpha > T(0)) beta = -beta;
    const T tj = (beta - alpha) / beta;
    const T scale = T(1) / (alpha - beta);
    for (int i = col + 1 + t; i < m; i += kQrThreads) A[(int64_t)i * lda + col] *= scale;
    __syncthreads();  // the scaled column and R_kk before any thread reads them back
    if (t == 0) {
      A[(int64_t)col * lda + col] = beta;
      tau[col] = tj;
    }
    for (int c0 = j + 1; c0 < nb; c0 += kQrChunk) {
      T d[kQrChunk];
#pragma unroll
      for (int u = 0; u < kQrChunk; u++) d[u] = T(0);
      for (int i = col + 1 + t; i < m; i += kQrThreads) {
        const T* row = A + (int64_t)i * lda;
        const T v = row[col];
#pragma unroll
        for (int u = 0; u < kQrChunk; u++)
          if (c0 + u < nb) d[u] += v * row[k + c0 + u];
      }
      block_sum<T, kQrChunk>(d, red, tot);
      // dot_c = c_k + v^T c_below: the unit entry of v meets row col
      T f[kQrChunk];
#pragma unroll
      for (int u = 0; u < kQrChunk; u++)
        f[u] = c0 + u < nb ? tj * (A[(int64_t)col * lda + k + c0 + u] + d[u]) : T(0);
      __syncthreads();  // row col read by every thread before it is rewritten
      if (t < kQrChunk && c0 + t < nb) {
        T* p = A + (int64_t)col * lda + k + c0 + t;
        *p -= tj * (*p + tot[t]);
      }
      for (int i = col + 1 + t; i < m; i += kQrThreads) {
        T* row = A + (int64_t)i * lda;
        const T v = row[col];
#pragma unroll
        for (int u = 0; u < kQrChunk; u++)
          if (c0 + u < nb) row[k + c0 + u] -= f[u] * v;
      }
      __syncthreads();
    }
  }
}

// V (rows x nb, ld nb) explicit from the panel: unit diagonal, zeros above.
template <typename T>
__global__ void explicit_v_kernel(int rows, int nb, const T* __restrict__ A, int lda, int k, T* __restrict__ V) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)rows * nb) return;
  const int r = (int)(idx / nb), c = (int)(idx % nb);
  V[idx] = r == c ? T(1) : (r > c ? A[(int64_t)(k + r) * lda + k + c] : T(0));
}

// larft_panel's T (forward, columnwise, upper, nb x nb) from the Gram matrix
// G = V^T V and tau: T_ii = tau_i, T[0:i, i] = -tau_i T[0:i, 0:i] G[0:i, i].
template <typename E>
__global__ void __launch_bounds__(64) larft_kernel(int nb, const E* __restrict__ G, const E* __restrict__ tau, int k,
                                                   E* __restrict__ T) {
  __shared__ E Ts[kQrMaxBlock][kQrMaxBlock + 1];
  __shared__ E g[kQrMaxBlock];
  const int q = threadIdx.x;
  for (int r = 0; r < nb; r++)
    if (q < nb) Ts[r][q] = E(0);
  if (q == 0) Ts[0][0] = tau[k];
  __syncthreads();
  for (int i = 1; i < nb; i++) {
    const E ti = tau[k + i];
    if (q < i) g[q] = -ti * G[q * nb + i];
    __syncthreads();
    E s = E(0);
    if (q < i)
      for (int r = q; r < i; r++) s += Ts[q][r] * g[r];
    __syncthreads();
    if (q < i) Ts[q][i] = s;
    if (q == 0) Ts[i][i] = ti;
    __syncthreads();
  }
  for (int r = 0; r < nb; r++)
    if (q < nb) T[r * nb + q] = Ts[r][q];
}

template <typename T>
__global__ void identity_kernel(int m, int n, T* __restrict__ Q, int ldq) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)m * n) return;
  const int r = (int)(idx / n), c = (int)(idx % n);
  Q[(int64_t)r * ldq + c] = r == c ? T(1) : T(0);
}

template <typename T>
__global__ void upper_kernel(int n, const T* __restrict__ A, int lda, T* __restrict__ R, int ldr) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * n) return;
  const int r = (int)(idx / n), c = (int)(idx % n);
  R[(int64_t)r * ldr + c] = r <= c ? A[(int64_t)r * lda + c] : T(0);
}

// Workspace of one block reflector: V (rows x nb), G and T (nb x nb), W and Y
// (nb x width), carved from the context workspace.
template <typename T>
struct Reflector {
  T* V;
  T* G;
  T* T_;
  T* W;
  T* Y;
};

template <typename T>
size_t reflector_bytes(int m, int nb, int width) {
  const size_t e = (size_t)m * nb + 2 * (size_t)nb * nb + 2 * (size_t)nb * (width > 0 ? width : 1);
  return e * sizeof(T);
}

template <typename T>
Reflector<T> carve(void* ws, int m, int nb, int width) {
  Reflector<T> r;
  r.V = (T*)ws;
  r.G = r.V + (size_t)m * nb;
  r.T_ = r.G + (size_t)nb * nb;
  r.W = r.T_ + (size_t)nb * nb;
  r.Y = r.W + (size_t)nb * (width > 0 ? width : 1);
  return r;
}

// V and T of the panel at column k (nb wide) of the m x n factored matrix.
template <typename T>
int build_reflector(nmx_ctx* ctx, const Reflector<T>& w, int m, const T* QR, int lda, const T* tau, int k, int nb) {
  const int rows = m - k;
  explicit_v_kernel<T><<<(unsigned)(((int64_t)rows * nb + 255) / 256), 256, 0, ctx->stream>>>(rows, nb, QR, lda, k,
                                                                                             w.V);
  NMX_HIP(ctx, hipGetLastError());
  NMX_TRY(gemm_launch<T>(ctx, 1, 0, nb, nb, rows, T(1), w.V, nb, w.V, nb, T(0), w.G, nb));
  larft_kernel<T><<<1, 64, 0, ctx->stream>>>(nb, w.G, tau, k, w.T_);
  NMX_HIP(ctx, hipGetLastError());
  return NMX_OK;
}

// C (rows x width, ld ldc; rows = m - k) := (I - V op(T) V^T) C, op(T) = T^T
// when `trans` (Q^T, the factorization's and apply_q_transpose's direction),
// T otherwise (Q, QR.q's).
template <typename T>
int apply_reflector(nmx_ctx* ctx, const Reflector<T>& w, int rows, int nb, int width, int trans, T* C, int ldc) {
  if (width <= 0) return NMX_OK;
  NMX_TRY(gemm_launch<T>(ctx, 1, 0, nb, width, rows, T(1), w.V, nb, C, ldc, T(0), w.W, width));
  NMX_TRY(gemm_launch<T>(ctx, trans, 0, nb, width, nb, T(1), w.T_, nb, w.W, width, T(0), w.Y, width));
  return gemm_launch<T>(ctx, 0, 0, rows, width, nb, T(-1), w.V, nb, w.Y, width, T(1), C, ldc);
}

template <typename T>
int geqrf_impl(nmx_ctx* ctx, int m, int n, T* A, int lda, T* tau, int block) {
  void* ws = nullptr;
  NMX_TRY(nmx_workspace(ctx, reflector_bytes<T>(m, block, n), &ws));
  const Reflector<T> w = carve<T>(ws, m, block, n);
  for (int k = 0; k < n; k += block) {
    const int nb = std::min(block, n - k);
    {
      ProfScope ps(ctx, K_QR_PANEL, 4.0 * (double)(m - k) * nb * nb, 2.0 * sizeof(T) * (double)(m - k) * nb * nb);
      geqr2_kernel<T><<<1, kQrThreads, 0, ctx->stream>>>(m, A, lda, tau, k, nb);
      NMX_HIP(ctx, hipGetLastError());
    }
    if (k + nb < n) {
      NMX_TRY(build_reflector(ctx, w, m, A, lda, tau, k, nb));
      NMX_TRY(apply_reflector(ctx, w, m - k, nb, n - k - nb, 1, A + (int64_t)k * lda + k + nb, lda));
    }
  }
  return NMX_OK;
}

// B (m x nrhs) := Q^T B (trans) or Q B, panels in the order each needs.
template <typename T>
int ormqr_impl(nmx_ctx* ctx, int trans, int m, int n, int nrhs, const T* QR, int lda, const T* tau, T* B, int ldb,
               int block) {
  void* ws = nullptr;
  NMX_TRY(nmx_workspace(ctx, reflector_bytes<T>(m, block, nrhs), &ws));
  const Reflector<T> w = carve<T>(ws, m, block, nrhs);
  const int steps = ceil_div(n, block);
  for (int s = 0; s < steps; s++) {
    const int k = (trans ? s : steps - 1 - s) * block;
    const int nb = std::min(block, n - k);
    NMX_TRY(build_reflector(ctx, w, m, QR, lda, tau, k, nb));
    NMX_TRY(apply_reflector(ctx, w, m - k, nb, nrhs, trans, B + (int64_t)k * ldb, ldb));
  }
  return NMX_OK;
}

int check_qr_args(nmx_ctx* ctx, int m, int n, int lda, int block) {
  if (!ctx) return -1;
  if (m < 0) return -2;
  if (n < 0 || n > m) return -3;
  if (lda < std::max(1, n)) return -5;
  if (block < 1 || block > kQrMaxBlock) return -7;
  return NMX_OK;
}

}  // namespace

namespace {

template <typename T>
int geqrf_api(nmx_ctx* ctx, int m, int n, T* A, int lda, T* tau, int block) {
  NMX_TRY(check_qr_args(ctx, m, n, lda, block));
  if (n == 0) return NMX_OK;
  if (!A) return -4;
  if (!tau) return -6;
  return geqrf_impl<T>(ctx, m, n, A, lda, tau, block);
}

template <typename T>
int ormqr_api(nmx_ctx* ctx, int trans, int m, int n, int nrhs, const T* QR, int lda, const T* tau, T* B, int ldb,
              int block) {
  if (!ctx) return -1;
  if (m < 0) return -3;
  if (n < 0 || n > m) return -4;
  if (nrhs < 0) return -5;
  if (lda < std::max(1, n)) return -7;
  if (ldb < std::max(1, nrhs)) return -10;
  if (block < 1 || block > kQrMaxBlock) return -11;
  if (n == 0 || nrhs == 0) return NMX_OK;
  if (!QR) return -6;
  if (!tau) return -8;
  if (!B) return -9;
  return ormqr_impl<T>(ctx, trans ? 1 : 0, m, n, nrhs, QR, lda, tau, B, ldb, block);
}

template <typename T>
int orgqr_api(nmx_ctx* ctx, int m, int n, const T* QR, int lda, const T* tau, T* Q, int ldq, int block) {
  NMX_TRY(check_qr_args(ctx, m, n, lda, block));
  if (ldq < std::max(1, n)) return -8;
  if (n == 0) return NMX_OK;
  if (!QR) return -4;
  if (!tau) return -6;
  if (!Q) return -7;
  identity_kernel<T><<<(unsigned)(((int64_t)m * n + 255) / 256), 256, 0, ctx->stream>>>(m, n, Q, ldq);
  NMX_HIP(ctx, hipGetLastError());
  // last panel first; columns left of the panel are unchanged (V^T e_j = 0)
  void* ws = nullptr;
  NMX_TRY(nmx_workspace(ctx, reflector_bytes<T>(m, block, n), &ws));
  const Reflector<T> w = carve<T>(ws, m, block, n);
  const int steps = ceil_div(n, block);
  for (int s = steps - 1; s >= 0; s--) {
    const int k = s * block;
    const int nb = std::min(block, n - k);
    NMX_TRY(build_reflector<T>(ctx, w, m, QR, lda, tau, k, nb));
    NMX_TRY(apply_reflector<T>(ctx, w, m - k, nb, n - k, 0, Q + (int64_t)k * ldq + k, ldq));
  }
  return NMX_OK;
}

template <typename T>
int qr_r_api(nmx_ctx* ctx, int n, const T* QR, int lda, T* R, int ldr) {
  if (!ctx) return -1;
  if (n < 0) return -2;
  if (lda < std::max(1, n)) return -4;
  if (ldr < std::max(1, n)) return -6;
  if (n == 0) return NMX_OK;
  if (!QR) return -3;
  if (!R) return -5;
  upper_kernel<T><<<(unsigned)(((int64_t)n * n + 255) / 256), 256, 0, ctx->stream>>>(n, QR, lda, R, ldr);
  NMX_HIP(ctx, hipGetLastError());
  return NMX_OK;
}

template <typename T>
int gels_api(nmx_ctx* ctx, int m, int n, int nrhs, T* A, int lda, T* tau, T* B, int ldb, int block) {
  if (!ctx) return -1;
  if (m < 0) return -2;
  if (n < 1 || n > m) return -3;
  if (nrhs < 0) return -4;
  if (lda < n) return -6;
  if (ldb < std::max(1, nrhs)) return -9;
  if (block < 1 || block > kQrMaxBlock) return -10;
  if (nrhs == 0) return NMX_OK;
  if (!A) return -5;
  if (!tau) return -7;
  if (!B) return -8;
  NMX_TRY(geqrf_impl<T>(ctx, m, n, A, lda, tau, block));
  NMX_TRY(ormqr_impl<T>(ctx, 1, m, n, nrhs, A, lda, tau, B, ldb, block));
  return trsm_left<T>(ctx, 1, 0, 0, n, nrhs, A, lda, B, ldb);
}

}  // namespace

NMX_API int nmx_geqrf_f64(nmx_ctx* ctx, int m, int n, double* A, int lda, double* tau, int block) {
  return geqrf_api<double>(ctx, m, n, A, lda, tau, block);
}
NMX_API int nmx_geqrf_f32(nmx_ctx* ctx, int m, int n, float* A, int lda, float* tau, int block) {
  return geqrf_api<float>(ctx, m, n, A, lda, tau, block);
}

NMX_API int nmx_ormqr_f64(nmx_ctx* ctx, int trans, int m, int n, int nrhs, const double* QR, int lda,
                          const double* tau, double* B, int ldb, int block) {
  return ormqr_api<double>(ctx, trans, m, n, nrhs, QR, lda, tau, B, ldb, block);
}
NMX_API int nmx_ormqr_f32(nmx_ctx* ctx, int trans, int m, int n, int nrhs, const float* QR, int lda,
                          const float* tau, float* B, int ldb, int block) {
  return ormqr_api<float>(ctx, trans, m, n, nrhs, QR, lda, tau, B, ldb, block);
}

NMX_API int nmx_orgqr_f64(nmx_ctx* ctx, int m, int n, const double* QR, int lda, const double* tau, double* Q,
                          int ldq, int block) {
  return orgqr_api<double>(ctx, m, n, QR, lda, tau, Q, ldq, block);
}
NMX_API int nmx_orgqr_f32(nmx_ctx* ctx, int m, int n, const float* QR, int lda, const float* tau, float* Q, int ldq,
                          int block) {
  return orgqr_api<float>(ctx, m, n, QR, lda, tau, Q, ldq, block);
}

NMX_API int nmx_qr_r_f64(nmx_ctx* ctx, int n, const double* QR, int lda, double* R, int ldr) {
  return qr_r_api<double>(ctx, n, QR, lda, R, ldr);
}
NMX_API int nmx_qr_r_f32(nmx_ctx* ctx, int n, const float* QR, int lda, float* R, int ldr) {
  return qr_r_api<float>(ctx, n, QR, lda, R, ldr);
}

NMX_API int nmx_gels_f64(nmx_ctx* ctx, int m, int n, int nrhs, double* A, int lda, double* tau, double* B, int ldb,
                         int block) {
  return gels_api<double>(ctx, m, n, nrhs, A, lda, tau, B, ldb, block);
}
NMX_API int nmx_gels_f32(nmx_ctx* ctx, int m, int n, int nrhs, float* A, int lda, float* tau, float* B, int ldb,
                         int block) {
  return gels_api<float>(ctx, m, n, nrhs, A, lda, tau, B, ldb, block);
}
