// Batched tall-skinny least squares, FP32: `count` independent problems
// min ||A x - b||, A m x n row-major (n <= 32, m >= n), one wavefront each.
//
// numax's `lstsq` (numax/linalg/qr.mojo:746-821) is blocked Householder QR
// (geqr2_panel, panel.mojo:1975-2073, with LAPACK dlarfg's convention:
// beta = -sign(alpha) sqrt(alpha^2 + sigma), tau = (beta - alpha) / beta,
// v = x / (alpha - beta), tau = 0 when sigma = 0), Q^T b applied panel by
// panel (qr.mojo:470-534), then the upper-triangular solve with R
// (qr.mojo:536-612).  Here the same reflectors are applied as a streaming
// (TSQR-style) factorization: the wavefront keeps the running 32 x 33
// triangle [R | Q^T b] in lanes 0..31 and brings the next 32 rows of [A | b]
// into lanes 32..63; each of the 32 column steps builds one reflector over
// row k of the triangle and the 32 new rows (the triangle is zero below its
// diagonal), so the 1024 x 33 problem is read exactly once.  Column sums
// across the wavefront use recursive halving; the reflector's coefficients
// are broadcast from lanes through scalar registers.
#include <math.h>

#include "nmx_internal.h"

namespace {

constexpr int kCols = 33;  // 32 columns of A, then b

// Sum NV per-lane partials over the 64 lanes (recursive halving); lane l
// ends with the total of partial l / (64 / NV).
template <int NV>
__device__ __forceinline__ float halving_sum(float (&s)[NV], int lane) {
  int m = 32;
#pragma unroll
  for (int w = NV / 2; w >= 1; w >>= 1, m >>= 1) {
    const bool hi = (lane & m) != 0;
#pragma unroll
    for (int k = 0; k < w; k++) {
      const float keep = hi ? s[w + k] : s[k];
      const float send = hi ? s[k] : s[w + k];
      s[k] = keep + __shfl_xor(send, m);
    }
  }
  float v = s[0];
#pragma unroll
  for (; m >= 1; m >>= 1) v += __shfl_xor(v, m);
  return v;
}

constexpr int pow2_at_least(int x) { return x <= 1 ? 1 : 2 * pow2_at_least((x + 1) / 2); }

// One reflector: column K of the stacked [triangle; new rows], applied to
// columns K+1 .. 32 (b included).  `in` marks the lanes the reflector spans
// (lane K of the triangle and every new row).
template <int K>
__device__ __forceinline__ void reflect(float (&r)[kCols], int lane) {
  constexpr int NJ = kCols - 1 - K;  // columns to the right, b included
  constexpr int P2 = pow2_at_least(NJ);
  const bool below = lane >= 32;
  const float alpha = readlane_f(r[K], K);
  // sigma = sum over the new rows of x^2
  float sq[1] = {below ? r[K] * r[K] : 0.f};
  const float sigma = readlane_f(halving_sum<1>(sq, lane), 0);
  float tau = 0.f, beta = alpha, scale = 0.f;
  if (sigma != 0.f) {
    beta = -copysignf(sqrtf(alpha * alpha + sigma), alpha);
    tau = (beta - alpha) / beta;
    scale = 1.f / (alpha - beta);
  }
  // v: 1 on row K of the triangle, x * scale on the new rows, 0 elsewhere
  const float vl = below ? r[K] * scale : (lane == K ? 1.f : 0.f);
  if (below) r[K] = 0.f;           // annihilated by the reflector
  if (lane == K) r[K] = beta;      // R_kk
  if (NJ > 0 && tau != 0.f) {
    // w_j = sum_lanes v * r[j], j = K+1 .. 32
    float p[P2];
#pragma unroll
    for (int j = 0; j < P2; j++) p[j] = (j < NJ) ? vl * r[K + 1 + j] : 0.f;
    const float tot = halving_sum<P2>(p, lane);
    constexpr int SPREAD = 64 / P2;
#pragma unroll
    for (int j = 0; j < NJ; j++) {
      const float wj = readlane_f(tot, j * SPREAD);
      r[K + 1 + j] -= tau * vl * wj;
    }
  }
}

template <int K>
__device__ __forceinline__ void reflect_all(float (&r)[kCols], int lane) {
  if constexpr (K < 32) {
    reflect<K>(r, lane);
    reflect_all<K + 1>(r, lane);
  }
}

__global__ void __launch_bounds__(256) gels_tsqr_kernel(int count, int m, int n, const float* __restrict__ A,
                                                        const float* __restrict__ B, float* __restrict__ X) {
  const int lane = threadIdx.x & 63;
  const int waves = gridDim.x * (blockDim.x / 64);
  for (int p = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6); p < count; p += waves) {
    const float* Ap = A + (int64_t)p * m * n;
    const float* Bp = B + (int64_t)p * m;
    const __amdgpu_buffer_rsrc_t ra = buf_rsrc(Ap);
    const __amdgpu_buffer_rsrc_t rb = buf_rsrc(Bp);
    float r[kCols];
    // the triangle starts as [I_pad | 0]: a unit diagonal on the columns past
    // n keeps R non-singular there, and those unknowns come out 0
#pragma unroll
    for (int j = 0; j < kCols; j++) r[j] = (lane < 32 && j == lane && j >= n && j < 32) ? 1.f : 0.f;
    for (int r0 = 0; r0 < m; r0 += 32) {
      // next 32 rows of [A | b] into lanes 32..63 (zero past m / past n)
      const int row = r0 + lane - 32;
      const bool ld = lane >= 32 && row < m;
      const uint32_t rowoff = (uint32_t)(ld ? row : 0) * (uint32_t)n;
#pragma unroll
      for (int j = 0; j < 32; j++) {
        const float a = buf_load<float>(ra, (ld & (j < n)) ? 4u * (rowoff + (uint32_t)j) : kBufOob);
        if (lane >= 32) r[j] = a;
      }
      const float bv = buf_load<float>(rb, ld ? 4u * (uint32_t)row : kBufOob);
      if (lane >= 32) r[32] = bv;
      reflect_all<0>(r, lane);
    }
    // R x = (Q^T b)[0:32] by back substitution (qr.mojo:536-612's _trsv)
    float acc = r[32];
#pragma unroll
    for (int c = 31; c >= 0; c--) {
      const float xc = readlane_f(acc, c) / readlane_f(r[c], c);
      if (lane == c) acc = xc;
      else if (lane < c) acc -= r[c] * xc;
    }
    if (lane < n) X[(int64_t)p * n + lane] = acc;
  }
}

// ---------------------------------------------------------------------------
// Flat Householder for m <= 1024: one 256-thread workgroup per problem, the
// whole m x 33 [A | b] in registers (thread t holds rows t, t+256, t+512,
// t+768), one reflector per column over all m rows -- 32 steps per problem
// instead of the streaming form's 32 per 32 rows.  Per step: sigma and the
// 33 - K column dot products reduced in the wavefront (recursive halving)
// and across the four wavefronts through LDS (one barrier each), alpha from
// the row-K thread through LDS; the update is four rows per thread.  The
// reflectors are geqr2_panel's (the same dlarfg convention as above), so R
// and Q^T b are numax's up to rounding.  Rows are loaded coalesced (eight
// rows per wavefront instruction) and transposed through LDS.
// ---------------------------------------------------------------------------
constexpr int kFlatT = 256;               // threads per problem
constexpr int kFlatR = 4;                 // rows per thread
constexpr int kFlatMaxM = kFlatT * kFlatR;
constexpr int kFlatWaves = kFlatT / 64;

struct FlatLds {
  float stage[64][kCols + 3];    // 64 rows of [A | b] per load round (stride 36: 16-byte aligned)
  // per step (double-buffered): each wavefront's column sums (sigma last),
  // then row K's own values
  float part[2][kFlatWaves + 1][kCols];
  float alpha[2];
};

template <int K>
__device__ __forceinline__ void flat_step(float (&r)[kFlatR][kCols], int t, int n, FlatLds& L) {
  if (K >= n) return;  // uniform
  constexpr int NJ = kCols - 1 - K;    // columns to the right, b included
  constexpr int P2 = pow2_at_least(NJ);
  constexpr int SPREAD = 64 / P2;
  const int lane = t & 63, wave = t >> 6;
  constexpr int buf = K & 1;
  // with x the column below K: sigma = x^T x and u_j = x^T A[:, j] now; the
  // reflector's v is 1 on row K and x * scale below, so v^T A[:, j] =
  // A[K][j] + scale * u_j once alpha gives scale
  float sg = 0.f;
  float p[P2];
#pragma unroll
  for (int q = 0; q < P2; q++) p[q] = 0.f;
#pragma unroll
  for (int i = 0; i < kFlatR; i++) {
    const int row = t + kFlatT * i;
    const float x = row > K ? r[i][K] : 0.f;
    sg += x * x;
#pragma unroll
    for (int j = 0; j < NJ; j++) p[j] += x * r[i][K + 1 + j];
    if (row == K) {
      L.alpha[buf] = r[i][K];
#pragma unroll
      for (int j = 0; j < NJ; j++) L.part[buf][kFlatWaves][j] = r[i][K + 1 + j];  // row K's A[K][j]
    }
  }
  const float tot = halving_sum<P2>(p, lane);
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) sg += __shfl_xor(sg, m);
  if (lane % SPREAD == 0 && lane / SPREAD < NJ) L.part[buf][wave][lane / SPREAD] = tot;
  if (lane == 0) L.part[buf][wave][kCols - 1] = sg;
  __syncthreads();
  float sigma = 0.f;
#pragma unroll
  for (int w = 0; w < kFlatWaves; w++) sigma += L.part[buf][w][kCols - 1];
  const float alpha = L.alpha[buf];
  if (sigma == 0.f) return;  // uniform: tau = 0, column K already triangular
  const float beta = -copysignf(sqrtf(alpha * alpha + sigma), alpha);
  const float tau = (beta - alpha) / beta;
  const float scale = 1.f / (alpha - beta);
  float f[NJ];  // tau * v^T A[:, j]
#pragma unroll
  for (int j = 0; j < NJ; j++) {
    float u = 0.f;
#pragma unroll
    for (int w = 0; w < kFlatWaves; w++) u += L.part[buf][w][j];
    f[j] = tau * (L.part[buf][kFlatWaves][j] + scale * u);
  }
#pragma unroll
  for (int i = 0; i < kFlatR; i++) {
    const int row = t + kFlatT * i;
    if (row == K) {
      r[i][K] = beta;
#pragma unroll
      for (int j = 0; j < NJ; j++) r[i][K + 1 + j] -= f[j];
    } else if (row > K) {
      const float v = r[i][K] * scale;
#pragma unroll
      for (int j = 0; j < NJ; j++) r[i][K + 1 + j] -= f[j] * v;
    }
  }
}

template <int K>
__device__ __forceinline__ void flat_all(float (&r)[kFlatR][kCols], int t, int n, FlatLds& L) {
  if constexpr (K < 32) {
    flat_step<K>(r, t, n, L);
    flat_all<K + 1>(r, t, n, L);
  }
}

__global__ void __launch_bounds__(kFlatT) gels_flat_kernel(int count, int m, int n, const float* __restrict__ A,
                                                           const float* __restrict__ B, float* __restrict__ X) {
  __shared__ FlatLds L;
  const int t = threadIdx.x;
  for (int p = blockIdx.x; p < count; p += gridDim.x) {
    const float* Ap = A + (int64_t)p * m * n;
    const float* Bp = B + (int64_t)p * m;
    float r[kFlatR][kCols];
    // 64 rows per round: thread t loads element t % 64 .. of rows (t / 64) + 4 k
    // -- consecutive threads read consecutive elements of the packed problem
#pragma unroll
    for (int i = 0; i < kFlatR; i++) {
#pragma unroll
      for (int round = 0; round < 4; round++) {
        const int r0 = i * kFlatT + round * 64;  // first row of this round
        __syncthreads();
        for (int e = t; e < 64 * n; e += kFlatT) {
          const int rr = e / n, cc = e % n;
          L.stage[rr][cc] = r0 + rr < m ? Ap[(int64_t)(r0 + rr) * n + cc] : 0.f;
        }
        if (t < 64) L.stage[t][32] = r0 + t < m ? Bp[r0 + t] : 0.f;
        __syncthreads();
        if ((t >> 6) == round) {
#pragma unroll
          for (int c = 0; c < kCols; c++) r[i][c] = c < n || c == 32 ? L.stage[t & 63][c] : 0.f;
        }
      }
    }
    __syncthreads();
    flat_all<0>(r, t, n, L);
    // R x = (Q^T b)[0:n] in wavefront 0 (rows 0..31 are its lanes' first rows)
    if (t < 64) {
      float acc = t < n ? r[0][32] : 0.f;
#pragma unroll
      for (int c = 31; c >= 0; c--) {
        if (c < n) {  // uniform
          const float xc = readlane_f(acc, c) / readlane_f(r[0][c], c);
          if (t == c) acc = xc;
          else if (t < c) acc -= r[0][c] * xc;
        }
      }
      if (t < n) X[(int64_t)p * n + t] = acc;
    }
    __syncthreads();
  }
}

}  // namespace

NMX_API int nmx_batched_gels_f32(nmx_ctx* ctx, int count, int m, int n, const float* A, const float* B, float* X) {
  if (!ctx) return -1;
  if (count < 0) return -2;
  if (m < 1) return -3;
  if (n < 1 || n > 32 || n > m) return -4;
  if (count == 0) return NMX_OK;
  if (!A) return -5;
  if (!B) return -6;
  if (!X) return -7;
  if ((int64_t)m * n * 4 >= 0x7FFFFFFF) return NMX_ERR_UNSUPPORTED;
  ProfScope ps(ctx, K_TSQR, 0.0, 4.0 * (double)count * ((double)m * n + m + n));
  if (m <= kFlatMaxM) {
    // one workgroup per problem, the whole problem in registers
    const int grid = std::min(count, ctx->num_cus * 16);
    gels_flat_kernel<<<grid, kFlatT, 0, ctx->stream>>>(count, m, n, A, B, X);
  } else {
    // taller problems stream through one wavefront, 32 rows at a time
    const int waves = 4;
    int grid = ceil_div(count, waves);
    if (grid > ctx->num_cus * 8) grid = ctx->num_cus * 8;
    gels_tsqr_kernel<<<grid, 64 * waves, 0, ctx->stream>>>(count, m, n, A, B, X);
  }
  NMX_HIP(ctx, hipGetLastError());
  return NMX_OK;
}
