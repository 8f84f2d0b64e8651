// Mixed-precision LU with iterative refinement (no reference counterpart;
// the reference has no mixed precision and no refinement).
//
// Factor:  A' = W^T A V with W, V depth-2 random butterflies (Parker's RBT;
//          Baboulin, Dongarra et al.), stored FP32, then blocked LU without
//          pivoting.  The diagonal block and both panel solves run in FP32;
//          the trailing update A22 -= L21 U12 -- the cubic term -- runs on
//          FP16 or BF16 MFMA with FP32 accumulation, the operands scaled by
//          powers of two so their largest entry sits at 2^14 (Higham, Pranesh
//          & Zounon 2019).  The "split" update keeps a second FP16 term per
//          operand (x ~ hi + 2^-11 lo) and adds the two cross products, which
//          recovers ~FP32 accuracy for 3x the MFMA work -- free while the
//          update is bound by its FP32 C traffic.
// Refine:  r = b - A x in FP64 on the original matrix, d = V U^-1 L^-1 W^T r
//          in FP32, x += d in FP64, until the normwise backward error meets
//          the target.  Past the point where the low-precision factor stops
//          contracting the error (kappa u_f ~ 1, kappa ~ 1e4 for FP16), the
//          correction equation A d = r is solved instead by GMRES in
//          FP64, left-preconditioned by the same factor (GMRES-IR, Carson and
//          Higham, SIAM J. Sci. Comput. 40(2), 2018): the Krylov basis,
//          the Arnoldi products (classical Gram-Schmidt, twice) and the
//          matrix products are FP64; the preconditioner solves are FP32.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "nmx_internal.h"

namespace {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));

constexpr float kRsqrt2 = 0.70710678118654752440f;
constexpr float kPivotFloorF = 1e-30f;

// ---------------------------------------------------------------------------
// Random butterflies.  A depth-2 butterfly closes over index groups
// {i, i+h, i+2h, i+3h} (h = n/4), so W^T A V is computed 4x4 block by 4x4
// block: T_V on each row, then T_W on each column.  Diagonals per transform
// (2n floats): r0, r1 (level 0, n/2 each), s0, s1, t0, t1 (level 1, n/4 each).
// ---------------------------------------------------------------------------
struct Bfly {
  const float* r0;
  const float* r1;
  const float* s0;
  const float* s1;
  const float* t0;
  const float* t1;
};

__device__ __forceinline__ Bfly bfly_of(const float* d, int n) {
  const int half = n / 2, q = n / 4;
  return Bfly{d, d + half, d + 2 * half, d + 2 * half + q, d + 2 * half + 2 * q, d + 2 * half + 3 * q};
}

// y <- B^T y on group i (the transpose transform).
template <typename T>
__device__ __forceinline__ void bfly_t(T y[4], const Bfly& f, int i, int h) {
  T z0 = f.s0[i] * (y[0] + y[1]) * kRsqrt2;
  T z1 = f.s1[i] * (y[0] - y[1]) * kRsqrt2;
  T z2 = f.t0[i] * (y[2] + y[3]) * kRsqrt2;
  T z3 = f.t1[i] * (y[2] - y[3]) * kRsqrt2;
  y[0] = f.r0[i] * (z0 + z2) * kRsqrt2;
  y[2] = f.r1[i] * (z0 - z2) * kRsqrt2;
  y[1] = f.r0[i + h] * (z1 + z3) * kRsqrt2;
  y[3] = f.r1[i + h] * (z1 - z3) * kRsqrt2;
}

// y <- B y on group i (level 0 first, then level 1).
template <typename T>
__device__ __forceinline__ void bfly_n(T y[4], const Bfly& f, int i, int h) {
  T z0 = (f.r0[i] * y[0] + f.r1[i] * y[2]) * kRsqrt2;
  T z2 = (f.r0[i] * y[0] - f.r1[i] * y[2]) * kRsqrt2;
  T z1 = (f.r0[i + h] * y[1] + f.r1[i + h] * y[3]) * kRsqrt2;
  T z3 = (f.r0[i + h] * y[1] - f.r1[i + h] * y[3]) * kRsqrt2;
  y[0] = (f.s0[i] * z0 + f.s1[i] * z1) * kRsqrt2;
  y[1] = (f.s0[i] * z0 - f.s1[i] * z1) * kRsqrt2;
  y[2] = (f.t0[i] * z2 + f.t1[i] * z3) * kRsqrt2;
  y[3] = (f.t0[i] * z2 - f.t1[i] * z3) * kRsqrt2;
}

// A' (FP32) = W^T A V for the leading n4 = 4*floor(n/4) block; the tail rows
// and columns pass through (the transform is diag(W, I) / diag(V, I)).
__global__ void rbt_matrix_kernel(int n, const double* __restrict__ A, int lda, float* __restrict__ Ap, int ldp,
                                  const float* __restrict__ dw, const float* __restrict__ dv) {
  const int n4 = n & ~3, h = n4 / 4;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  // groups: h x h interior blocks, then the pass-through tail
  if (idx < (int64_t)h * h) {
    const int gi = (int)(idx / h), gj = (int)(idx % h);
    const Bfly fw = bfly_of(dw, n4), fv = bfly_of(dv, n4);
    float x[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++) {
      float row[4];
#pragma unroll
      for (int b = 0; b < 4; b++) row[b] = (float)A[(int64_t)(gi + a * h) * lda + gj + b * h];
      bfly_t(row, fv, gj, h);  // (A V) row = V^T applied to the row as a column
#pragma unroll
      for (int b = 0; b < 4; b++) x[a][b] = row[b];
    }
#pragma unroll
    for (int b = 0; b < 4; b++) {
      float col[4] = {x[0][b], x[1][b], x[2][b], x[3][b]};
      bfly_t(col, fw, gi, h);
#pragma unroll
      for (int a = 0; a < 4; a++) Ap[(int64_t)(gi + a * h) * ldp + gj + b * h] = col[a];
    }
    return;
  }
  // tail: rows >= n4 (all columns) and columns >= n4 of rows < n4
  int64_t t = idx - (int64_t)h * h;
  const int tail = n - n4;
  if (tail == 0) return;
  const int64_t rows_tail = (int64_t)tail * n;
  if (t < rows_tail) {
    int r = n4 + (int)(t / n), c = (int)(t % n);
    // row r untransformed on the left; columns transformed by V on groups
    if (c < n4) {
      const Bfly fv = bfly_of(dv, n4);
      int gj = c % h, b = c / h;
      float row[4];
#pragma unroll
      for (int q = 0; q < 4; q++) row[q] = (float)A[(int64_t)r * lda + gj + q * h];
      bfly_t(row, fv, gj, h);
      Ap[(int64_t)r * ldp + c] = row[b];
    } else {
      Ap[(int64_t)r * ldp + c] = (float)A[(int64_t)r * lda + c];
    }
    return;
  }
  t -= rows_tail;
  if (t < (int64_t)n4 * tail) {
    int r = (int)(t / tail), c = n4 + (int)(t % tail);
    const Bfly fw = bfly_of(dw, n4);
    int gi = r % h, a = r / h;
    float col[4];
#pragma unroll
    for (int q = 0; q < 4; q++) col[q] = (float)A[(int64_t)(gi + q * h) * lda + c];
    bfly_t(col, fw, gi, h);
    Ap[(int64_t)r * ldp + c] = col[a];
  }
}

// v (FP64 in, FP32 out) <- W^T v, per column of an n x nrhs block.
__global__ void rbt_vec_t_kernel(int n, int nrhs, const double* __restrict__ v, int ldv, float* __restrict__ out,
                                 int ldo, const float* __restrict__ dw) {
  const int n4 = n & ~3, h = n4 / 4;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t groups = (int64_t)h * nrhs;
  if (idx < groups) {
    const int gi = (int)(idx / nrhs), c = (int)(idx % nrhs);
    const Bfly fw = bfly_of(dw, n4);
    float y[4];
#pragma unroll
    for (int a = 0; a < 4; a++) y[a] = (float)v[(int64_t)(gi + a * h) * ldv + c];
    bfly_t(y, fw, gi, h);
#pragma unroll
    for (int a = 0; a < 4; a++) out[(int64_t)(gi + a * h) * ldo + c] = y[a];
  } else {
    int64_t t = idx - groups;
    if (t < (int64_t)(n - n4) * nrhs) {
      int r = n4 + (int)(t / nrhs), c = (int)(t % nrhs);
      out[(int64_t)r * ldo + c] = (float)v[(int64_t)r * ldv + c];
    }
  }
}

// x (FP64) += V y (y FP32), per column.
__global__ void rbt_vec_n_add_kernel(int n, int nrhs, const float* __restrict__ y, int ldy, double* __restrict__ x,
                                     int ldx, const float* __restrict__ dv, int accumulate,
                                     const int* __restrict__ stop) {
  if (stop && *stop) return;
  const int n4 = n & ~3, h = n4 / 4;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t groups = (int64_t)h * nrhs;
  if (idx < groups) {
    const int gi = (int)(idx / nrhs), c = (int)(idx % nrhs);
    const Bfly fv = bfly_of(dv, n4);
    double z[4];
#pragma unroll
    for (int a = 0; a < 4; a++) z[a] = (double)y[(int64_t)(gi + a * h) * ldy + c];
    bfly_n(z, fv, gi, h);
#pragma unroll
    for (int a = 0; a < 4; a++) {
      double* p = x + (int64_t)(gi + a * h) * ldx + c;
      *p = accumulate ? *p + z[a] : z[a];
    }
  } else {
    int64_t t = idx - groups;
    if (t < (int64_t)(n - n4) * nrhs) {
      int r = n4 + (int)(t / nrhs), c = (int)(t % nrhs);
      double* p = x + (int64_t)r * ldx + c;
      double z = (double)y[(int64_t)r * ldy + c];
      *p = accumulate ? *p + z : z;
    }
  }
}

// Butterfly diagonals exp(rho/10), rho ~ U[-1/2, 1/2) from Philox.
__device__ __forceinline__ uint32_t philox_word(uint64_t seed, uint64_t idx) {
  uint32_t c[4] = {(uint32_t)(idx >> 2), (uint32_t)(idx >> 34), 0x52425400u, 0u};
  uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  for (int r = 0; r < 10; r++) {
    if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
  }
  return c[idx & 3];
}

__global__ void rbt_diag_kernel(int count, uint64_t seed, float* __restrict__ d) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  float u = (float)(philox_word(seed, (uint64_t)i) >> 8) * 5.9604644775390625e-08f - 0.5f;
  d[i] = expf(u * 0.1f);
}

// ---------------------------------------------------------------------------
// No-pivot diagonal leaf (nb <= 64) in LDS: right-looking LU, an exact-zero
// pivot floored to 1e-30 with its first occurrence in info.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) getf2_nopiv_kernel(float* __restrict__ A, int lda, int k0, int nb,
                                                          int32_t* __restrict__ info) {
  __shared__ float S[64][65];
  const int t = threadIdx.x;
  for (int idx = t; idx < nb * nb; idx += 256) S[idx / nb][idx % nb] = A[(int64_t)(k0 + idx / nb) * lda + k0 + idx % nb];
  __syncthreads();
  for (int j = 0; j < nb; j++) {
    float p = S[j][j];
    if (p == 0.0f) {
      if (t == 0 && *info == 0) *info = k0 + j + 1;
      p = kPivotFloorF;
    }
    for (int i = j + 1 + t; i < nb; i += 256) S[i][j] = S[i][j] / p;
    __syncthreads();
    for (int idx = t; idx < (nb - j - 1) * (nb - j - 1); idx += 256) {
      int i = j + 1 + idx / (nb - j - 1), c = j + 1 + idx % (nb - j - 1);
      S[i][c] = S[i][c] - S[i][j] * S[j][c];
    }
    __syncthreads();
  }
  for (int idx = t; idx < nb * nb; idx += 256) A[(int64_t)(k0 + idx / nb) * lda + k0 + idx % nb] = S[idx / nb][idx % nb];
}

// ---------------------------------------------------------------------------
// Trailing-update operands: diagonal power-of-two scaling, then FP16 / BF16
// (hi [, lo]) copies.  L21 stays m x k (row-major); U12 (k x m) is stored
// transposed (m x k) so both MFMA operands read 8 consecutive K values.
// Each block of 32 rows of L21 and each column of U12 gets its own scale
// (the product L21 U12 = D_r^-1 (D_r L21)(U12 D_c) D_c^-1 is unscaled in the
// GEMM's epilogue, where a lane's 16 outputs of an MFMA tile share one row
// block and one column), so a block or column far below the operand's
// largest entry still fills the 16-bit format's range instead of losing bits
// to its exponent floor.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float pow2_scale(unsigned maxbits) {
  float m = __uint_as_float(maxbits);
  if (!(m > 0.f) || !isfinite(m)) return 1.0f;
  int e;
  frexpf(m, &e);          // m = f * 2^e, f in [0.5, 1)
  return ldexpf(1.0f, 14 - e);  // max * scale in [2^13, 2^14)
}

// Scales of one update step: workgroups [0, ga) take L21's 32-row blocks
// (one workgroup a block, a row per thread group of 8), the rest U12's
// columns (64 columns a workgroup, its four wavefronts splitting the k rows,
// combined in LDS).  rs[b] and cs[c] are pow2_scale of the block's /
// column's largest magnitude (NaN never raises it); rs[ld + b] and cs[ld + c]
// hold the reciprocals (exact: powers of two).
__global__ void __launch_bounds__(256) scales2_kernel(int ga, int m, int k, const float* __restrict__ L21, int ldl,
                                                      float* __restrict__ rs, const float* __restrict__ U12, int ldu,
                                                      float* __restrict__ cs, int ld) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if ((int)blockIdx.x < ga) {
    __shared__ float red[4];
    const int rb = blockIdx.x;  // rows 32 rb .. 32 rb + 31
    float mx = 0.f;
    for (int r = 32 * rb + (threadIdx.x >> 3); r < min(m, 32 * rb + 32); r += 32) {
      const float* row = L21 + (int64_t)r * ldl;
      for (int c = threadIdx.x & 7; c < k; c += 8) {
        const float v = fabsf(row[c]);
        if (v > mx) mx = v;
      }
    }
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
    if (lane == 0) red[w] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
      const float sc = pow2_scale(__float_as_uint(fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]))));
      rs[rb] = sc;
      rs[ld + rb] = 1.0f / sc;
    }
  } else {
    __shared__ float part[4][64];
    const int c = (blockIdx.x - ga) * 64 + lane;
    float mx = 0.f;
    if (c < m) {
      for (int r = w; r < k; r += 4) {
        const float v = fabsf(U12[(int64_t)r * ldu + c]);
        if (v > mx) mx = v;
      }
    }
    part[w][lane] = mx;
    __syncthreads();
    if (w == 0 && c < m) {
      mx = fmaxf(fmaxf(part[0][lane], part[1][lane]), fmaxf(part[2][lane], part[3][lane]));
      const float sc = pow2_scale(__float_as_uint(mx));
      cs[c] = sc;
      cs[ld + c] = 1.0f / sc;
    }
  }
}

template <int BF>
struct Lo;
template <>
struct Lo<0> {
  typedef _Float16 t;
  static __device__ __forceinline__ t cvt(float x) { return (_Float16)x; }
  static __device__ __forceinline__ float back(t x) { return (float)x; }
};
template <>
struct Lo<1> {
  typedef __bf16 t;
  static __device__ __forceinline__ t cvt(float x) { return (__bf16)x; }
  static __device__ __forceinline__ float back(t x) { return (float)x; }
};

// dst[r][c] = sc[r >> 5] src[r][c] (transpose = 0: a scale per 32-row block)
// or sc[r] src[c][r] (transpose = 1: a scale per row of dst, a column of src);
// dst is rows x cols row-major with ld = cols.
template <int BF, int SPLIT>
__device__ __forceinline__ void convert_part(int bid, int nblk, int rows, int cols, const float* __restrict__ src,
                                             int lds, int transpose, const float* __restrict__ sc,
                                             typename Lo<BF>::t* __restrict__ hi, typename Lo<BF>::t* __restrict__ lo) {
  for (int64_t idx = (int64_t)bid * blockDim.x + threadIdx.x; idx < (int64_t)rows * cols;
       idx += (int64_t)nblk * blockDim.x) {
    int r = (int)(idx / cols), c = (int)(idx % cols);
    float v = (transpose ? src[(int64_t)c * lds + r] * sc[r] : src[(int64_t)r * lds + c] * sc[r >> 5]);
    typename Lo<BF>::t h = Lo<BF>::cvt(v);
    hi[idx] = h;
    if (SPLIT) lo[idx] = Lo<BF>::cvt((v - Lo<BF>::back(h)) * 2048.0f);
  }
}

// A (rows x cols, as is) on workgroups [0, ga), B (transposed) on the rest
template <int BF, int SPLIT>
__global__ void convert2_kernel(int ga, int rows, int cols, const float* __restrict__ A, int lda,
                                const float* __restrict__ sA, typename Lo<BF>::t* __restrict__ ahi,
                                typename Lo<BF>::t* __restrict__ alo, const float* __restrict__ B, int ldb,
                                const float* __restrict__ sB, typename Lo<BF>::t* __restrict__ bhi,
                                typename Lo<BF>::t* __restrict__ blo) {
  if ((int)blockIdx.x < ga)
    convert_part<BF, SPLIT>(blockIdx.x, ga, rows, cols, A, lda, 0, sA, ahi, alo);
  else
    convert_part<BF, SPLIT>(blockIdx.x - ga, gridDim.x - ga, rows, cols, B, ldb, 1, sB, bhi, blo);
}

// ---------------------------------------------------------------------------
// C (FP32, in place) -= (A . B^T) iA[row / 32] iB[col] (the reciprocal scales) with A m x k, B n x k in FP16 /
// BF16, FP32 accumulation on v_mfma_f32_32x32x16_{f16,bf16}.  128x128 C tile
// per 512-thread workgroup, 64x32 per wavefront (2x1 MFMA tiles, so four
// wavefronts fit a SIMD's registers and hide the slab loads); the C tile
// is read after the products, its latency hidden by the other wavefronts.
// SPLIT adds 2^-11 (A_hi B_lo^T + A_lo B_hi^T).
// ---------------------------------------------------------------------------
template <int BF>
struct Mma;
template <>
struct Mma<0> {
  static __device__ __forceinline__ f16v op(h8 a, h8 b, f16v c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
  }
};
template <>
struct Mma<1> {
  static __device__ __forceinline__ f16v op(b8 a, b8 b, f16v c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
  }
};

constexpr int MT = 128, KT = 32, KP = 8;  // C tile, K slab, LDS pad (halves)
constexpr int kMpThreads = 512;           // 8 wavefronts, each a 64 x 32 piece of the C tile

template <int BF, int SPLIT>
__global__ void __launch_bounds__(kMpThreads) mp_gemm_kernel(int M, int N, int K, const typename Lo<BF>::t* __restrict__ Ah,
                                                      const typename Lo<BF>::t* __restrict__ Al,
                                                      const typename Lo<BF>::t* __restrict__ Bh,
                                                      const typename Lo<BF>::t* __restrict__ Bl,
                                                      const float* __restrict__ iA,
                                                      const float* __restrict__ iB, float* __restrict__ C, int ldc,
                                                      int tilesM, int tilesN) {
  typedef typename Lo<BF>::t lt;
  typedef lt v8 __attribute__((ext_vector_type(8)));
  constexpr int NOP = SPLIT ? 2 : 1;
  __shared__ __attribute__((aligned(16))) lt As[2][NOP][MT][KT + KP];
  __shared__ __attribute__((aligned(16))) lt Bs[2][NOP][MT][KT + KP];

  const int total = tilesM * tilesN;
  const int b = blockIdx.x;
  const int per = (total + kXcds - 1) / kXcds;
  const int lin = (b % kXcds) * per + b / kXcds;
  if (lin >= total) return;
  const int GM = 8;
  const int group = lin / (GM * tilesN);
  const int first = group * GM;
  const int gm = min(tilesM - first, GM);
  const int in = lin % (GM * tilesN);
  const int m0 = (first + in % gm) * MT, n0 = (in / gm) * MT;

  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave >> 2, wn = wave & 3;  // 2 x 4 wavefronts of 64 x 32

  f16v acc[2][1], accx[2][1];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 1; j++) {
      acc[i][j] = f16v{};
      accx[i][j] = f16v{};
    }

  // global -> registers -> LDS, double-buffered: slab s+1 is in flight while
  // slab s feeds the MFMAs.  Each slab is 128 rows x 32 K values per operand
  // = 512 chunks of 8, one per thread.
  const __amdgpu_buffer_rsrc_t rAh = buf_rsrc(Ah + (int64_t)m0 * K);
  const __amdgpu_buffer_rsrc_t rBh = buf_rsrc(Bh + (int64_t)n0 * K);
  const __amdgpu_buffer_rsrc_t rAl = buf_rsrc((SPLIT ? Al : Ah) + (int64_t)m0 * K);
  const __amdgpu_buffer_rsrc_t rBl = buf_rsrc((SPLIT ? Bl : Bh) + (int64_t)n0 * K);
  v8 pa[NOP][1], pb[NOP][1];
  auto load_slab = [&](int k0) {
#pragma unroll
    for (int c = 0; c < 1; c++) {
      const int chunk = t + c * kMpThreads;
      const int r = chunk >> 2, kk = (chunk & 3) * 8;
      const int gk = k0 + kk;
      const uint32_t off = ((uint32_t)r * (uint32_t)K + (uint32_t)gk) * 2u;
      const uint32_t offA = ((m0 + r < M) & (gk < K)) ? off : kBufOob;
      const uint32_t offB = ((n0 + r < N) & (gk < K)) ? off : kBufOob;
      pa[0][c] = __builtin_bit_cast(v8, __builtin_amdgcn_raw_buffer_load_b128(rAh, offA, 0, 0));
      pb[0][c] = __builtin_bit_cast(v8, __builtin_amdgcn_raw_buffer_load_b128(rBh, offB, 0, 0));
      if (SPLIT) {
        pa[NOP - 1][c] = __builtin_bit_cast(v8, __builtin_amdgcn_raw_buffer_load_b128(rAl, offA, 0, 0));
        pb[NOP - 1][c] = __builtin_bit_cast(v8, __builtin_amdgcn_raw_buffer_load_b128(rBl, offB, 0, 0));
      }
    }
  };
  auto store_slab = [&](int buf) {
#pragma unroll
    for (int op = 0; op < NOP; op++)
#pragma unroll
      for (int c = 0; c < 1; c++) {
        const int chunk = t + c * kMpThreads;
        const int r = chunk >> 2, kk = (chunk & 3) * 8;
        *(v8*)&As[buf][op][r][kk] = pa[op][c];
        *(v8*)&Bs[buf][op][r][kk] = pb[op][c];
      }
  };
  const int nslab = (K + KT - 1) / KT;
  load_slab(0);
  store_slab(0);
  __syncthreads();
  for (int sl = 0; sl < nslab; sl++) {
    const int buf = sl & 1;
    if (sl + 1 < nslab) load_slab((sl + 1) * KT);
#pragma unroll
    for (int ks = 0; ks < KT; ks += 16) {
      const int ko = ks + 8 * (lane >> 5);
      v8 a0[2], b0[1], a1[2], b1[1];
#pragma unroll
      for (int i = 0; i < 2; i++) {
        a0[i] = *(const v8*)&As[buf][0][wm * 64 + i * 32 + (lane & 31)][ko];
        if (SPLIT) a1[i] = *(const v8*)&As[buf][NOP - 1][wm * 64 + i * 32 + (lane & 31)][ko];
      }
#pragma unroll
      for (int j = 0; j < 1; j++) {
        b0[j] = *(const v8*)&Bs[buf][0][wn * 32 + j * 32 + (lane & 31)][ko];
        if (SPLIT) b1[j] = *(const v8*)&Bs[buf][NOP - 1][wn * 32 + j * 32 + (lane & 31)][ko];
      }
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 1; j++) {
          acc[i][j] = Mma<BF>::op(a0[i], b0[j], acc[i][j]);
          if (SPLIT) {
            accx[i][j] = Mma<BF>::op(a0[i], b1[j], accx[i][j]);
            accx[i][j] = Mma<BF>::op(a1[i], b0[j], accx[i][j]);
          }
        }
    }
    if (sl + 1 < nslab) {
      store_slab(buf ^ 1);
      __syncthreads();
    }
  }

  // this lane's C values (the MFMA output layout), read after the products:
  // keeping them live through the K loop would cost a wavefront per SIMD
  float cpre[2][1][16];
  {
    const __amdgpu_buffer_rsrc_t rc = buf_rsrc(C + (int64_t)m0 * ldc + n0);
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 1; j++)
#pragma unroll
        for (int e = 0; e < 16; e++) {
          const int lr = wm * 64 + i * 32 + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
          const int lc = wn * 32 + j * 32 + (lane & 31);
          const uint32_t off =
              ((m0 + lr < M) & (n0 + lc < N)) ? ((uint32_t)lr * (uint32_t)ldc + (uint32_t)lc) * 4u : kBufOob;
          cpre[i][j][e] = buf_load<float>(rc, off);
        }
  }

  // the scales are powers of two: multiplying by their reciprocals is exact;
  // MFMA tile i's 32 rows are one row block of A
  const int colj = n0 + wn * 32 + (lane & 31);
  const float ibc = colj < N ? iB[colj] : 1.0f;
  float iar[2];
#pragma unroll
  for (int i = 0; i < 2; i++) {
    const int r0 = m0 + wm * 64 + i * 32;
    iar[i] = r0 < M ? iA[r0 >> 5] : 1.0f;
  }
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 1; j++)
#pragma unroll
      for (int e = 0; e < 16; e++) {
        const int row = m0 + wm * 64 + i * 32 + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
        const int col = n0 + wn * 32 + j * 32 + (lane & 31);
        float prod = acc[i][j][e];
        if (SPLIT) prod += accx[i][j][e] * (1.0f / 2048.0f);
        const float out = cpre[i][j][e] - prod * iar[i] * ibc;
        if ((row < M) & (col < N)) C[(int64_t)row * ldc + col] = out;
      }
}

// ---------------------------------------------------------------------------
// Inverses of the 128 x 128 diagonal blocks (unit-lower L or upper U), one
// workgroup per block, for the blocked triangular solves of the refinement.
// ---------------------------------------------------------------------------
constexpr int TB = 128;

template <int UPPER>
__global__ void __launch_bounds__(256) tri_inv_kernel(int n, const float* __restrict__ F, int ldf,
                                                      float* __restrict__ inv) {
  __shared__ float S[TB][TB + 1];
  __shared__ float X[TB][TB + 1];
  const int blk = blockIdx.x, t = threadIdx.x;
  const int k0 = blk * TB, nb = min(TB, n - k0);
  float* out = inv + (int64_t)blk * TB * TB;
  for (int idx = t; idx < TB * TB; idx += 256) {
    int i = idx / TB, c = idx % TB;
    S[i][c] = (i < nb && c < nb) ? F[(int64_t)(k0 + i) * ldf + k0 + c] : (i == c ? 1.f : 0.f);
  }
  __syncthreads();
  if (t < TB) {
    // column c of the inverse: solve T x = e_c by substitution
    const int c = t;
    if (!UPPER) {
      for (int i = 0; i < TB; i++) {
        float s = (i == c) ? 1.f : 0.f;
        for (int q = c; q < i; q++) s -= S[i][q] * X[q][c];
        X[i][c] = (i < c) ? 0.f : s;  // unit diagonal
      }
    } else {
      for (int i = TB - 1; i >= 0; i--) {
        float s = (i == c) ? 1.f : 0.f;
        for (int q = i + 1; q <= c; q++) s -= S[i][q] * X[q][c];
        const float d = S[i][i];
        X[i][c] = (i > c) ? 0.f : s / (d == 0.f ? kPivotFloorF : d);
      }
    }
  }
  __syncthreads();
  for (int idx = t; idx < TB * TB; idx += 256) out[idx] = X[idx / TB][idx % TB];
}

// ---------------------------------------------------------------------------
// Blocked triangular solve with one right-hand side, one workgroup per
// 128-row block, all resident (cooperative launch).  Block i folds in the
// products of its row strip with each finished solution block as soon as that
// block's flag is set, then applies its diagonal inverse and publishes.
// ---------------------------------------------------------------------------
constexpr long kTrsvSpin = 1L << 23;

// The pause between two polls of a block's published words, longer the
// further the poller is from that block in the solve order: only the next
// block's wait is on the chain's critical path, and the many later blocks
// waiting for the same words would otherwise keep their few cache lines
// busy with uncached reads.
__device__ __forceinline__ void poll_pause(int distance) {
  __builtin_amdgcn_s_sleep(1);
  for (int z = 1; z < min(distance, 32); z++) __builtin_amdgcn_s_sleep(4);
}

// A 128 x 128 tile as two register slabs per lane: with 8 wavefronts, row
// r = wave + 8 q (q < 16), columns lane and lane + 64 (coalesced row reads).
// All 32 loads of a lane are in flight at once; rows >= rows or columns
// >= cols read 0.
constexpr int kTrsvThreads = 512, kTrsvWaves = kTrsvThreads / 64, kTrsvRows = TB / kTrsvWaves;
struct TileRegs {
  float a0[kTrsvRows], a1[kTrsvRows];
};

__device__ __forceinline__ void tile_load(TileRegs& T, const float* __restrict__ tile, int64_t ld, int rows, int cols,
                                          int wave, int lane) {
  const __amdgpu_buffer_rsrc_t rt = buf_rsrc(tile);
  // The empty asm makes the lane offset opaque, so the offsets below are
  // recomputed per tile instead of being hoisted out of the caller's loop.
  uint32_t vb = ((uint32_t)wave * (uint32_t)ld + (uint32_t)lane) * 4u;
  asm volatile("" : "+v"(vb));
  const bool c0 = lane < cols, c1 = lane + 64 < cols;
  const uint32_t rstep = 4u * kTrsvWaves * (uint32_t)ld;  // kTrsvWaves rows, in bytes
#pragma unroll
  for (int q = 0; q < kTrsvRows; q++) {
    const bool rv = wave + kTrsvWaves * q < rows;
    const uint32_t off = vb + (uint32_t)q * rstep;
    T.a0[q] = buf_load<float>(rt, (rv & c0) ? off : kBufOob);
    T.a1[q] = buf_load<float>(rt, (rv & c1) ? off + 256u : kBufOob);
  }
}

// s[q] += tile[wave + 8 q][.] . v  (per-lane partial sums)
__device__ __forceinline__ void tile_fma(float s[kTrsvRows], const TileRegs& T, const float* v, int lane) {
  const float v0 = v[lane], v1 = v[lane + 64];
#pragma unroll
  for (int q = 0; q < kTrsvRows; q++) s[q] += T.a0[q] * v0 + T.a1[q] * v1;
}

// Sum N per-lane partials over the wavefront by recursive halving
// (reduce-scatter, N - 1 + log2(64 / N) exchanges instead of 6 N).  Lane l
// ends with the total of partial l / (64 / N).
// a where the lane's bit of m is set, else b: one v_cndmask on a constant
// lane mask (written as a plain select, `hi ? s[w + k] : s[k]` becomes an
// indexed access and the array is moved to LDS)
__device__ __forceinline__ float lane_sel(uint64_t m, float a, float b) {
  float r;
  asm("v_cndmask_b32_e64 %0, %1, %2, %3" : "=v"(r) : "v"(b), "v"(a), "s"(m));
  return r;
}
__host__ __device__ constexpr uint64_t lanes_with_bit(int m) {
  return m == 32 ? 0xFFFFFFFF00000000ull
                 : m == 16 ? 0xFFFF0000FFFF0000ull
                           : m == 8 ? 0xFF00FF00FF00FF00ull
                                    : m == 4 ? 0xF0F0F0F0F0F0F0F0ull : m == 2 ? 0xCCCCCCCCCCCCCCCCull : 0xAAAAAAAAAAAAAAAAull;
}

template <int N>
__device__ __forceinline__ float wave_reduce(float s[N], int lane) {
  int m = 32;
#pragma unroll
  for (int w = N / 2; w >= 1; w >>= 1, m >>= 1) {
    const uint64_t hm = lanes_with_bit(m);
#pragma unroll
    for (int k = 0; k < w; k++) {
      const float keep = lane_sel(hm, s[w + k], s[k]);
      const float send = lane_sel(hm, s[k], s[w + k]);
      s[k] = keep + __shfl_xor(send, m);
    }
  }
  float v = s[0];
#pragma unroll
  for (; m >= 1; m >>= 1) v += __shfl_xor(v, m);
  return v;
}

template <int UPPER>
__global__ void __launch_bounds__(kTrsvThreads) trsv_blocked_kernel(int n, const float* __restrict__ F, int ldf,
                                                           const float* __restrict__ inv, float* __restrict__ x,
                                                           uint64_t* __restrict__ xpub,
                                                           const uint32_t* __restrict__ ebase, uint32_t eoff,
                                                           const int* __restrict__ stop, int* __restrict__ err,
                                                           uint64_t* __restrict__ stamps) {
  // a stopped refinement skips the whole solve (every workgroup returns, so
  // none waits); the epoch of this launch is the device base plus its offset
  if (stop && *stop) return;
  const uint32_t epoch = *ebase + eoff;
  // Blocks are solved in dispatch order (lower: first block first; upper:
  // last block first), and a block only waits for blocks dispatched before
  // it, so the chain needs no co-residency guarantee.  Each solved entry is
  // published as one 64-bit word (value bits, epoch): a reader validates
  // every word it uses on its own, with no fence.  The strip's tiles and the
  // diagonal inverse do not depend on the solution, so they are loaded one
  // step ahead of the wait for the block they multiply.
  __shared__ float xs[2][TB];
  __shared__ float rhs[TB];
  __shared__ int aborted;
  const int nblk = (n + TB - 1) / TB;
  const int bi = UPPER ? nblk - 1 - (int)blockIdx.x : (int)blockIdx.x;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int k0 = bi * TB, nb = min(TB, n - k0);
  if (t == 0) aborted = 0;
  float s[kTrsvRows];
#pragma unroll
  for (int q = 0; q < kTrsvRows; q++) s[q] = 0.f;
  TileRegs Ia, Ib, Iv;
  const int steps = UPPER ? nblk - 1 - bi : bi;
  // the diagonal inverse and this block's right-hand side depend on no other
  // block: loaded first, so the end of the chain waits on neither
  tile_load(Iv, inv + (int64_t)bi * TB * TB, TB, nb, nb, wave, lane);
  constexpr int kSpreadRows = 64 / kTrsvRows;  // lanes holding one row's total
  const int my_row = wave + kTrsvWaves * (lane / kSpreadRows);
  const float my_rhs = (lane % kSpreadRows == 0 && my_row < nb) ? x[k0 + my_row] : 0.f;
  auto col_block = [&](int st) { return UPPER ? nblk - 1 - st : st; };
  auto fetch = [&](TileRegs& T, int st) {
    const int j0 = col_block(st) * TB;
    tile_load(T, F + (int64_t)k0 * ldf + j0, ldf, nb, min(TB, n - j0), wave, lane);
  };
  // wait for block st's solution, stage it in xs[st & 1]
  auto take = [&](int st) {
    const int j0 = col_block(st) * TB, jb = min(TB, n - j0);
    float* dst = xs[st & 1];
    if (t < TB) {
      float v = 0.f;
      if (t < jb) {
        uint64_t wv;
        long spins = 0;
        for (;;) {
          wv = __hip_atomic_load(&xpub[j0 + t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if ((uint32_t)wv == epoch) break;
          if (aborted || ++spins > kTrsvSpin) { aborted = 1; break; }
          poll_pause(steps - st);
        }
        v = __uint_as_float((uint32_t)(wv >> 32));
      }
      dst[t] = v;
    }
    __syncthreads();
  };
  __syncthreads();
  if (steps > 0) fetch(Ia, 0);
  int st = 0;
  for (; st + 1 < steps; st += 2) {
    fetch(Ib, st + 1);
    take(st);
    tile_fma(s, Ia, xs[st & 1], lane);
    if (st + 2 < steps) fetch(Ia, st + 2);
    take(st + 1);
    tile_fma(s, Ib, xs[(st + 1) & 1], lane);
  }
  // rhs = b - L_i,<i x_<i, then x_i = inv_ii rhs, stored plain and
  // published tagged
  auto finish = [&](const TileRegs& Inv) {
    constexpr int kSpread = kSpreadRows;
    const int r = my_row;
    {
      const float v = wave_reduce<kTrsvRows>(s, lane);
      if (lane % kSpread == 0) rhs[r] = (r < nb) ? my_rhs - v : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kTrsvRows; q++) s[q] = 0.f;
    tile_fma(s, Inv, rhs, lane);
    const float v = wave_reduce<kTrsvRows>(s, lane);
    if (lane % kSpread == 0 && r < nb) {
      x[k0 + r] = v;
      __hip_atomic_store(&xpub[k0 + r], ((uint64_t)__float_as_uint(v) << 32) | epoch, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
  };
  if (st < steps) {
    take(st);
    tile_fma(s, Ia, xs[st & 1], lane);
  }
  finish(Iv);
  if (t == 0 && aborted) atomicExch(err, NMX_ERR_SYNC_TIMEOUT);
  if (stamps && t == 0) stamps[bi] = __builtin_amdgcn_s_memrealtime();  // 100 MHz clock
}

// ---------------------------------------------------------------------------
// M_i = inv_i . op_(i, i -+ 1): block i's diagonal inverse times the tile
// that couples it to the block solved just before it (lower: left of the
// diagonal, upper: right of it); zero for the first block solved.  With it
// a block solve x_i = inv_i (t_i - sum_j op_ij x_j) needs, once x of that
// previous block arrives, one product: x_i = pre_i - M_i x_(i -+ 1).
// ---------------------------------------------------------------------------
template <int UPPER>
__global__ void __launch_bounds__(512) couple_kernel(int n, const float* __restrict__ F, int ldf,
                                                     const float* __restrict__ inv, float* __restrict__ out) {
  __shared__ float Iv[TB][TB + 1];
  __shared__ float Tt[TB][TB + 1];
  const int nblk = (n + TB - 1) / TB, bi = blockIdx.x, t = threadIdx.x;
  const int k0 = bi * TB, nb = min(TB, n - k0), bj = UPPER ? bi + 1 : bi - 1;
  float* o = out + (int64_t)bi * TB * TB;
  if (bj < 0 || bj >= nblk) {
    for (int idx = t; idx < TB * TB; idx += 512) o[idx] = 0.f;
    return;
  }
  const int j0 = bj * TB, jb = min(TB, n - j0);
  const float* iv = inv + (int64_t)bi * TB * TB;
  for (int idx = t; idx < TB * TB; idx += 512) {
    const int r = idx / TB, c = idx % TB;
    Iv[r][c] = (r < nb && c < nb) ? iv[idx] : 0.f;
    Tt[r][c] = (r < nb && c < jb) ? F[(int64_t)(k0 + r) * ldf + j0 + c] : 0.f;
  }
  __syncthreads();
  const int r = t >> 2, c0 = (t & 3) * 32;
  float acc[32];
#pragma unroll
  for (int cc = 0; cc < 32; cc++) acc[cc] = 0.f;
  for (int k = 0; k < TB; k++) {
    const float a = Iv[r][k];
#pragma unroll
    for (int cc = 0; cc < 32; cc++) acc[cc] += a * Tt[k][c0 + cc];
  }
#pragma unroll
  for (int cc = 0; cc < 32; cc++) o[r * TB + c0 + cc] = acc[cc];
}

// ---------------------------------------------------------------------------
// The same blocked solve for a block of right-hand sides, Y = op^-1 T: one
// workgroup per (128-row block, group of 16 NS right-hand sides), dispatched
// block-major in solve order, so again a workgroup only waits for
// workgroups dispatched before it; each group is its own chain.  Tiles are
// loaded coalesced and staged through each wavefront's own LDS rows; the
// products with the finished blocks and with the diagonal inverse run on
// v_mfma_f32_16x16x4_f32, NS 16-column accumulators per wavefront sharing
// each fragment.  Each solved entry is published as one 64-bit word (value
// bits, epoch) in pub (ld ldp).
// ---------------------------------------------------------------------------
constexpr int kChainR = 16, kChainThreads = 512;
constexpr int kChainLd = TB + 4;  // LDS row stride of a wavefront's staged tile rows
constexpr int kChainGroup = TB * 16;
typedef float chain_f4 __attribute__((ext_vector_type(4)));

// A 128 x 128 tile in registers, loaded coalesced: instruction j of lane l
// of wavefront w reads row 16 w + j / 2, column 64 (j & 1) + l; rows >= rows
// or columns >= cols read 0.
struct ChainTile {
  float v[32];
};
__device__ __forceinline__ void chain_load(ChainTile& Tl, const float* __restrict__ tile, int64_t ld, int rows,
                                           int cols, int wave, int lane) {
  const __amdgpu_buffer_rsrc_t rt = buf_rsrc(tile);
  uint32_t vb = ((uint32_t)(16 * wave) * (uint32_t)ld + (uint32_t)lane) * 4u;
  asm volatile("" : "+v"(vb));  // recomputed per tile, not hoisted out of the caller's loop
#pragma unroll
  for (int j = 0; j < 32; j++) {
    const int r = 16 * wave + j / 2, c = 64 * (j & 1) + lane;
    const uint32_t off = vb + ((uint32_t)(j / 2) * (uint32_t)ld + 64u * (uint32_t)(j & 1)) * 4u;
    Tl.v[j] = buf_load<float>(rt, (r < rows && c < cols) ? off : kBufOob);
  }
}
// The wavefront's 16 rows into its own LDS rows (stride kChainLd): only this
// wavefront reads them, after its own writes, so no barrier.
__device__ __forceinline__ void chain_stage(const ChainTile& Tl, float* wl, int lane) {
#pragma unroll
  for (int j = 0; j < 32; j++) wl[(j / 2) * kChainLd + 64 * (j & 1) + lane] = Tl.v[j];
}

// acc[g] += (the staged 16 rows) . X_g for the 128 x 16 column groups X_g of a
// staged block, row-major at xs + g * kChainGroup.  MFMA step s: lane l gives
// row l & 15 at K index 4 s + (l >> 4) (bank-conflict-free with the row
// stride kChainLd) and X_g's K index 4 s + (l >> 4), column l & 15.  Lane l,
// item e of acc[g] is row 16 w + 4 (l >> 4) + e, column l & 15 of group g.
__device__ __forceinline__ int chain_slot(int k, int c) { return (c >> 4) * kChainGroup + k * 16 + (c & 15); }
template <int NS>
__device__ __forceinline__ void chain_mma(const float* wl, const float* xs, int lane, chain_f4 (&acc)[NS]) {
  const float* arow = wl + (lane & 15) * kChainLd + (lane >> 4);
#pragma unroll
  for (int s = 0; s < 32; s++) {
    const float a = arow[4 * s];
#pragma unroll
    for (int g = 0; g < NS; g++)
      acc[g] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, xs[g * kChainGroup + 64 * s + lane], acc[g], 0, 0, 0);
  }
}

template <int UPPER, int NS>
__global__ void __launch_bounds__(kChainThreads) trsm_chain_kernel(int n, int nrhs, const float* __restrict__ F,
                                                                   int ldf, const float* __restrict__ inv,
                                                                   const float* __restrict__ cpl,
                                                                   const float* __restrict__ T, int ldt,
                                                                   float* __restrict__ Y, int ldy,
                                                                   uint64_t* __restrict__ pub, int ldp,
                                                                   uint32_t epoch, int* __restrict__ err,
                                                                   uint64_t* __restrict__ stamps) {
  constexpr int CR = 16 * NS;                 // right-hand sides per workgroup
  constexpr int WPT = TB * CR / kChainThreads;  // published words each thread takes per block
  __shared__ float xs[3][NS * kChainGroup];  // two for the waits, one for the block's own right-hand side
  __shared__ float wtile[kChainThreads / 64][16 * kChainLd];
  __shared__ int aborted;
  const int nblk = (n + TB - 1) / TB, groups = (nrhs + CR - 1) / CR;
  const int step = (int)blockIdx.x / groups, grp = (int)blockIdx.x % groups;
  const int bi = UPPER ? nblk - 1 - step : step;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int k0 = bi * TB, nb = min(TB, n - k0), c0 = grp * CR, nc = min(CR, nrhs - c0);
  if (t == 0) aborted = 0;
  if (stamps && t == 0) stamps[2 * blockIdx.x] = __builtin_amdgcn_s_memrealtime();  // 100 MHz clock
  // this block's right-hand sides and diagonal inverse depend on no other
  // block: loaded first
  const int rbase = 16 * wave + 4 * (lane >> 4), cl = lane & 15;
  float tv[NS][4];
#pragma unroll
  for (int g = 0; g < NS; g++)
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int r = rbase + e, c = 16 * g + cl;
      tv[g][e] = (r < nb && c < nc) ? T[(int64_t)(k0 + r) * ldt + c0 + c] : 0.f;
    }
  ChainTile Fa, Fb, Fi, Fm;
  float* wl = wtile[wave];
  chain_load(Fi, inv + (int64_t)bi * TB * TB, TB, nb, nb, wave, lane);
  chain_load(Fm, cpl + (int64_t)bi * TB * TB, TB, nb, TB, wave, lane);
  auto col_block = [&](int st) { return UPPER ? nblk - 1 - st : st; };
  auto fetch = [&](ChainTile& Tl, int st) {
    const int j0 = col_block(st) * TB;
    chain_load(Tl, F + (int64_t)k0 * ldf + j0, ldf, nb, min(TB, n - j0), wave, lane);
  };
  // wait for block st's solution in this group (WPT consecutive words per
  // thread, all polled together), staged in xs[st & 1]
  auto take = [&](int st) {
    const int j0 = col_block(st) * TB, jb = min(TB, n - j0);
    const int k = t / (CR / WPT), cb = (t % (CR / WPT)) * WPT;
    const uint64_t* src = pub + (int64_t)(j0 + k) * ldp + c0 + cb;
    // every word is loaded on every pass, an unwanted one from pub[0], so
    // that all the loads are in flight before the one wait
    bool want[WPT];
    const uint64_t* at[WPT];
#pragma unroll
    for (int i = 0; i < WPT; i++) {
      want[i] = k < jb && cb + i < nc;
      at[i] = want[i] ? src + i : pub;
    }
    uint64_t wv[WPT];
    long spins = 0;
    for (;;) {
#pragma unroll
      for (int i = 0; i < WPT; i++) wv[i] = __hip_atomic_load(at[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      bool ok = true;
#pragma unroll
      for (int i = 0; i < WPT; i++) ok = ok && (!want[i] || (uint32_t)wv[i] == epoch);
      if (ok) break;
      if (aborted || ++spins > kTrsvSpin) {
        aborted = 1;
        break;
      }
      poll_pause(step - st);
    }
    float* dst = xs[st & 1];
#pragma unroll
    for (int i = 0; i < WPT; i++) dst[chain_slot(k, cb + i)] = want[i] ? __uint_as_float((uint32_t)(wv[i] >> 32)) : 0.f;
    __syncthreads();
  };
  __syncthreads();
  // every block solved before this one except the last: its product is
  // folded into the coupling tile Fm = inv . op_(bi, last)
  const int steps = step > 0 ? step - 1 : 0;
  chain_f4 acc[NS];
#pragma unroll
  for (int g = 0; g < NS; g++) acc[g] = chain_f4{0.f, 0.f, 0.f, 0.f};
  if (steps > 0) fetch(Fa, 0);
  int st = 0;
  // each tile is staged before the wait for the block it multiplies
  for (; st + 1 < steps; st += 2) {
    fetch(Fb, st + 1);
    chain_stage(Fa, wl, lane);
    take(st);
    chain_mma<NS>(wl, xs[st & 1], lane, acc);
    if (st + 2 < steps) fetch(Fa, st + 2);
    chain_stage(Fb, wl, lane);
    take(st + 1);
    chain_mma<NS>(wl, xs[(st + 1) & 1], lane, acc);
  }
  if (st < steps) {
    chain_stage(Fa, wl, lane);
    take(st);
    chain_mma<NS>(wl, xs[st & 1], lane, acc);
  }
  // pre = inverse . (T - the products), the right-hand side staged in its
  // own buffer; then, once the last block arrives, x = pre - Fm . x_last
  float* rb = xs[2];
  chain_stage(Fi, wl, lane);
#pragma unroll
  for (int g = 0; g < NS; g++)
#pragma unroll
    for (int e = 0; e < 4; e++) rb[chain_slot(rbase + e, 16 * g + cl)] = tv[g][e] - acc[g][e];
  __syncthreads();
  chain_f4 x[NS];
#pragma unroll
  for (int g = 0; g < NS; g++) x[g] = chain_f4{0.f, 0.f, 0.f, 0.f};
  chain_mma<NS>(wl, rb, lane, x);
  if (step > 0) {
    chain_stage(Fm, wl, lane);
    take(steps);  // the block solved just before this one
    chain_f4 cm[NS];
#pragma unroll
    for (int g = 0; g < NS; g++) cm[g] = chain_f4{0.f, 0.f, 0.f, 0.f};
    chain_mma<NS>(wl, xs[steps & 1], lane, cm);
#pragma unroll
    for (int g = 0; g < NS; g++) x[g] -= cm[g];
  }
#pragma unroll
  for (int g = 0; g < NS; g++)
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int r = rbase + e, c = 16 * g + cl;
      if (r < nb && c < nc) {
        Y[(int64_t)(k0 + r) * ldy + c0 + c] = x[g][e];
        __hip_atomic_store(&pub[(int64_t)(k0 + r) * ldp + c0 + c], ((uint64_t)__float_as_uint(x[g][e]) << 32) | epoch,
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  if (t == 0 && aborted) atomicExch(err, NMX_ERR_SYNC_TIMEOUT);
  if (stamps && t == 0) stamps[2 * blockIdx.x + 1] = __builtin_amdgcn_s_memrealtime();
}

// ---------------------------------------------------------------------------
// r = b - A x (b == nullptr: r = -A x) in FP64, one wavefront per row, the
// rows strided over at most kResidualBlocks workgroups; with `norms`, the
// max |r|, max |x| and (with rowabs) max row sum |A| of the stopping test,
// one atomic per workgroup each.
constexpr int kResidualBlocks = 1024;
constexpr int kResidualChunks = 16;
__global__ void __launch_bounds__(256) residual_kernel(int n, const double* __restrict__ A, int lda,
                                                       const double* __restrict__ x, const double* __restrict__ b,
                                                       double* __restrict__ r, double* __restrict__ rowabs,
                                                       unsigned long long* __restrict__ norms,
                                                       const int* __restrict__ stop) {
  if (stop && *stop) return;
  __shared__ unsigned long long red[3][4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const __amdgpu_buffer_rsrc_t rx = buf_rsrc(x);
  // running maxima as bit patterns of non-negative doubles (a NaN stays the max)
  unsigned long long mr = 0, mx = 0, ma = 0;
  for (int row = blockIdx.x * 4 + wv; row < n; row += gridDim.x * 4) {
    const __amdgpu_buffer_rsrc_t ra = buf_rsrc(A + (int64_t)row * lda);
    double s = 0.0, sa = 0.0;
    // kResidualChunks column chunks of 64 in flight per step (8 KB of the
    // row per wavefront); past-the-end reads are 0
    for (int c0 = 0; c0 < n; c0 += 64 * kResidualChunks) {
      double av[kResidualChunks], xv[kResidualChunks];
#pragma unroll
      for (int u = 0; u < kResidualChunks; u++) {
        const int c = c0 + u * 64 + lane;
        const uint32_t off = c < n ? (uint32_t)c * 8u : kBufOob;
        av[u] = buf_load<double>(ra, off);
        xv[u] = buf_load<double>(rx, off);
      }
#pragma unroll
      for (int u = 0; u < kResidualChunks; u++) {
        s += av[u] * xv[u];
        sa += fabs(av[u]);
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      s += __shfl_xor(s, off);
      sa += __shfl_xor(sa, off);
    }
    if (lane == 0) {
      const double rv = (b ? b[row] : 0.0) - s;
      r[row] = rv;
      if (rowabs) rowabs[row] = sa;
      mr = max(mr, (unsigned long long)__double_as_longlong(fabs(rv)));
      mx = max(mx, (unsigned long long)__double_as_longlong(fabs(x[row])));
      ma = max(ma, (unsigned long long)__double_as_longlong(sa));
    }
  }
  if (!norms) return;  // uniform: the whole grid takes the same branch
  if (lane == 0) {
    red[0][wv] = mr;
    red[1][wv] = mx;
    red[2][wv] = ma;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int q = threadIdx.x;
    const unsigned long long v = max(max(red[q][0], red[q][1]), max(red[q][2], red[q][3]));
    if (q < 2 || rowabs) atomicMax(&norms[q], v);
  }
}

__global__ void zero_u64(unsigned long long* p, int n) {
  int i = threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// F = (float) A, a wavefront a row, eight 512-byte loads in flight a lane;
// with `anorm`, also ||A||_inf (the largest row sum of |a_ij|, each row
// summed lane-strided then across the wavefront, as the blocked residual
// does) into *anorm as the bits of a non-negative double.
constexpr int kCastBlocks = 1024;
__global__ void __launch_bounds__(256) cast_f64_f32(int n, const double* __restrict__ A, int lda,
                                                    float* __restrict__ F, int ldf,
                                                    unsigned long long* __restrict__ anorm) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  unsigned long long m = 0;
  for (int row = blockIdx.x * 4 + wv; row < n; row += gridDim.x * 4) {
    const double* a = A + (int64_t)row * lda;
    float* f = F + (int64_t)row * ldf;
    double s = 0.0;
    int c = lane;
    for (; c + 64 * 7 < n; c += 64 * 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) v[u] = a[c + 64 * u];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        f[c + 64 * u] = (float)v[u];
        s += fabs(v[u]);
      }
    }
    for (; c < n; c += 64) {
      const double v = a[c];
      f[c] = (float)v;
      s += fabs(v);
    }
    if (anorm) {  // uniform
      for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
      m = max(m, (unsigned long long)__double_as_longlong(s));
    }
  }
  if (anorm && lane == 0) atomicMax(anorm, m);
}

// Device-side state of one classical refinement: its steps run as a fixed
// launch sequence (a graph can hold it), each kernel returning at once when
// `stop` is set; ir_decide_kernel makes the host loop's decision per step.
struct IrDev {
  int32_t stop, it, converged, pad;
  double eta, prev, anorm;
  double history[32];
  uint32_t ebase, pad2;  // TRSV epoch base: each solve moves it by kEpochSpan
};
constexpr uint32_t kEpochSpan = 1u << 20;  // TRSV launches one solve may make

__global__ void ir_init_kernel(IrDev* s) {
  if (threadIdx.x == 0) {
    *s = IrDev{};
    s->ebase = 0;
  }
}
__global__ void ir_epoch_bump_kernel(IrDev* s) {
  if (threadIdx.x == 0) s->ebase += kEpochSpan;
}
__global__ void ir_reset_kernel(IrDev* s) {
  if (threadIdx.x == 0) {
    s->stop = 0;
    s->it = 0;
    s->converged = 0;
    s->eta = __builtin_inf();
    s->prev = __builtin_inf();
    s->anorm = 0.0;
    for (int i = 0; i < 32; i++) s->history[i] = 0.0;
    s->ebase += kEpochSpan;
  }
}
// After a residual: eta = max|r| / (||A||_inf max|x|) (first: ||A||_inf from
// the residual's row sums), recorded; stop when converged, out of steps or
// no longer halving; otherwise count the correction about to be applied.
__global__ void ir_decide_kernel(IrDev* s, const unsigned long long* __restrict__ norms, double tol, int max_iter,
                                 int first) {
  if (threadIdx.x != 0 || s->stop) return;
  const double rm = __longlong_as_double((long long)norms[0]), xm = __longlong_as_double((long long)norms[1]);
  if (first) s->anorm = __longlong_as_double((long long)norms[2]);
  const double eta = rm / (s->anorm * xm);
  const int it = s->it;
  s->eta = eta;
  s->history[it < 32 ? it : 31] = eta;
  if (!(eta > tol) || it >= max_iter || (it > 0 && !(eta < 0.5 * s->prev))) {
    s->stop = 1;
    s->converged = eta <= tol ? 1 : 0;
  } else {
    s->prev = eta;
    s->it = it + 1;
  }
}

__global__ void vec_f32_add_f64(int n, const float* __restrict__ y, double* __restrict__ x, int accumulate,
                                const int* __restrict__ stop) {
  if (stop && *stop) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = accumulate ? x[i] + (double)y[i] : (double)y[i];
}

}  // namespace

// ===========================================================================
// Host side
// ===========================================================================
struct MpFactor {
  int n = 0, block = 256, prec = 2;  // prec: 0 fp16, 1 bf16, 2 fp16 split, 3 fp32
  int pivot = 1;                     // 1: partial pivoting (GEPP); 2: tournament pivoting; 0: butterflies, no pivoting
  int32_t* ipiv = nullptr;
  int32_t* perm = nullptr;           // net row permutation of ipiv
  float* F = nullptr;                // n x n FP32 factors (of W^T A V without pivoting)
  float* dw = nullptr;               // 2n butterfly diagonals for W, then V
  float* dv = nullptr;
  float* invL = nullptr;             // nblk x 128 x 128
  float* invU = nullptr;
  int32_t* info = nullptr;
  void* lo_buf = nullptr;            // FP16/BF16 operand copies
  float* rs = nullptr;               // scales of L21's 32-row blocks, then their reciprocals (at + n), per step
  float* cs = nullptr;               // per-column scales of U12 and their reciprocals (2n), per step
  uint64_t* xpub = nullptr;          // TRSV published solution words (n)
  IrDev* ird = nullptr;              // device-side refinement state and TRSV epoch base
  uint32_t eoff = 0;                 // TRSV epoch offset of the next launch in this solve
  uint64_t* mpub = nullptr;          // block-solve published words (n x mpub_ld), when set up
  int mpub_ld = 0;
  float* cplL = nullptr;             // block-solve coupling tiles (couple_kernel), when set up
  float* cplU = nullptr;
  void* ts = nullptr;                // tournament leaf scratch
  size_t ts_bytes = 0;
  uint32_t mepoch = 0;
  bool ok = false;
};

// Device buffers of the factorization, cached in the context and grown on
// demand so repeated solves do not allocate.
struct MpBuffers {
  size_t F_elems = 0, d_elems = 0, inv_elems = 0, piv_elems = 0, lo_bytes = 0, pub_elems = 0, ts_bytes = 0;
  void* ts = nullptr;
  float* F = nullptr;
  float* d = nullptr;
  float* inv = nullptr;
  int32_t* ipiv = nullptr;
  int32_t* info = nullptr;
  void* lo = nullptr;
  float* sc = nullptr;
  size_t sc_elems = 0;
  uint64_t* xpub = nullptr;
  IrDev* ird = nullptr;
  size_t ird_elems = 0;
  uint64_t* mpub = nullptr;
  size_t mpub_elems = 0;
  float* cpl = nullptr;
  size_t cpl_elems = 0;
  uint32_t mepoch = 0;
};

static void mp_buffers_free(void* p) {
  MpBuffers* b = (MpBuffers*)p;
  if (!b) return;
  hipFree(b->F);
  hipFree(b->d);
  hipFree(b->inv);
  hipFree(b->ipiv);
  hipFree(b->info);
  hipFree(b->lo);
  hipFree(b->sc);
  hipFree(b->xpub);
  hipFree(b->ird);
  hipFree(b->mpub);
  hipFree(b->cpl);
  hipFree(b->ts);
  delete b;
}

// Grows a cached buffer.  Never under capture (run the call once first);
// while captured graphs live, the old buffer is retired, not freed (they may
// still read it), and released with the last graph.
template <typename P>
static int grow(nmx_ctx* ctx, P*& ptr, size_t& have, size_t want, size_t elem) {
  if (want <= have && ptr) return NMX_OK;
  if (ctx->capturing) return NMX_ERR_UNSUPPORTED;
  if (ptr) {
    if (ctx->live_graphs > 0)
      ctx->retired.push_back((void*)ptr);
    else
      hipFree(ptr);
  }
  ptr = nullptr;
  have = 0;
  if (hipMalloc((void**)&ptr, want * elem + 256) != hipSuccess) return NMX_ERR_ALLOC;
  have = want;
  return NMX_OK;
}

static int mp_alloc(nmx_ctx* ctx, MpFactor& f, int n, int block, int prec) {
  f.n = n;
  f.block = block;
  f.prec = prec;
  MpBuffers* b = (MpBuffers*)ctx->mp_cache;
  if (!b) {
    b = new MpBuffers();
    ctx->mp_cache = b;
    ctx->mp_cache_free = mp_buffers_free;
  }
  if (!ctx->capturing) NMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const int nblk = ceil_div(n, TB);
  size_t lo_bytes = (size_t)n * block * 4 * 2;  // A_hi, A_lo, B_hi, B_lo (lo unused unless split)
  size_t unused = 0;
  NMX_TRY(grow(ctx, b->F, b->F_elems, (size_t)n * n, sizeof(float)));
  NMX_TRY(grow(ctx, b->d, b->d_elems, 4 * (size_t)n + 16, sizeof(float)));
  NMX_TRY(grow(ctx, b->inv, b->inv_elems, (size_t)nblk * TB * TB * 2, sizeof(float)));
  NMX_TRY(grow(ctx, b->ipiv, b->piv_elems, 2 * (size_t)n, sizeof(int32_t)));
  NMX_TRY(grow(ctx, b->lo, b->lo_bytes, lo_bytes, 1));
  if (!b->info) NMX_TRY(grow(ctx, b->info, unused, 64, sizeof(int32_t)));
  NMX_TRY(grow(ctx, b->sc, b->sc_elems, 4 * (size_t)n, sizeof(float)));
  if ((size_t)n > b->pub_elems) {
    NMX_TRY(grow(ctx, b->xpub, b->pub_elems, (size_t)n, sizeof(uint64_t)));
    hipMemsetAsync(b->xpub, 0, sizeof(uint64_t) * (size_t)n, ctx->stream);
    NMX_HIP(ctx, hipGetLastError());
  }
  if (!b->ird) {
    NMX_TRY(grow(ctx, b->ird, b->ird_elems, 1, sizeof(IrDev)));
    ir_init_kernel<<<1, 64, 0, ctx->stream>>>(b->ird);
    NMX_HIP(ctx, hipGetLastError());
  }
  f.ird = b->ird;
  f.eoff = 0;
  f.F = b->F;
  f.dw = b->d;
  f.dv = b->d + 2 * (size_t)n;
  f.invL = b->inv;
  f.invU = b->inv + (size_t)nblk * TB * TB;
  f.ipiv = b->ipiv;
  f.perm = b->ipiv + n;
  f.info = b->info;
  f.lo_buf = b->lo;
  f.rs = b->sc;                  // [0, n) scales, [n, 2n) reciprocals (L21: one per 32-row block)
  f.cs = b->sc + 2 * (size_t)n;
  f.xpub = b->xpub;
  if (f.pivot == 2) {
    const size_t want = tslu_scratch_bytes(n);
    if (want > b->ts_bytes) {
      NMX_TRY(grow(ctx, b->ts, b->ts_bytes, want, 1));
      NMX_TRY(tslu_scratch_init(ctx, b->ts));
    }
    f.ts = b->ts;
    f.ts_bytes = b->ts_bytes;
  }
  return NMX_OK;
}

// The published words of the block solves for nrhs right-hand sides (after
// mp_alloc): n x round_up(nrhs, kChainR), cleared when they grow.
static int mp_alloc_chains(nmx_ctx* ctx, MpFactor& f, int nrhs) {
  MpBuffers* b = (MpBuffers*)ctx->mp_cache;
  const int ld = ceil_div(nrhs, kChainR) * kChainR;
  const size_t want = (size_t)f.n * ld;
  if (want > b->mpub_elems) {
    NMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    NMX_TRY(grow(ctx, b->mpub, b->mpub_elems, want, sizeof(uint64_t)));
    NMX_HIP(ctx, hipMemsetAsync(b->mpub, 0, sizeof(uint64_t) * want, ctx->stream));
    b->mepoch = 0;
  }
  f.mpub = b->mpub;
  f.mpub_ld = ld;
  f.mepoch = b->mepoch;
  const size_t tiles = (size_t)ceil_div(f.n, TB) * TB * TB;
  NMX_TRY(grow(ctx, b->cpl, b->cpl_elems, 2 * tiles, sizeof(float)));
  f.cplL = b->cpl;
  f.cplU = b->cpl + tiles;
  return NMX_OK;
}

// The coupling tiles of both triangles, after the factorization.
static int mp_couple(nmx_ctx* ctx, MpFactor& f) {
  const int nblk = ceil_div(f.n, TB);
  couple_kernel<0><<<nblk, 512, 0, ctx->stream>>>(f.n, f.F, f.n, f.invL, f.cplL);
  couple_kernel<1><<<nblk, 512, 0, ctx->stream>>>(f.n, f.F, f.n, f.invU, f.cplU);
  NMX_HIP(ctx, hipGetLastError());
  return NMX_OK;
}

// Hands the block-solve epoch back to the cache (published words keep their
// last values).
static void mp_free(nmx_ctx* ctx, MpFactor& f) {
  MpBuffers* b = (MpBuffers*)ctx->mp_cache;
  if (b && f.mpub) b->mepoch = f.mepoch;
  f = MpFactor();
}

int zero_info(nmx_ctx* ctx, int32_t* info);

// The operands of step k's update A22 -= L21 U12 (L21 m x nb, U12 nb x m)
// scaled and converted to FP16 / BF16: L21 as is, U12 transposed (row j of
// the copy is trailing column j).
template <int BF, int SPLIT>
static int mp_convert(nmx_ctx* ctx, MpFactor& f, int k, int nb, int m) {
  typedef typename Lo<BF>::t lt;
  const int n = f.n;
  const float* L21 = f.F + (int64_t)(k + nb) * n + k;  // m x nb, ld n
  const float* U12 = f.F + (int64_t)k * n + k + nb;    // nb x m, ld n
  lt* base = (lt*)f.lo_buf;
  const size_t op = (size_t)m * nb;
  const int g1 = (int)std::min<int64_t>(4096, ((int64_t)m * nb + 255) / 256);
  // both operands' scales in one launch, then both conversions in one
  const int ga = ceil_div(m, 32), gb = ceil_div(m, 64);
  scales2_kernel<<<ga + gb, 256, 0, ctx->stream>>>(ga, m, nb, L21, n, f.rs, U12, n, f.cs, n);
  convert2_kernel<BF, SPLIT><<<2 * g1, 256, 0, ctx->stream>>>(g1, m, nb, L21, n, f.rs, base, base + op, U12, n,
                                                               f.cs, base + 2 * op, base + 3 * op);
  NMX_HIP(ctx, hipGetLastError());
  return NMX_OK;
}

// Trailing columns j0 .. j1 of step k's update, from mp_convert's copies.
template <int BF, int SPLIT>
static int mp_gemm_cols(nmx_ctx* ctx, MpFactor& f, int k, int nb, int m, int j0, int j1) {
  typedef typename Lo<BF>::t lt;
  const int n = f.n, cols = j1 - j0;
  if (cols <= 0) return NMX_OK;
  lt* base = (lt*)f.lo_buf;
  const size_t op = (size_t)m * nb;
  const int tm = ceil_div(m, MT), tn = ceil_div(cols, MT);
  const int grid = ceil_div(tm * tn, kXcds) * kXcds;
  ProfScope ps(ctx, K_MP_GEMM, 2.0 * m * (double)cols * nb * (SPLIT ? 3 : 1),
               8.0 * (double)m * cols + (double)sizeof(lt) * (double)(m + cols) * nb * (SPLIT ? 2 : 1));
  mp_gemm_kernel<BF, SPLIT><<<grid, kMpThreads, 0, ctx->stream>>>(
      m, cols, nb, base, base + op, base + 2 * op + (size_t)j0 * nb, base + 3 * op + (size_t)j0 * nb, f.rs + n,
      f.cs + n + j0, f.F + (int64_t)(k + nb) * n + k + nb + j0, n, tm, tn);
  NMX_HIP(ctx, hipGetLastError());
  return NMX_OK;
}

template <int BF, int SPLIT>
static int mp_update(nmx_ctx* ctx, MpFactor& f, int k, int nb, int m) {
  NMX_TRY((mp_convert<BF, SPLIT>(ctx, f, k, nb, m)));
  return mp_gemm_cols<BF, SPLIT>(ctx, f, k, nb, m, 0, m);
}

// The tournament-pivoted factorization with one panel of look-ahead: after
// panel k, its U12 and the update of the next panel's columns run on the
// main stream, the rest of the update -- most of the MFMA work -- on the
// context's second stream, while the main stream factors panel k+1, a
// latency-bound run of single-wavefront tournaments that leaves most of the
// chip to the update.  Panel k+1's interchanges reach the columns right of
// it only once that update has finished (the conversion buffers and the
// per-row / per-column scales are reused only after that wait too).
template <int BF, int SPLIT>
static int mp_getrf_lookahead(nmx_ctx* ctx, MpFactor& f) {
  const int n = f.n, bw = f.block;
  float* F = f.F;
  NMX_TRY(nmx_aux_stream(ctx));
  NMX_TRY(zero_info(ctx, f.info));
  bool behind = false;  // an update is queued on the second stream
  for (int k = 0; k < n; k += bw) {
    const int nb = std::min(bw, n - k), slot = (k / bw) & 1;
    NMX_TRY(tslu_panel_launch(ctx, f.ts, f.ts_bytes, n, F, n, k, nb, f.ipiv, f.info, 0, slot));
    if (behind) {
      NMX_HIP(ctx, hipStreamWaitEvent(ctx->stream, ctx->behind_ev, 0));
      behind = false;
    }
    NMX_TRY(tslu_swap_columns(ctx, f.ts, F, n, k, nb, k + nb, n, slot));
    const int m = n - k - nb;
    if (m <= 0) {
      NMX_TRY(tslu_swap_columns(ctx, f.ts, F, n, k, nb, 0, k, slot));
      break;
    }
    NMX_TRY(trsm_left<float>(ctx, 0, 1, 0, nb, m, F + (int64_t)k * n + k, n, F + (int64_t)k * n + k + nb, n));
    NMX_TRY((mp_convert<BF, SPLIT>(ctx, f, k, nb, m)));
    const int near = std::min(bw, m);
    NMX_TRY((mp_gemm_cols<BF, SPLIT>(ctx, f, k, nb, m, 0, near)));
    if (m > near) {
      NMX_HIP(ctx, hipEventRecord(ctx->ahead_ev, ctx->stream));
      NMX_HIP(ctx, hipStreamWaitEvent(ctx->aux, ctx->ahead_ev, 0));
      {
        // the interchanges of the columns left of the panel feed nothing
        // until the end: they trail on the second stream too, from this
        // panel's record slot, which the panel after next refills only once
        // the main stream has waited for this work
        StreamSwap on_aux(ctx, ctx->aux);
        NMX_TRY((mp_gemm_cols<BF, SPLIT>(ctx, f, k, nb, m, near, m)));
        NMX_TRY(tslu_swap_columns(ctx, f.ts, F, n, k, nb, 0, k, slot));
      }
      NMX_HIP(ctx, hipEventRecord(ctx->behind_ev, ctx->aux));
      behind = true;
    } else {
      NMX_TRY(tslu_swap_columns(ctx, f.ts, F, n, k, nb, 0, k, slot));
    }
  }
  if (behind) NMX_HIP(ctx, hipStreamWaitEvent(ctx->stream, ctx->behind_ev, 0));
  return NMX_OK;
}

// diagonal block without pivoting: recursive down to 64 columns
static int nopiv_diag(nmx_ctx* ctx, MpFactor& f, int k, int nb) {
  const int n = f.n;
  float* F = f.F;
  if (nb <= 64) {
    ProfScope ps(ctx, K_MP_PANEL, 2.0 / 3.0 * nb * nb * (double)nb, 8.0 * nb * nb);
    getf2_nopiv_kernel<<<1, 256, 0, ctx->stream>>>(F, n, k, nb, f.info);
    NMX_HIP(ctx, hipGetLastError());
    return NMX_OK;
  }
  int half = ((nb / 2 + 63) / 64) * 64;
  if (half >= nb) half = nb - 64;
  const int rest = nb - half;
  NMX_TRY(nopiv_diag(ctx, f, k, half));
  float* A11 = F + (int64_t)k * n + k;
  NMX_TRY(trsm_left<float>(ctx, 0, 1, 0, half, rest, A11, n, F + (int64_t)k * n + k + half, n));
  // L21 := A21 U11^-1 (rows solve U11^T x^T = a^T)
  NMX_TRY(trsm_rec_rows_upper(ctx, rest, half, A11, n, F + (int64_t)(k + half) * n + k, n));
  NMX_TRY(gemm_launch<float>(ctx, 0, 0, rest, rest, half, -1.0f, F + (int64_t)(k + half) * n + k, n,
                             F + (int64_t)k * n + k + half, n, 1.0f, F + (int64_t)(k + half) * n + k + half, n, 0));
  return nopiv_diag(ctx, f, k + half, rest);
}

static int mp_update_hook(nmx_ctx* ctx, void* user, int k, int nb, int m) {
  MpFactor& f = *(MpFactor*)user;
  switch (f.prec) {
    case 0: return mp_update<0, 0>(ctx, f, k, nb, m);
    case 1: return mp_update<1, 0>(ctx, f, k, nb, m);
    default: return mp_update<0, 1>(ctx, f, k, nb, m);
  }
}

static int tslu_hook(nmx_ctx* ctx, void* user, int n, float* A, int lda, int k, int nb, int32_t* ipiv,
                     int32_t* info) {
  MpFactor& f = *(MpFactor*)user;
  return tslu_panel_launch(ctx, f.ts, f.ts_bytes, n, A, lda, k, nb, ipiv, info, 1);
}

// anorm (pivoting paths only): ||A||_inf into *anorm (zeroed by the caller)
// from the same pass that casts A.
static int mp_factor(nmx_ctx* ctx, MpFactor& f, const double* A, int lda, uint64_t seed, int stop_after_transform = 0,
                     unsigned long long* anorm = nullptr) {
  const int n = f.n;
  if (f.pivot) {
    // GEPP in FP32 (the cooperative pivoting panel), the block-level trailing
    // update on FP16 / BF16 MFMA
    const int64_t total = (int64_t)n * n;
    {
      ProfScope ps(ctx, K_OTHER, 0.0, 12.0 * (double)total);
      cast_f64_f32<<<std::min(ceil_div(n, 4), kCastBlocks), 256, 0, ctx->stream>>>(n, A, lda, f.F, n, anorm);
      NMX_HIP(ctx, hipGetLastError());
    }
    if (stop_after_transform) return zero_info(ctx, f.info);
    if (f.pivot == 2 && f.prec != 3) {
      switch (f.prec) {
        case 0: NMX_TRY((mp_getrf_lookahead<0, 0>(ctx, f))); break;
        case 1: NMX_TRY((mp_getrf_lookahead<1, 0>(ctx, f))); break;
        default: NMX_TRY((mp_getrf_lookahead<0, 1>(ctx, f))); break;
      }
    } else {
      NMX_TRY(getrf_impl_ex<float>(ctx, n, f.F, n, f.ipiv, f.info, f.block, kLeafW,
                                   f.prec == 3 ? nullptr : mp_update_hook, &f, f.pivot == 2 ? tslu_hook : nullptr));
    }
    NMX_TRY(perm_from_ipiv(ctx, n, f.ipiv, f.perm));
    const int nblk = ceil_div(n, TB);
    tri_inv_kernel<0><<<nblk, 256, 0, ctx->stream>>>(n, f.F, n, f.invL);
    tri_inv_kernel<1><<<nblk, 256, 0, ctx->stream>>>(n, f.F, n, f.invU);
    NMX_HIP(ctx, hipGetLastError());
    f.ok = true;
    return NMX_OK;
  }
  // butterflies and the transformed FP32 matrix
  rbt_diag_kernel<<<ceil_div(4 * n, 256), 256, 0, ctx->stream>>>(4 * n, seed, f.dw);
  const int n4 = n & ~3, h = n4 / 4;
  const int64_t work = (int64_t)h * h + (int64_t)(n - n4) * n + (int64_t)n4 * (n - n4);
  {
    ProfScope ps(ctx, K_OTHER, 0.0, 12.0 * (double)n * n);
    rbt_matrix_kernel<<<(unsigned)((work + 255) / 256), 256, 0, ctx->stream>>>(n, A, lda, f.F, n, f.dw, f.dv);
    NMX_HIP(ctx, hipGetLastError());
  }
  NMX_TRY(zero_info(ctx, f.info));
  if (stop_after_transform) return NMX_OK;
  for (int k = 0; k < n; k += f.block) {
    const int nb = std::min(f.block, n - k);
    NMX_TRY(nopiv_diag(ctx, f, k, nb));
    const int m = n - k - nb;
    if (m <= 0) break;
    float* A11 = f.F + (int64_t)k * n + k;
    NMX_TRY(trsm_left<float>(ctx, 0, 1, 0, nb, m, A11, n, f.F + (int64_t)k * n + k + nb, n));
    NMX_TRY(trsm_rec_rows_upper(ctx, m, nb, A11, n, f.F + (int64_t)(k + nb) * n + k, n));
    switch (f.prec) {
      case 0: NMX_TRY((mp_update<0, 0>(ctx, f, k, nb, m))); break;
      case 1: NMX_TRY((mp_update<1, 0>(ctx, f, k, nb, m))); break;
      case 2: NMX_TRY((mp_update<0, 1>(ctx, f, k, nb, m))); break;
      default:
        NMX_TRY(gemm_launch<float>(ctx, 0, 0, m, m, nb, -1.0f, f.F + (int64_t)(k + nb) * n + k, n,
                                   f.F + (int64_t)k * n + k + nb, n, 1.0f, f.F + (int64_t)(k + nb) * n + k + nb, n, 0));
    }
  }
  const int nblk = ceil_div(n, TB);
  tri_inv_kernel<0><<<nblk, 256, 0, ctx->stream>>>(n, f.F, n, f.invL);
  tri_inv_kernel<1><<<nblk, 256, 0, ctx->stream>>>(n, f.F, n, f.invU);
  NMX_HIP(ctx, hipGetLastError());
  f.ok = true;
  return NMX_OK;
}

// ---- multi-RHS refinement building blocks -------------------------------------
// R = B - (P_0 + ... + P_{kResSplit-1}) over an n x nrhs block, the split-K
// partial products of A X summed in index order
constexpr int kResSplit = 4;
__global__ void residual_combine_kernel(int n, int nrhs, const double* __restrict__ B, int ldb,
                                        const double* __restrict__ P, double* __restrict__ R) {
  const int64_t total = (int64_t)n * nrhs, idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const int r = (int)(idx / nrhs), c = (int)(idx % nrhs);
  double s = P[idx];
#pragma unroll
  for (int q = 1; q < kResSplit; q++) s += P[q * total + idx];
  R[idx] = B[(int64_t)r * ldb + c] - s;
}

// x (+)= y over an n x nrhs block (row-major, leading dimensions ldy / ldx)
__global__ void mat_f32_add_f64(int n, int nrhs, const float* __restrict__ y, int ldy, double* __restrict__ x,
                                int ldx, int accumulate) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * nrhs) return;
  const int r = (int)(idx / nrhs), c = (int)(idx % nrhs);
  const double v = (double)y[(int64_t)r * ldy + c];
  double* px = x + (int64_t)r * ldx + c;
  *px = accumulate ? *px + v : v;
}

// per-column max |M[:, c]| of an n x nrhs block as bit patterns (a NaN
// stays the max): rows strided over the workgroups, one atomic per column
// per workgroup
constexpr int kColMaxBlocks = 256;
__global__ void __launch_bounds__(256) colmax_kernel(int n, int nrhs, const double* __restrict__ M, int ldm,
                                                     unsigned long long* __restrict__ out) {
  for (int c0 = 0; c0 < nrhs; c0 += 256) {
    const int c = c0 + (int)threadIdx.x;
    if (c < nrhs) {
      unsigned long long m = 0;
      for (int r = blockIdx.x; r < n; r += gridDim.x)
        m = max(m, (unsigned long long)__double_as_longlong(fabs(M[(int64_t)r * ldm + c])));
      atomicMax(&out[c], m);
    }
  }
}

// max over rows of sum_j |A[i][j]| (the infinity norm), one wavefront per row
__global__ void __launch_bounds__(256) row_abs_max_kernel(int n, const double* __restrict__ A, int lda,
                                                          unsigned long long* __restrict__ out) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  unsigned long long m = 0;
  for (int row = blockIdx.x * 4 + wv; row < n; row += gridDim.x * 4) {
    const double* a = A + (int64_t)row * lda;
    double s = 0.0;
    for (int c = lane; c < n; c += 64) s += fabs(a[c]);
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    m = max(m, (unsigned long long)__double_as_longlong(s));
  }
  if (lane == 0) atomicMax(out, m);
}

// zero the columns whose flag is set (converged right-hand sides get no correction)
__global__ void zero_cols_kernel(int n, int nrhs, double* __restrict__ M, int ldm, const int32_t* __restrict__ flag) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * nrhs) return;
  const int r = (int)(idx / nrhs), c = (int)(idx % nrhs);
  if (flag[c]) M[(int64_t)r * ldm + c] = 0.0;
}

// ---- GMRES-IR building blocks (FP64) ------------------------------------------
// out[i] = V_i . w for the k basis vectors V_i (rows of V, length n, ld n).
constexpr int kGmDotThreads = 256;
__global__ void __launch_bounds__(kGmDotThreads) gm_dots_kernel(int n, const double* __restrict__ V,
                                                                const double* __restrict__ w, double* __restrict__ out) {
  __shared__ double red[kGmDotThreads / 64];
  const double* v = V + (int64_t)blockIdx.x * n;
  double s = 0.0;
  for (int e = threadIdx.x; e < n; e += kGmDotThreads) s += v[e] * w[e];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int q = 0; q < kGmDotThreads / 64; q++) t += red[q];
    out[blockIdx.x] = t;
  }
}

// w -= sum_{i<k} h_i V_i  (h on the device)
__global__ void gm_update_kernel(int n, int k, const double* __restrict__ V, const double* __restrict__ h,
                                 double* __restrict__ w) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double s = w[e];
  for (int i = 0; i < k; i++) s -= h[i] * V[(int64_t)i * n + e];
  w[e] = s;
}

// v = alpha w
__global__ void gm_scale_kernel(int n, double alpha, const double* __restrict__ w, double* __restrict__ v) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) v[e] = alpha * w[e];
}

// d (FP64, n) = V U^-1 L^-1 W^T r  (accumulate: x += d)
// x (+)= M^-1 r; with `stop` (device), every kernel returns at once once it
// is set.
static int mp_solve_vec(nmx_ctx* ctx, MpFactor& f, const double* r, double* x, float* tmp, int accumulate,
                        const int* stop = nullptr) {
  const int n = f.n, n4 = n & ~3, h = n4 / 4;
  const int64_t work = (int64_t)h + (n - n4);
  if (f.pivot) {
    NMX_TRY((gather_rows<double, float>(ctx, n, 1, f.perm, r, 1, tmp, 1)));
  } else {
    rbt_vec_t_kernel<<<(unsigned)((work + 255) / 256), 256, 0, ctx->stream>>>(n, 1, r, 1, tmp, 1, f.dw);
    NMX_HIP(ctx, hipGetLastError());
  }
  const int nblk = ceil_div(n, TB);
  for (int up = 0; up < 2; up++) {
    if (f.eoff + 1 >= kEpochSpan) {  // a very long solve: move the base on
      ir_epoch_bump_kernel<<<1, 64, 0, ctx->stream>>>(f.ird);
      f.eoff = 0;
    }
    f.eoff += 1;
    const uint32_t* ebase = &f.ird->ebase;
    ProfScope ps(ctx, K_MP_TRSV, 2.0 * (double)n * n / 2, 4.0 * (double)n * n / 2);
    if (up)
      trsv_blocked_kernel<1><<<nblk, kTrsvThreads, 0, ctx->stream>>>(n, f.F, n, f.invU, tmp, f.xpub, ebase, f.eoff,
                                                                      stop, ctx->err, nullptr);
    else
      trsv_blocked_kernel<0><<<nblk, kTrsvThreads, 0, ctx->stream>>>(n, f.F, n, f.invL, tmp, f.xpub, ebase, f.eoff,
                                                                      stop, ctx->err, nullptr);
    NMX_HIP(ctx, hipGetLastError());
  }
  if (f.pivot) {
    vec_f32_add_f64<<<ceil_div(n, 256), 256, 0, ctx->stream>>>(n, tmp, x, accumulate, stop);
  } else {
    rbt_vec_n_add_kernel<<<(unsigned)((work + 255) / 256), 256, 0, ctx->stream>>>(n, 1, tmp, 1, x, 1, f.dv,
                                                                                   accumulate, stop);
  }
  NMX_HIP(ctx, hipGetLastError());
  return NMX_OK;
}

// Y := op^-1 T for nrhs right-hand sides (n x nrhs, ld nrhs), op = L (unit
// lower) or U.  With the published words set up (mp_alloc_chains): one
// trsm_chain_kernel launch, T only read.  Otherwise right-looking over the
// 128-row blocks (T used as scratch): each diagonal block solved by its
// precomputed inverse (one GEMM), then the blocks still to solve updated with
// one GEMM -- two MFMA launches per block.
static int mp_trsm_blocks(nmx_ctx* ctx, MpFactor& f, int upper, int nrhs, float* T, float* Y) {
  const int n = f.n, nblk = ceil_div(n, TB);
  const float* inv = upper ? f.invU : f.invL;
  if (f.mpub && f.cplL && nrhs <= f.mpub_ld) {
    // one launch: the chains of trsm_chain_kernel (T is only read)
    f.mepoch += 1;
    ProfScope ps(ctx, K_TRSM, (double)n * n * nrhs, 4.0 * ((double)n * n / 2 + 2.0 * n * nrhs));
    // one 16-column group per workgroup: a hop's two 128 x 128 x 16 products
    // already take a CU's FP32 MFMA rate for about 1.7 us, so wider groups
    // lengthen the chain (measured: 32 or 64 columns per workgroup 1.5x and
    // 3x slower at 64 right-hand sides); the groups run as separate chains
    const int grid = nblk * ceil_div(nrhs, kChainR);
    if (upper)
      trsm_chain_kernel<1, 1><<<grid, kChainThreads, 0, ctx->stream>>>(n, nrhs, f.F, n, inv, f.cplU, T, nrhs, Y, nrhs,
                                                                        f.mpub, f.mpub_ld, f.mepoch, ctx->err, nullptr);
    else
      trsm_chain_kernel<0, 1><<<grid, kChainThreads, 0, ctx->stream>>>(n, nrhs, f.F, n, inv, f.cplL, T, nrhs, Y, nrhs,
                                                                        f.mpub, f.mpub_ld, f.mepoch, ctx->err, nullptr);
    NMX_HIP(ctx, hipGetLastError());
    return NMX_OK;
  }
  for (int s = 0; s < nblk; s++) {
    const int kb = upper ? nblk - 1 - s : s;
    const int r0 = kb * TB, bs = std::min(TB, n - r0);
    NMX_TRY(gemm_launch<float>(ctx, 0, 0, bs, nrhs, bs, 1.0f, inv + (int64_t)kb * TB * TB, TB,
                               T + (int64_t)r0 * nrhs, nrhs, 0.0f, Y + (int64_t)r0 * nrhs, nrhs));
    if (!upper && r0 + bs < n)
      NMX_TRY(gemm_launch<float>(ctx, 0, 0, n - r0 - bs, nrhs, bs, -1.0f, f.F + (int64_t)(r0 + bs) * n + r0, n,
                                 Y + (int64_t)r0 * nrhs, nrhs, 1.0f, T + (int64_t)(r0 + bs) * nrhs, nrhs));
    if (upper && r0 > 0)
      NMX_TRY(gemm_launch<float>(ctx, 0, 0, r0, nrhs, bs, -1.0f, f.F + r0, n, Y + (int64_t)r0 * nrhs, nrhs, 1.0f, T,
                                 nrhs));
  }
  return NMX_OK;
}

// D (FP64, n x nrhs, ld ldd) (+)= V U^-1 L^-1 W^T R for a block of
// right-hand sides: the FP32 solves are blocked over the factorization's
// 128 x 128 diagonal-block inverses with MFMA GEMM updates.  T and Y are
// n x nrhs FP32 scratch.
static int mp_solve_mat(nmx_ctx* ctx, MpFactor& f, int nrhs, const double* R, int ldr, double* D, int ldd,
                        float* T, float* Y, int accumulate) {
  const int n = f.n, n4 = n & ~3, h = n4 / 4;
  const int64_t work = ((int64_t)h + (n - n4)) * nrhs;
  if (f.pivot) {
    NMX_TRY((gather_rows<double, float>(ctx, n, nrhs, f.perm, R, ldr, T, nrhs)));
  } else {
    rbt_vec_t_kernel<<<(unsigned)((work + 255) / 256), 256, 0, ctx->stream>>>(n, nrhs, R, ldr, T, nrhs, f.dw);
    NMX_HIP(ctx, hipGetLastError());
  }
  NMX_TRY(mp_trsm_blocks(ctx, f, 0, nrhs, T, Y));  // Y = L^-1 T
  NMX_TRY(mp_trsm_blocks(ctx, f, 1, nrhs, Y, T));  // T = U^-1 Y
  if (f.pivot) {
    mat_f32_add_f64<<<(unsigned)(((int64_t)n * nrhs + 255) / 256), 256, 0, ctx->stream>>>(n, nrhs, T, nrhs, D, ldd,
                                                                                          accumulate);
  } else {
    rbt_vec_n_add_kernel<<<(unsigned)((work + 255) / 256), 256, 0, ctx->stream>>>(n, nrhs, T, nrhs, D, ldd, f.dv,
                                                                                  accumulate, nullptr);
  }
  NMX_HIP(ctx, hipGetLastError());
  return NMX_OK;
}

static int residual(nmx_ctx* ctx, int n, const double* A, int lda, const double* x, const double* b, double* r,
                    double* rowabs, unsigned long long* norms, double out[3]) {
  zero_u64<<<1, 64, 0, ctx->stream>>>(norms, 3);
  {
    ProfScope ps(ctx, K_RESIDUAL, 2.0 * (double)n * n, 8.0 * (double)n * n + 24.0 * n);
    residual_kernel<<<std::min(ceil_div(n, 4), kResidualBlocks), 256, 0, ctx->stream>>>(n, A, lda, x, b, r, rowabs,
                                                                                       norms, nullptr);
    NMX_HIP(ctx, hipGetLastError());
  }
  unsigned long long h[3];
  NMX_HIP(ctx, hipMemcpyAsync(h, norms, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  NMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int i = 0; i < 3; i++) memcpy(&out[i], &h[i], 8);
  return NMX_OK;
}

// ---- GMRES-IR correction ------------------------------------------------------
struct Gmres {
  double* V;   // (restart + 1) x n Krylov basis, FP64
  double* w;   // n: -A v_j
  double* h;   // 3 (restart + 2): two Gram-Schmidt passes and the norm
  double* y;   // restart: the combination added to x
  int restart;
};

static size_t gmres_doubles(int n, int restart) {
  return (size_t)(restart + 1) * n + n + 3 * (size_t)(restart + 2) + restart;
}

// x += d with d from GMRES(restart) on M^-1 A d = M^-1 r started at d = 0,
// stopped when the preconditioned residual estimate falls under tol_in of
// its start (or the Krylov space is exhausted).  *inner: iterations taken.
static int gmres_correction(nmx_ctx* ctx, MpFactor& f, const double* A, int lda, const double* r, double* x,
                            float* tmp, const Gmres& g, double tol_in, int* inner) {
  const int n = f.n, m = g.restart;
  const unsigned nbk = (unsigned)ceil_div(n, 256);
  double* hv = g.h;
  *inner = 0;
  NMX_TRY(mp_solve_vec(ctx, f, r, g.V, tmp, 0));  // z = M^-1 r
  gm_dots_kernel<<<1, kGmDotThreads, 0, ctx->stream>>>(n, g.V, g.V, hv);
  NMX_HIP(ctx, hipGetLastError());
  double beta2 = 0.0;
  NMX_HIP(ctx, hipMemcpyAsync(&beta2, hv, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  NMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const double beta = sqrt(beta2);
  if (!(beta > 0.0) || !std::isfinite(beta)) return NMX_OK;
  gm_scale_kernel<<<nbk, 256, 0, ctx->stream>>>(n, 1.0 / beta, g.V, g.V);
  NMX_HIP(ctx, hipGetLastError());
  std::vector<double> H((size_t)(m + 1) * m, 0.0), cs(m, 0.0), sn(m, 0.0), gv(m + 1, 0.0), hh(3 * (m + 2));
  gv[0] = beta;
  int j = 0;
  while (j < m) {
    const double* vj = g.V + (int64_t)j * n;
    double* vn = g.V + (int64_t)(j + 1) * n;
    {
      ProfScope ps(ctx, K_RESIDUAL, 2.0 * (double)n * n, 8.0 * (double)n * n);
      residual_kernel<<<std::min(ceil_div(n, 4), kResidualBlocks), 256, 0, ctx->stream>>>(n, A, lda, vj, nullptr,
                                                                                         g.w, nullptr, nullptr,
                                                                                         nullptr);
      NMX_HIP(ctx, hipGetLastError());
    }
    NMX_TRY(mp_solve_vec(ctx, f, g.w, vn, tmp, 0));  // u = -M^-1 A v_j
    // classical Gram-Schmidt, twice, against v_0 .. v_j; then |u|^2
    gm_dots_kernel<<<j + 1, kGmDotThreads, 0, ctx->stream>>>(n, g.V, vn, hv);
    gm_update_kernel<<<nbk, 256, 0, ctx->stream>>>(n, j + 1, g.V, hv, vn);
    gm_dots_kernel<<<j + 1, kGmDotThreads, 0, ctx->stream>>>(n, g.V, vn, hv + (m + 2));
    gm_update_kernel<<<nbk, 256, 0, ctx->stream>>>(n, j + 1, g.V, hv + (m + 2), vn);
    gm_dots_kernel<<<1, kGmDotThreads, 0, ctx->stream>>>(n, vn, vn, hv + 2 * (m + 2));
    NMX_HIP(ctx, hipGetLastError());
    NMX_HIP(ctx, hipMemcpyAsync(hh.data(), hv, sizeof(double) * hh.size(), hipMemcpyDeviceToHost, ctx->stream));
    NMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    // u held -(M^-1 A v_j), so the Hessenberg column is the negated projections
    for (int i = 0; i <= j; i++) H[(size_t)i * m + j] = -(hh[i] + hh[(m + 2) + i]);
    const double hn = sqrt(hh[2 * (m + 2)]);
    H[(size_t)(j + 1) * m + j] = hn;
    if (hn > 0.0) {
      gm_scale_kernel<<<nbk, 256, 0, ctx->stream>>>(n, -1.0 / hn, vn, vn);
      NMX_HIP(ctx, hipGetLastError());
    }
    // earlier rotations, then the one that zeroes H[j+1][j]
    for (int i = 0; i < j; i++) {
      const double a = H[(size_t)i * m + j], b2 = H[(size_t)(i + 1) * m + j];
      H[(size_t)i * m + j] = cs[i] * a + sn[i] * b2;
      H[(size_t)(i + 1) * m + j] = -sn[i] * a + cs[i] * b2;
    }
    const double a = H[(size_t)j * m + j], b2 = H[(size_t)(j + 1) * m + j];
    const double rho = hypot(a, b2);
    cs[j] = rho > 0.0 ? a / rho : 1.0;
    sn[j] = rho > 0.0 ? b2 / rho : 0.0;
    H[(size_t)j * m + j] = rho;
    H[(size_t)(j + 1) * m + j] = 0.0;
    gv[j + 1] = -sn[j] * gv[j];
    gv[j] = cs[j] * gv[j];
    j++;
    if (!(fabs(gv[j]) > tol_in * beta) || !(hn > 0.0)) break;
  }
  // y = H[0:j, 0:j]^-1 g[0:j]; x += V y (as x -= V (-y))
  std::vector<double> yv(j, 0.0);
  for (int i = j - 1; i >= 0; i--) {
    double t = gv[i];
    for (int c = i + 1; c < j; c++) t -= H[(size_t)i * m + c] * yv[c];
    yv[i] = H[(size_t)i * m + i] != 0.0 ? t / H[(size_t)i * m + i] : 0.0;
  }
  for (double& v : yv) v = -v;
  NMX_HIP(ctx, hipMemcpyAsync(g.y, yv.data(), sizeof(double) * j, hipMemcpyHostToDevice, ctx->stream));
  gm_update_kernel<<<nbk, 256, 0, ctx->stream>>>(n, j, g.V, g.y, x);
  NMX_HIP(ctx, hipGetLastError());
  NMX_HIP(ctx, hipStreamSynchronize(ctx->stream));  // yv's host memory is released on return
  *inner = j;
  return NMX_OK;
}

// ---- C ABI -------------------------------------------------------------------
NMX_API int nmx_gesv_mpir(nmx_ctx* ctx, int n, const double* A, int lda, const double* b, double* x,
                          const nmx_ir_cfg* cfg_in, nmx_ir_report* rep) {
  if (!ctx) return -1;
  if (n <= 0) return -2;
  if (!A) return -3;
  if (lda < n) return -4;
  if (!b) return -5;
  if (!x) return -6;
  nmx_ir_cfg cfg;
  if (cfg_in) {
    cfg = *cfg_in;
  } else {
    nmx_ir_cfg_default(&cfg);
  }
  if (cfg.precision < 0 || cfg.precision > 3) return -7;
  if (cfg.block <= 0) cfg.block = 256;
  if (cfg.block % 64) return -7;
  if (cfg.solver < 0 || cfg.solver > 2) return -7;
  if (cfg.solver != 1 && (cfg.gmres_restart < 1 || cfg.gmres_restart > 100)) return -7;
  if (cfg.max_iter < 0) return -7;
  if (n > 65536) return NMX_ERR_UNSUPPORTED;
  if (cfg.pivot < 0 || cfg.pivot > 2) return -7;
  if (cfg.pivot == 1 && n > kLeafMaxRows) return NMX_ERR_UNSUPPORTED;  // checked before anything is allocated
  // Under capture only classical refinement (its stop test runs on the
  // device); GMRES-IR decides on the host.
  const bool cap = ctx->capturing;
  if (cap && (cfg.solver == 2 || cfg.pivot == 1)) return NMX_ERR_UNSUPPORTED;  // host decisions / host-side epochs
  nmx_ir_report local{};
  nmx_ir_report* R = rep ? rep : &local;
  *R = nmx_ir_report{};

  MpFactor f;
  f.pivot = cfg.pivot;
  int st = mp_alloc(ctx, f, n, cfg.block, cfg.precision);
  f.pivot = cfg.pivot;
  void* wsp = nullptr;
  const size_t gm_doubles = cfg.solver == 1 ? 0 : gmres_doubles(n, cfg.gmres_restart);
  if (st == NMX_OK)
    st = nmx_workspace(ctx, sizeof(double) * (3 * (size_t)n + 8 + gm_doubles) + sizeof(float) * (size_t)n, &wsp);
  if (st != NMX_OK) {
    mp_free(ctx, f);
    return st;
  }
  double* r = (double*)wsp;
  double* rowabs = r + n;
  double* d = rowabs + n;
  unsigned long long* norms = (unsigned long long*)(d + n);
  Gmres gm{};
  gm.restart = cfg.gmres_restart;
  if (gm_doubles) {
    gm.V = (double*)(norms + 8);
    gm.w = gm.V + (size_t)(gm.restart + 1) * n;
    gm.h = gm.w + n;
    gm.y = gm.h + 3 * (size_t)(gm.restart + 2);
  }
  float* tmp = (float*)((double*)(norms + 8) + gm_doubles);

  hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
  if (!cap) {
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    hipEventCreate(&e2);
    hipEventRecord(e0, ctx->stream);
  }
  ir_reset_kernel<<<1, 64, 0, ctx->stream>>>(f.ird);  // also moves the TRSV epoch base on
  NMX_HIP(ctx, hipGetLastError());
  st = mp_factor(ctx, f, A, lda, cfg.seed);
  if (!cap) hipEventRecord(e1, ctx->stream);
  if (st == NMX_OK) st = mp_solve_vec(ctx, f, b, x, tmp, 0);  // x0 = M^-1 b
  int it = 0;
  double eta = INFINITY, prev = INFINITY, anorm = 0.0;
  int inner_total = 0;
  IrDev hs{};
  bool gmres_phase = cfg.solver == 2;
  if (st == NMX_OK && !gmres_phase) {
    // Classical refinement as a fixed launch sequence: step s is the FP64
    // residual, the device-side stop test, and (s < max_iter) the
    // correction, every kernel returning at once after the stop.  Eagerly,
    // the host looks at the stop word every kStepsPerCheck steps; a captured
    // graph holds all max_iter + 1 steps.
    constexpr int kStepsPerCheck = 4;
    const int* stop = &f.ird->stop;
    const int steps = cfg.max_iter + 1, chunk = cap ? steps : kStepsPerCheck;
    for (int s0 = 0; s0 < steps && st == NMX_OK; s0 += chunk) {
      for (int sk = s0; sk < std::min(s0 + chunk, steps) && st == NMX_OK; sk++) {
        zero_u64<<<1, 64, 0, ctx->stream>>>(norms, 3);
        {
          ProfScope ps(ctx, K_RESIDUAL, 2.0 * (double)n * n, 8.0 * (double)n * n + 24.0 * n);
          residual_kernel<<<std::min(ceil_div(n, 4), kResidualBlocks), 256, 0, ctx->stream>>>(
              n, A, lda, x, b, r, sk == 0 ? rowabs : nullptr, norms, stop);
        }
        ir_decide_kernel<<<1, 64, 0, ctx->stream>>>(f.ird, norms, cfg.tol, cfg.max_iter, sk == 0);
        NMX_HIP(ctx, hipGetLastError());
        if (sk < cfg.max_iter) st = mp_solve_vec(ctx, f, r, x, tmp, 1, stop);
      }
      if (!cap && st == NMX_OK) {
        int32_t hstop = 0;
        NMX_HIP(ctx, hipMemcpyAsync(&hstop, stop, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
        NMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (hstop) break;
      }
    }
    if (!cap && st == NMX_OK) {
      NMX_HIP(ctx, hipMemcpyAsync(&hs, f.ird, sizeof(IrDev), hipMemcpyDeviceToHost, ctx->stream));
      NMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
      it = hs.it;
      eta = hs.eta;
      prev = hs.prev;
      anorm = hs.anorm;
      for (int i = 0; i < 32; i++) R->history[i] = hs.history[i];
      // auto: a classical run that stopped without converging and with steps
      // left stagnated -- GMRES-IR takes over from here (solver 0: classical
      // while each step at least halves the backward error; the
      // preconditioner is applied in FP32, so where classical steps still
      // contract GMRES costs more solves for the same accuracy)
      if (cfg.solver == 0 && !hs.converged && it < cfg.max_iter && eta == eta) {
        gmres_phase = true;
        prev = INFINITY;
      }
    }
  }
  if (gmres_phase) {
    double nrm[3] = {0, 0, 0};
    while (st == NMX_OK) {
      st = residual(ctx, n, A, lda, x, b, r, it == 0 && anorm == 0.0 ? rowabs : nullptr, norms, nrm);
      if (st != NMX_OK) break;
      if (anorm == 0.0) anorm = nrm[2];
      eta = nrm[0] / (anorm * nrm[1]);
      R->history[it < 32 ? it : 31] = eta;
      if (!(eta > cfg.tol) || it >= cfg.max_iter) break;
      if (it > 0 && !(eta < 0.5 * prev)) break;  // stagnation
      prev = eta;
      // reduce the preconditioned residual by what the target still needs, with margin
      const double tol_in = std::min(1e-2, std::max(1e-10, 0.1 * cfg.tol / eta));
      int inner = 0;
      st = gmres_correction(ctx, f, A, lda, r, x, tmp, gm, tol_in, &inner);
      inner_total += inner;
      it++;
    }
  }
  if (cap) {
    // the graph records the report on the device: nmx_ir_report_read after a replay
    mp_free(ctx, f);
    return st;
  }
  R->gmres_iterations = inner_total;
  hipEventRecord(e2, ctx->stream);
  hipEventSynchronize(e2);
  float fms = 0.f, rms = 0.f;
  hipEventElapsedTime(&fms, e0, e1);
  hipEventElapsedTime(&rms, e1, e2);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipEventDestroy(e2);
  int32_t hinfo = 0;
  hipMemcpy(&hinfo, f.info, sizeof(int32_t), hipMemcpyDeviceToHost);
  R->iterations = it;
  R->backward_error = eta;
  R->converged = (eta <= cfg.tol) ? 1 : 0;
  R->factor_ms = fms;
  R->refine_ms = rms;
  R->info = hinfo;
  mp_free(ctx, f);
  if (st != NMX_OK) return st;
  int dev = nmx_ctx_check(ctx);
  if (dev != NMX_OK) return dev;
  return R->converged ? NMX_OK : NMX_ERR_NOT_CONVERGED;
}

NMX_API int nmx_ir_report_read(nmx_ctx* ctx, nmx_ir_report* rep) {
  if (!ctx) return -1;
  if (!rep) return -2;
  if (ctx->capturing) return NMX_ERR_UNSUPPORTED;
  MpBuffers* b = (MpBuffers*)ctx->mp_cache;
  if (!b || !b->ird || !b->info) return NMX_ERR_UNSUPPORTED;  // no mixed-precision solve has run on this context
  IrDev hs{};
  int32_t hinfo = 0;
  NMX_HIP(ctx, hipMemcpyAsync(&hs, b->ird, sizeof(IrDev), hipMemcpyDeviceToHost, ctx->stream));
  NMX_HIP(ctx, hipMemcpyAsync(&hinfo, b->info, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  NMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *rep = nmx_ir_report{};
  rep->iterations = hs.it;
  rep->converged = hs.converged;
  rep->backward_error = hs.eta;
  rep->info = hinfo;
  for (int i = 0; i < 32; i++) rep->history[i] = hs.history[i];
  return nmx_ctx_check(ctx);
}

NMX_API int nmx_gesv_mpir_nrhs(nmx_ctx* ctx, int n, int nrhs, const double* A, int lda, const double* B, int ldb,
                               double* X, int ldx, const nmx_ir_cfg* cfg_in, nmx_ir_report* reps) {
  if (!ctx) return -1;
  if (ctx->capturing) return NMX_ERR_UNSUPPORTED;
  if (n <= 0) return -2;
  if (nrhs <= 0) return -3;
  if (!A) return -4;
  if (lda < n) return -5;
  if (!B) return -6;
  if (ldb < nrhs) return -7;
  if (!X) return -8;
  if (ldx < nrhs) return -9;
  nmx_ir_cfg cfg;
  if (cfg_in) {
    cfg = *cfg_in;
  } else {
    nmx_ir_cfg_default(&cfg);
  }
  if (cfg.precision < 0 || cfg.precision > 3) return -10;
  if (cfg.block <= 0) cfg.block = 256;
  if (cfg.block % 64) return -10;
  if (n > 65536) return NMX_ERR_UNSUPPORTED;
  if (cfg.pivot < 0 || cfg.pivot > 2) return -10;
  if (cfg.pivot == 1 && n > kLeafMaxRows) return NMX_ERR_UNSUPPORTED;
  std::vector<nmx_ir_report> local;
  if (!reps) {
    local.resize(nrhs);
    reps = local.data();
  }
  for (int c = 0; c < nrhs; c++) reps[c] = nmx_ir_report{};

  MpFactor f;
  f.pivot = cfg.pivot;
  int st = mp_alloc(ctx, f, n, cfg.block, cfg.precision);
  f.pivot = cfg.pivot;
  if (st == NMX_OK) st = mp_alloc_chains(ctx, f, nrhs);
  void* wsp = nullptr;
  const size_t blk = (size_t)n * nrhs;
  if (st == NMX_OK)
    st = nmx_workspace(ctx, sizeof(double) * blk * (1 + kResSplit) + sizeof(unsigned long long) * (2 * (size_t)nrhs + 8) +
                                sizeof(int32_t) * (size_t)nrhs + 2 * sizeof(float) * blk + 1024,
                       &wsp);
  if (st != NMX_OK) {
    mp_free(ctx, f);
    return st;
  }
  double* R = (double*)wsp;
  unsigned long long* norms = (unsigned long long*)(R + blk);  // [0, nrhs) |r|, [nrhs, 2 nrhs) |x|, then |A|
  int32_t* flag = (int32_t*)(norms + 2 * (size_t)nrhs + 8);
  float* T = (float*)(((uintptr_t)(flag + nrhs) + 255) & ~(uintptr_t)255);
  float* Yb = (float*)(((uintptr_t)(T + blk) + 255) & ~(uintptr_t)255);
  double* P = (double*)(((uintptr_t)(Yb + blk) + 255) & ~(uintptr_t)255);  // kResSplit partial products

  hipEvent_t e0, e1, e2;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  hipEventCreate(&e2);
  hipEventRecord(e0, ctx->stream);
  // ||A||_inf: from the pass that casts A when the factor pivots, else its own
  zero_u64<<<1, 64, 0, ctx->stream>>>(norms + 2 * nrhs, 1);
  st = mp_factor(ctx, f, A, lda, cfg.seed, 0, f.pivot ? norms + 2 * nrhs : nullptr);
  if (st == NMX_OK) st = mp_couple(ctx, f);
  hipEventRecord(e1, ctx->stream);
  if (st == NMX_OK) st = mp_solve_mat(ctx, f, nrhs, B, ldb, X, ldx, T, Yb, 0);  // X0 = M^-1 B
  double anorm = 0.0;
  if (st == NMX_OK) {
    if (!f.pivot)
      row_abs_max_kernel<<<std::min(ceil_div(n, 4), kResidualBlocks), 256, 0, ctx->stream>>>(n, A, lda,
                                                                                            norms + 2 * nrhs);
    unsigned long long ha = 0;
    if (hipMemcpyAsync(&ha, norms + 2 * nrhs, 8, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
      st = NMX_ERR_HIP;
    memcpy(&anorm, &ha, 8);
  }
  std::vector<unsigned long long> hn(2 * (size_t)nrhs);
  // two host copies of the column flags, used alternately: the copy queued
  // at step it is not rewritten before the norms read-back of step it + 1
  // has synchronized past it
  std::vector<int32_t> hflag_ring(2 * (size_t)nrhs, 0);
  std::vector<double> eta(nrhs, INFINITY), prev(nrhs, INFINITY);
  std::vector<int> done(nrhs, 0);
  int it = 0;
  while (st == NMX_OK) {
    // R = B - A X in FP64: the MFMA GEMM split over K (an n x nrhs output
    // alone is one workgroup per 64 rows), the partials summed in order;
    // then the per-column norms
    {
      ProfScope ps(ctx, K_RESIDUAL, 2.0 * (double)n * n * nrhs, 8.0 * (double)n * n + 24.0 * (double)n * nrhs);
      st = gemm_launch_splitk<double>(ctx, n, nrhs, n, 1.0, A, lda, X, ldx, P, nrhs, kResSplit, (int64_t)blk);
      if (st != NMX_OK) break;
      residual_combine_kernel<<<(unsigned)((blk + 255) / 256), 256, 0, ctx->stream>>>(n, nrhs, B, ldb, P, R);
      NMX_HIP(ctx, hipGetLastError());
    }
    {
      const int z = 2 * nrhs;
      for (int o = 0; o < z; o += 64) zero_u64<<<1, 64, 0, ctx->stream>>>(norms + o, std::min(64, z - o));
    }
    colmax_kernel<<<std::min(n, kColMaxBlocks), 256, 0, ctx->stream>>>(n, nrhs, R, nrhs, norms);
    colmax_kernel<<<std::min(n, kColMaxBlocks), 256, 0, ctx->stream>>>(n, nrhs, X, ldx, norms + nrhs);
    NMX_HIP(ctx, hipGetLastError());
    NMX_HIP(ctx, hipMemcpyAsync(hn.data(), norms, sizeof(unsigned long long) * hn.size(), hipMemcpyDeviceToHost,
                                ctx->stream));
    NMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    int active = 0;
    for (int c = 0; c < nrhs; c++) {
      if (done[c]) continue;
      double rm, xm;
      memcpy(&rm, &hn[c], 8);
      memcpy(&xm, &hn[nrhs + c], 8);
      eta[c] = rm / (anorm * xm);
      reps[c].history[it < 32 ? it : 31] = eta[c];
      reps[c].iterations = it;
      // converged, out of steps, or stagnating: this column is finished
      if (!(eta[c] > cfg.tol) || it >= cfg.max_iter || (it > 0 && !(eta[c] < 0.5 * prev[c]))) {
        done[c] = 1;
      } else {
        prev[c] = eta[c];
        active++;
      }
    }
    for (int c = 0; c < nrhs; c++) hflag_ring[(size_t)(it & 1) * nrhs + c] = done[c];
    if (!active) break;
    NMX_HIP(ctx, hipMemcpyAsync(flag, hflag_ring.data() + (size_t)(it & 1) * nrhs, sizeof(int32_t) * nrhs,
                                hipMemcpyHostToDevice, ctx->stream));
    zero_cols_kernel<<<(unsigned)((blk + 255) / 256), 256, 0, ctx->stream>>>(n, nrhs, R, nrhs, flag);
    NMX_HIP(ctx, hipGetLastError());
    st = mp_solve_mat(ctx, f, nrhs, R, nrhs, X, ldx, T, Yb, 1);
    it++;
  }
  hipEventRecord(e2, ctx->stream);
  hipEventSynchronize(e2);
  float fms = 0.f, rms = 0.f;
  hipEventElapsedTime(&fms, e0, e1);
  hipEventElapsedTime(&rms, e1, e2);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipEventDestroy(e2);
  int32_t hinfo = 0;
  hipMemcpy(&hinfo, f.info, sizeof(int32_t), hipMemcpyDeviceToHost);
  int all = 1;
  for (int c = 0; c < nrhs; c++) {
    reps[c].backward_error = eta[c];
    reps[c].converged = eta[c] <= cfg.tol ? 1 : 0;
    reps[c].factor_ms = fms;
    reps[c].refine_ms = rms;
    reps[c].info = hinfo;
    all &= reps[c].converged;
  }
  mp_free(ctx, f);
  if (st != NMX_OK) return st;
  int dev = nmx_ctx_check(ctx);
  if (dev != NMX_OK) return dev;
  return all ? NMX_OK : NMX_ERR_NOT_CONVERGED;
}

// Autotuner of the panel width: one factorization (plus the first solve and
// residual) of a synthetic N(0,1) matrix of order n per candidate width, the
// fastest factorization time's width into cfg->block.  Other fields kept.
NMX_API int nmx_ir_autotune(nmx_ctx* ctx, int n, nmx_ir_cfg* cfg) {
  if (!ctx) return -1;
  if (n < 64 || n > 65536) return -2;
  if (!cfg) return -3;
  if (ctx->capturing) return NMX_ERR_UNSUPPORTED;
  static const int kCand[] = {128, 256, 512};
  double *A = nullptr, *b = nullptr, *x = nullptr;
  if (hipMalloc(&A, sizeof(double) * (size_t)n * n) != hipSuccess) return NMX_ERR_ALLOC;
  if (hipMalloc(&b, sizeof(double) * (size_t)n) != hipSuccess || hipMalloc(&x, sizeof(double) * (size_t)n) != hipSuccess) {
    hipFree(A);
    if (b) hipFree(b);
    return NMX_ERR_ALLOC;
  }
  int st = nmx_fill_random_f64(ctx, 1, 0x7ab1eULL, 0, (int64_t)n * n, 0.0, 0.0, A);
  if (st == NMX_OK) st = nmx_fill_random_f64(ctx, 1, 0x7ab1fULL, 0, n, 0.0, 0.0, b);
  double best = INFINITY;
  int best_block = cfg->block > 0 ? cfg->block : 256;
  for (int ci = 0; ci < 3 && st == NMX_OK; ci++) {
    if (kCand[ci] > n) continue;
    nmx_ir_cfg c = *cfg;
    c.block = kCand[ci];
    c.max_iter = 0;
    c.solver = 1;
    double tot = 0.0;
    for (int rep = 0; rep < 3 && st == NMX_OK; rep++) {  // the first run is a warm-up
      nmx_ir_report r;
      const int s2 = nmx_gesv_mpir(ctx, n, A, n, b, x, &c, &r);
      if (s2 != NMX_OK && s2 != NMX_ERR_NOT_CONVERGED) st = s2;
      if (rep > 0) tot += r.factor_ms;
    }
    if (st == NMX_OK && tot < best) {
      best = tot;
      best_block = kCand[ci];
    }
  }
  hipStreamSynchronize(ctx->stream);
  hipFree(A);
  hipFree(b);
  hipFree(x);
  if (st != NMX_OK) return st;
  cfg->block = best_block;
  return NMX_OK;
}

NMX_API void nmx_ir_cfg_default(nmx_ir_cfg* cfg) {
  if (!cfg) return;
  cfg->precision = 2;
  cfg->block = 256;
  cfg->max_iter = 30;
  cfg->solver = 0;
  cfg->gmres_restart = 30;
  cfg->tol = 1e-15;
  cfg->seed = 0x5eedULL;
  cfg->pivot = 2;
}

// Test hook: the transformed matrix (stage 0) or its no-pivot factors
// (stage 1) copied to F (n x n FP32, ld n), the 4n butterfly diagonals (W's
// 2n then V's 2n) to diags.
NMX_API int nmx_mp_factor_debug(nmx_ctx* ctx, int n, const double* A, int lda, int precision, int block, uint64_t seed,
                                int stage, int pivot, float* F, float* diags, int32_t* info, int32_t* ipiv) {
  if (!ctx) return -1;
  if (n <= 0) return -2;
  if (lda < n) return -4;
  if (precision < 0 || precision > 3) return -5;
  if (block <= 0 || block % 64) return -6;
  if (pivot < 0 || pivot > 2) return -9;
  if (pivot == 1 && n > kLeafMaxRows) return NMX_ERR_UNSUPPORTED;
  MpFactor f;
  f.pivot = pivot;
  int st = mp_alloc(ctx, f, n, block, precision);
  if (st == NMX_OK) st = mp_factor(ctx, f, A, lda, seed, stage == 0);
  if (st == NMX_OK) {
    hipMemcpyAsync(F, f.F, sizeof(float) * (size_t)n * n, hipMemcpyDeviceToDevice, ctx->stream);
    hipMemcpyAsync(diags, f.dw, sizeof(float) * 4 * (size_t)n, hipMemcpyDeviceToDevice, ctx->stream);
    hipMemcpyAsync(info, f.info, sizeof(int32_t), hipMemcpyDeviceToDevice, ctx->stream);
    if (ipiv) hipMemcpyAsync(ipiv, f.ipiv, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToDevice, ctx->stream);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) st = NMX_ERR_HIP;
  }
  mp_free(ctx, f);
  return st;
}
