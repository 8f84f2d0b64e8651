// MFMA GEMM update C = beta C + alpha op(A) op(B) for the blocked drivers.
//
// This is the trailing update numax hands to MAX's linalg.matmul with a
// subtracting epilogue (numax/linalg/lu.mojo:629-673, panel.mojo:557-582,
// cholesky.mojo:126-225, triangular.mojo:193-218, qr.mojo:262-327).  Unlike
// MAX's matmul it reads strided operands in place, so the drivers never pack
// L21 / U12, and it writes the epilogue straight into the strided matrix.
//
// Tiling for CDNA4: a 256-thread workgroup owns a 128x128 tile of C, four
// 64-wide wavefronts in a 2x2 arrangement each owning 64x64 = 4x4 MFMA
// 16x16 tiles (64x64 tiles of 2x2 per wavefront when 128x128 ones would not
// fill the device).  K advances 16 at a time through double-buffered LDS with
// the next K slab prefetched into registers while the current one is consumed.
// FP64 runs v_mfma_f64_16x16x4_f64, FP32 v_mfma_f32_16x16x4_f32.
//
// Tile order is XCD-aware: workgroup b runs on XCD b % 8, so each XCD is
// given a contiguous run of tiles (grouped 8 tile-rows at a time) and the
// A row-panels and B column-panels it reads stay in that XCD's L2.
#include "nmx_internal.h"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;

template <>
struct Mfma<double> {
  typedef d4 acc_t;
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // Output item i of lane l sits at row l/16 + 4 i (measured on gfx950).
  static __device__ __forceinline__ int row(int lane, int i) { return (lane >> 4) + 4 * i; }
};

template <>
struct Mfma<float> {
  typedef f4 acc_t;
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int i) { return 4 * (lane >> 4) + i; }
};

constexpr int BM = 128, BK = 16, PAD = 4;  // the large tile, K slab, LDS row pad

// Two workgroups per CU (two wavefronts per SIMD): the FP64 accumulators then
// live in the 256 unified registers with a few epilogue-only spills, and the
// second wave hides the LDS and global latency of the first (1.5x over one
// wave per SIMD on the 16128x16128x256 trailing update).  Two double-buffered
// FP64 tiles take 68 KB of LDS each, which is what 160 KB holds.
template <typename T, int TA, int TB, int TILE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
gemm_kernel(int M, int N, int K, T alpha, const T* __restrict__ A, int lda, const T* __restrict__ B, int ldb, T beta,
            T* __restrict__ C, int ldc, int tilesM, int tilesN, int lower_only, int ksplit, int kchunk, int64_t csplit) {
  typedef typename Mfma<T>::acc_t acc_t;
  // TILE x TILE of C per workgroup; each wavefront owns WT x WT = FR x FR
  // MFMA tiles; each thread stages LPT elements of each K slab
  constexpr int TM = TILE, TN = TILE, WT = TILE / 2, FR = WT / 16, LPT = TILE * BK / 256;
  static_assert(FR >= 1 && LPT >= 1 && BK % LPT == 0, "tile shape");
  __shared__ T As[2][BK][TM + PAD];
  __shared__ T Bs[2][BK][TN + PAD];

  // ---- split K: slice q of the grid (a multiple of kXcds workgroups) takes
  // K range [q kchunk, (q + 1) kchunk) into its own C at C + q csplit ----
  const int per_slice = gridDim.x / ksplit, q = blockIdx.x / per_slice;
  {
    const int k_lo = q * kchunk;
    K = max(0, min(kchunk, K - k_lo));
    A += TA ? (int64_t)k_lo * lda : k_lo;
    B += TB ? k_lo : (int64_t)k_lo * ldb;
    C += q * csplit;
  }
  // ---- tile assignment: contiguous chunk per XCD, grouped rows inside ----
  const int total = lower_only ? tilesM * (tilesM + 1) / 2 : tilesM * tilesN;
  const int b = blockIdx.x - q * per_slice;
  const int per = (total + kXcds - 1) / kXcds;
  const int lin = (b % kXcds) * per + b / kXcds;
  if (lin >= total) return;
  int tm, tn;
  if (lower_only) {
    // lin -> (tm, tn) with tn <= tm, row-major over the lower block triangle
    int r = (int)((sqrtf(8.0f * (float)lin + 1.0f) - 1.0f) * 0.5f);
    while ((r + 1) * (r + 2) / 2 <= lin) r++;
    while (r * (r + 1) / 2 > lin) r--;
    tm = r;
    tn = lin - r * (r + 1) / 2;
  } else {
    const int GM = 8;
    int group = lin / (GM * tilesN);
    int first = group * GM;
    int gm = min(tilesM - first, GM);
    int in = lin % (GM * tilesN);
    tm = first + in % gm;
    tn = in / gm;
  }
  const int m0 = tm * TM, n0 = tn * TN;

  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wm = wave >> 1, wn = wave & 1;

  acc_t acc[FR][FR];
#pragma unroll
  for (int i = 0; i < FR; i++)
#pragma unroll
    for (int j = 0; j < FR; j++) acc[i][j] = acc_t{0, 0, 0, 0};

  T ra[LPT], rb[LPT];
  // Slab loads through buffer descriptors: an element outside the operand
  // gets an offset past the descriptor's range and reads as 0, so edge and
  // interior tiles run the same branch-free code with all loads in flight.
  constexpr uint32_t E = sizeof(T);
  auto load_tiles = [&](int k0) {
    if (!TA) {
      const int r = t / (BK / LPT), kh = (t % (BK / LPT)) * LPT, gm = m0 + r;
      const __amdgpu_buffer_rsrc_t ra_ = buf_rsrc(A + (int64_t)m0 * lda);
#pragma unroll
      for (int i = 0; i < LPT; i++) {
        const int gk = k0 + kh + i;
        const uint32_t off = ((gm < M) & (gk < K)) ? ((uint32_t)r * (uint32_t)lda + (uint32_t)gk) * E : kBufOob;
        ra[i] = buf_load<T>(ra_, off);
      }
    } else {
      const int kk = t / (TM / LPT), mc = (t % (TM / LPT)) * LPT, gk = k0 + kk;
      const __amdgpu_buffer_rsrc_t ra_ = buf_rsrc(A + (int64_t)k0 * lda + m0);
#pragma unroll
      for (int i = 0; i < LPT; i++) {
        const int gm = m0 + mc + i;
        const uint32_t off = ((gm < M) & (gk < K)) ? ((uint32_t)kk * (uint32_t)lda + (uint32_t)(mc + i)) * E : kBufOob;
        ra[i] = buf_load<T>(ra_, off);
      }
    }
    if (!TB) {
      const int kk = t / (TN / LPT), nc = (t % (TN / LPT)) * LPT, gk = k0 + kk;
      const __amdgpu_buffer_rsrc_t rb_ = buf_rsrc(B + (int64_t)k0 * ldb + n0);
#pragma unroll
      for (int i = 0; i < LPT; i++) {
        const int gn = n0 + nc + i;
        const uint32_t off = ((gn < N) & (gk < K)) ? ((uint32_t)kk * (uint32_t)ldb + (uint32_t)(nc + i)) * E : kBufOob;
        rb[i] = buf_load<T>(rb_, off);
      }
    } else {
      const int r = t / (BK / LPT), kh = (t % (BK / LPT)) * LPT, gn = n0 + r;
      const __amdgpu_buffer_rsrc_t rb_ = buf_rsrc(B + (int64_t)n0 * ldb);
#pragma unroll
      for (int i = 0; i < LPT; i++) {
        const int gk = k0 + kh + i;
        const uint32_t off = ((gn < N) & (gk < K)) ? ((uint32_t)r * (uint32_t)ldb + (uint32_t)gk) * E : kBufOob;
        rb[i] = buf_load<T>(rb_, off);
      }
    }
  };
  auto store_tiles = [&](int buf) {
    if (!TA) {
      int r = t / (BK / LPT), kh = (t % (BK / LPT)) * LPT;
#pragma unroll
      for (int i = 0; i < LPT; i++) As[buf][kh + i][r] = ra[i];
    } else {
      int kk = t / (TM / LPT), mc = (t % (TM / LPT)) * LPT;
#pragma unroll
      for (int i = 0; i < LPT; i++) As[buf][kk][mc + i] = ra[i];
    }
    if (!TB) {
      int kk = t / (TN / LPT), nc = (t % (TN / LPT)) * LPT;
#pragma unroll
      for (int i = 0; i < LPT; i++) Bs[buf][kk][nc + i] = rb[i];
    } else {
      int r = t / (BK / LPT), kh = (t % (BK / LPT)) * LPT;
#pragma unroll
      for (int i = 0; i < LPT; i++) Bs[buf][kh + i][r] = rb[i];
    }
  };

  const int ksteps = (K + BK - 1) / BK;
  load_tiles(0);
  store_tiles(0);
  __syncthreads();
  for (int s = 0; s < ksteps; s++) {
    const int buf = s & 1;
    if (s + 1 < ksteps) load_tiles((s + 1) * BK);
    // fragments for k-step kq+1 are read while kq's MFMAs issue
    T af[2][FR], bf[2][FR];
#pragma unroll
    for (int i = 0; i < FR; i++) {
      af[0][i] = As[buf][lane >> 4][wm * WT + i * 16 + (lane & 15)];
      bf[0][i] = Bs[buf][lane >> 4][wn * WT + i * 16 + (lane & 15)];
    }
#pragma unroll
    for (int kq = 0; kq < BK / 4; kq++) {
      const int cur = kq & 1;
      if (kq + 1 < BK / 4) {
        const int kr = (kq + 1) * 4 + (lane >> 4);
#pragma unroll
        for (int i = 0; i < FR; i++) {
          af[cur ^ 1][i] = As[buf][kr][wm * WT + i * 16 + (lane & 15)];
          bf[cur ^ 1][i] = Bs[buf][kr][wn * WT + i * 16 + (lane & 15)];
        }
      }
#pragma unroll
      for (int i = 0; i < FR; i++)
#pragma unroll
        for (int j = 0; j < FR; j++) acc[i][j] = Mfma<T>::mma(af[cur][i], bf[cur][j], acc[i][j]);
    }
    if (s + 1 < ksteps) {
      store_tiles(buf ^ 1);
      __syncthreads();
    }
  }

  // ---- epilogue straight into the strided C: every C load is issued
  // before any is used (buffer loads, out-of-range lanes read 0)
  T cv[FR][FR][4];
  const bool use_c = beta != T(0);
  if (use_c) {
    const __amdgpu_buffer_rsrc_t rc = buf_rsrc(C + (int64_t)m0 * ldc + n0);
#pragma unroll
    for (int i = 0; i < FR; i++)
#pragma unroll
      for (int j = 0; j < FR; j++)
#pragma unroll
        for (int e = 0; e < 4; e++) {
          const int lr = wm * WT + i * 16 + Mfma<T>::row(lane, e);
          const int lc = wn * WT + j * 16 + (lane & 15);
          const uint32_t off =
              ((m0 + lr < M) & (n0 + lc < N)) ? ((uint32_t)lr * (uint32_t)ldc + (uint32_t)lc) * E : kBufOob;
          cv[i][j][e] = buf_load<T>(rc, off);
        }
  }
#pragma unroll
  for (int i = 0; i < FR; i++)
#pragma unroll
    for (int j = 0; j < FR; j++) {
      const int col = n0 + wn * WT + j * 16 + (lane & 15);
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int row = m0 + wm * WT + i * 16 + Mfma<T>::row(lane, e);
        T v = alpha * acc[i][j][e];
        if (use_c) v += beta * cv[i][j][e];
        if ((row < M) & (col < N)) C[(int64_t)row * ldc + col] = v;
      }
    }
}

template <typename T>
__global__ void scale_kernel(int M, int N, T beta, T* C, int ldc) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)M * N) return;
  int r = (int)(idx / N), c = (int)(idx % N);
  T* p = C + (int64_t)r * ldc + c;
  *p = beta == T(0) ? T(0) : beta * *p;
}

}  // namespace

template <typename T>
int gemm_launch(nmx_ctx* ctx, int transa, int transb, int m, int n, int k, T alpha, const T* A, int lda, const T* B,
                int ldb, T beta, T* C, int ldc, int lower_only) {
  if (m <= 0 || n <= 0) return NMX_OK;
  if (k <= 0 || alpha == T(0)) {
    if (beta == T(1)) return NMX_OK;
    int64_t total = (int64_t)m * n;
    scale_kernel<T><<<(unsigned)((total + 255) / 256), 256, 0, ctx->stream>>>(m, n, beta, C, ldc);
    NMX_HIP(ctx, hipGetLastError());
    return NMX_OK;
  }
  if (lower_only && m != n) return NMX_ERR_UNSUPPORTED;
  // 128x128 tiles unless they would leave the device under-filled (fewer
  // than two per CU, the in-panel and late trailing updates): 64x64 tiles
  // then quadruple the workgroups and quarter each one's MFMA chain.  The
  // k order of every C element is the same either way.
  auto tiles = [&](int tile) {
    const int tm = ceil_div(m, tile), tn = ceil_div(n, tile);
    return lower_only ? tm * (tm + 1) / 2 : tm * tn;
  };
  const int tile = tiles(BM) < 2 * ctx->num_cus ? 64 : BM;
  const int tilesM = ceil_div(m, tile), tilesN = ceil_div(n, tile);
  const int grid = ceil_div(tiles(tile), kXcds) * kXcds;
  dim3 g(grid), blk(256);
  const double flops = 2.0 * m * (double)n * k * (lower_only ? 0.5 : 1.0);
  ProfScope ps(ctx, K_GEMM, flops, (double)sizeof(T) * ((double)m * k + (double)k * n + 2.0 * m * n));
  auto run = [&](auto tile_tag) {
    constexpr int TL = decltype(tile_tag)::value;
    if (!transa && !transb)
      gemm_kernel<T, 0, 0, TL><<<g, blk, 0, ctx->stream>>>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tilesM, tilesN,
                                                            lower_only, 1, k, 0);
    else if (!transa && transb)
      gemm_kernel<T, 0, 1, TL><<<g, blk, 0, ctx->stream>>>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tilesM, tilesN,
                                                            lower_only, 1, k, 0);
    else if (transa && !transb)
      gemm_kernel<T, 1, 0, TL><<<g, blk, 0, ctx->stream>>>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tilesM, tilesN,
                                                            lower_only, 1, k, 0);
    else
      gemm_kernel<T, 1, 1, TL><<<g, blk, 0, ctx->stream>>>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tilesM, tilesN,
                                                            lower_only, 1, k, 0);
  };
  if (tile == 64)
    run(std::integral_constant<int, 64>());
  else
    run(std::integral_constant<int, BM>());
  NMX_HIP(ctx, hipGetLastError());
  return NMX_OK;
}

// C_q = alpha A[:, K_q] B[K_q, :] for the ksplit K slices K_q of length
// kchunk = round_up(ceil(k / ksplit), 16) (C_q at C + q csplit, no beta):
// one launch with ksplit times the workgroups, for products whose output
// alone would leave the device under-filled (n x 64 residuals).  The caller
// sums the C_q in a fixed order.
template <typename T>
int gemm_launch_splitk(nmx_ctx* ctx, int m, int n, int k, T alpha, const T* A, int lda, const T* B, int ldb, T* C,
                       int ldc, int ksplit, int64_t csplit) {
  if (m <= 0 || n <= 0 || k <= 0 || ksplit < 1) return NMX_ERR_UNSUPPORTED;
  const int kchunk = ceil_div(ceil_div(k, ksplit), 16) * 16;
  constexpr int TL = 64;
  const int tilesM = ceil_div(m, TL), tilesN = ceil_div(n, TL);
  const int per_slice = ceil_div(tilesM * tilesN, kXcds) * kXcds;
  ProfScope ps(ctx, K_GEMM, 2.0 * m * (double)n * k,
               (double)sizeof(T) * ((double)m * k + (double)k * n + (double)ksplit * m * n));
  gemm_kernel<T, 0, 0, TL><<<per_slice * ksplit, 256, 0, ctx->stream>>>(m, n, k, alpha, A, lda, B, ldb, T(0), C, ldc,
                                                                         tilesM, tilesN, 0, ksplit, kchunk, csplit);
  NMX_HIP(ctx, hipGetLastError());
  return NMX_OK;
}
template int gemm_launch_splitk<double>(nmx_ctx*, int, int, int, double, const double*, int, const double*, int,
                                        double*, int, int, int64_t);

template int gemm_launch<double>(nmx_ctx*, int, int, int, int, int, double, const double*, int, const double*, int,
                                 double, double*, int, int);
template int gemm_launch<float>(nmx_ctx*, int, int, int, int, int, float, const float*, int, const float*, int, float,
                                float*, int, int);

template <typename T>
static int gemm_api(nmx_ctx* ctx, int transa, int transb, int m, int n, int k, T alpha, const T* A, int lda,
                    const T* B, int ldb, T beta, T* C, int ldc) {
  if (!ctx) return -1;
  if (m < 0) return -4;
  if (n < 0) return -5;
  if (k < 0) return -6;
  if (lda < (transa ? m : k) || lda < 1) return -9;
  if (ldb < (transb ? k : n) || ldb < 1) return -11;
  if (ldc < n || ldc < 1) return -14;
  return gemm_launch<T>(ctx, transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, 0);
}

NMX_API int nmx_gemm_f64(nmx_ctx* ctx, int transa, int transb, int m, int n, int k, double alpha, const double* A,
                         int lda, const double* B, int ldb, double beta, double* C, int ldc) {
  return gemm_api<double>(ctx, transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

NMX_API int nmx_gemm_f32(nmx_ctx* ctx, int transa, int transb, int m, int n, int k, float alpha, const float* A,
                         int lda, const float* B, int ldb, float beta, float* C, int ldc) {
  return gemm_api<float>(ctx, transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}
