// Batched uneven-size solves: thousands of independent systems of order
// 1..256 in one call, FP32 factorization, FP64 residual and refinement.
//
// numax runs its small-matrix solvers one problem per GPU thread through
// `map_blocks` (numax/core/functional.mojo:691-834) over the register-tier
// algorithms -- pivoted LU (numax/linalg/_array/lu.mojo:165-216) and
// Cholesky (numax/linalg/_array/cholesky.mojo:26-101) -- with one compile-time
// n per launch and n <= 64 (numax/linalg/common.mojo:30-35).  Here one launch
// per size class takes every problem of that class from a queue sorted by
// size (largest first): a team of G*G wavefronts owns one problem at a time,
// G lanes per matrix row, each lane holding 64 consecutive columns of its row
// in registers:
//
//   class 0: 128 < n <= 256, G = 4, 1024 threads
//   class 1:  64 < n <= 128, G = 2,  256 threads
//   class 2:       n <=  64, G = 1,   64 threads (one wavefront, no barrier
//                                                 is ever waited on)
//
// The factorization is right-looking, one column per step, with every
// register index static: column k is read with a select tree on k, the
// pivot row is broadcast through LDS with its columns <= k zeroed, so the
// rank-1 update runs over all 64 registers and leaves column k's raw values
// in place; the multipliers (raw / pivot) are formed once, by the same
// division, when the factor is written out.  The finished factor goes to an
// FP32 scratch copy (L2-resident while the team works on it); the solves
// read it blockwise, 64 rows at a time; the residual b - A x is FP64 from
// the original matrix.
#include <limits.h>
#include <math.h>

#include "nmx_internal.h"

namespace {

constexpr int kBatchMaxN = 256;
constexpr int kSizeClasses = 3;
constexpr int kClasses = 2 * kSizeClasses;  // (LU, Cholesky) x size class
constexpr int kBins = 2 * (kBatchMaxN + 1);  // histogram bin = kind * 257 + n
constexpr float kFloorF = 1e-30f;  // numax/linalg/common.mojo:27-29

__host__ __device__ constexpr int class_of(int n) { return n > 128 ? 0 : (n > 64 ? 1 : 2); }

// guard_nonzero (numax/core/numeric.mojo:346-364): |x| raised to the floor,
// sign kept, +floor at +0, NaN passes through.
__device__ __forceinline__ float guard_f(float x) {
  const float m = fabsf(x);
  return (m >= kFloorF || m != m) ? x : copysignf(kFloorF, x);
}
__device__ __forceinline__ double guard_d(double x) {
  const double m = fabs(x);
  return (m >= 1e-30 || m != m) ? x : copysign(1e-30, x);
}

__device__ __forceinline__ float readlane_f(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

struct BatchArgs {
  int count;
  const int32_t* n;
  const int64_t* offA;
  const int64_t* offB;
  const uint8_t* kind;
  const double* A;
  const double* B;
  double* X;
  int32_t* info;
  int32_t* iters;
  double* berr;
  float* F;               // FP32 factor scratch, same offsets as A
  const int32_t* order;   // problem indices, largest n first
  const int32_t* cls;     // [0..3) class begin, [3..6) class count
  int32_t* head;          // per-class queue heads
  int max_iter;
  double tol;
};

template <int G>
struct Team {
  static constexpr int NMAX = 64 * G;
  static constexpr int T = 64 * G * G;
  static constexpr int NW = G * G;  // wavefronts
};

template <int G>
struct BatchShared {
  float P[Team<G>::NMAX];       // pivot row (LU), full
  float U[Team<G>::NMAX];       // pivot row with finished positions zeroed
  float K[Team<G>::NMAX];       // old row k (LU interchange)
  float col[2][Team<G>::NMAX];  // column k below the diagonal (Cholesky), double-buffered
  float dpiv[Team<G>::NMAX];    // guarded pivots (LU) / ljj (Cholesky)
  float pcur, kcur;
  float cmag[16];
  int crow[16];
  int perm[Team<G>::NMAX];
  float xs[Team<G>::NMAX];      // FP32 solve vector
  double x64[Team<G>::NMAX];
  double r64[Team<G>::NMAX];
  unsigned long long norms[4];  // max|r|, max|x|, ||A||_inf (bits of non-negative doubles)
  int slot, info;
};

// v[i] -= s * w[i] over the 64 registers, w in LDS; eight at a time with a
// scheduling barrier between groups, so at most eight LDS values are live
// (a fully hoisted 64-value prefetch would double the register footprint).
__device__ __forceinline__ void axpy_lds(float (&v)[64], float s, const float* w) {
#pragma unroll
  for (int c = 0; c < 64; c += 8) {
    const float4 a = *(const float4*)(w + c);
    const float4 b = *(const float4*)(w + c + 4);
    v[c + 0] -= s * a.x;
    v[c + 1] -= s * a.y;
    v[c + 2] -= s * a.z;
    v[c + 3] -= s * a.w;
    v[c + 4] -= s * b.x;
    v[c + 5] -= s * b.y;
    v[c + 6] -= s * b.z;
    v[c + 7] -= s * b.w;
    __builtin_amdgcn_sched_barrier(0);
  }
}

__device__ __forceinline__ void store_lds(const float (&v)[64], float* w) {
#pragma unroll
  for (int c = 0; c < 64; c += 4) *(float4*)(w + c) = make_float4(v[c], v[c + 1], v[c + 2], v[c + 3]);
}

__device__ __forceinline__ void load_lds(float (&v)[64], const float* w) {
#pragma unroll
  for (int c = 0; c < 64; c += 4) {
    const float4 a = *(const float4*)(w + c);
    v[c] = a.x;
    v[c + 1] = a.y;
    v[c + 2] = a.z;
    v[c + 3] = a.w;
    __builtin_amdgcn_sched_barrier(0);
  }
}

template <int G>
__device__ __forceinline__ float group_sum_f(float x) {
#pragma unroll
  for (int m = 1; m < G; m <<= 1) x += __shfl_xor(x, m);
  return x;
}
template <int G>
__device__ __forceinline__ double group_sum_d(double x) {
#pragma unroll
  for (int m = 1; m < G; m <<= 1) x += __shfl_xor(x, m);
  return x;
}

__device__ __forceinline__ bool better(float ma, int ra, float mb, int rb) {
  return ma > mb || (ma == mb && ra < rb);
}

// (m ? a : b) for a wave-uniform 64-bit mask, as an opaque v_cndmask: a
// plain ?: over two elements of v is folded into one load from a selected
// address, which would pin v in scratch memory.
__device__ __forceinline__ float sel_u(uint64_t m, float a, float b) {
  float r;
  asm("v_cndmask_b32_e64 %0, %1, %2, %3" : "=v"(r) : "v"(b), "v"(a), "s"(m));
  return r;
}

// v[k] for a uniform k in [0, 64): a six-level select tree (63 selects, no
// register-indexed access, nothing loop-carried).
__device__ __forceinline__ float pick64(const float (&v)[64], int k) {
  float a[32], b[16], c[8], d[4], e[2];
  const uint64_t m0 = (k & 1) ? ~0ull : 0ull, m1 = (k & 2) ? ~0ull : 0ull, m2 = (k & 4) ? ~0ull : 0ull;
  const uint64_t m3 = (k & 8) ? ~0ull : 0ull, m4 = (k & 16) ? ~0ull : 0ull, m5 = (k & 32) ? ~0ull : 0ull;
#pragma unroll
  for (int i = 0; i < 32; i++) a[i] = sel_u(m0, v[2 * i + 1], v[2 * i]);
#pragma unroll
  for (int i = 0; i < 16; i++) b[i] = sel_u(m1, a[2 * i + 1], a[2 * i]);
#pragma unroll
  for (int i = 0; i < 8; i++) c[i] = sel_u(m2, b[2 * i + 1], b[2 * i]);
#pragma unroll
  for (int i = 0; i < 4; i++) d[i] = sel_u(m3, c[2 * i + 1], c[2 * i]);
#pragma unroll
  for (int i = 0; i < 2; i++) e[i] = sel_u(m4, d[2 * i + 1], d[2 * i]);
  return sel_u(m5, e[1], e[0]);
}

// a[row][k] of the lane's row, for every lane of the row's group
template <int G>
__device__ __forceinline__ float column_value(const float (&v)[64], int k, int lane) {
  const float own = pick64(v, k & 63);
  if (G == 1) return own;
  return __shfl(own, (lane & ~(G - 1)) | (k >> 6));
}

// Pivoted LU of the team's problem (PivotedLU's arithmetic,
// _array/lu.mojo:165-216: first maximum of |a_ik| over i >= k with the
// diagonal as the incumbent, full-row interchange, pivot guard_nonzero(a_kk),
// multiplier a_ik / pivot).  The trailing update only touches columns > k,
// so a_ik stays in v and the multiplier is formed again, by the same
// division, when the factor is written out (S.dpiv holds the pivots).
// S.perm[i] ends as the original row now at row i; S.info the first k with
// an exactly-zero pivot.
template <int G>
__device__ __forceinline__ void lu_team(float (&v)[64], int n, BatchShared<G>& S, int t) {
  constexpr int NW = Team<G>::NW;
  const int lane = t & 63, wave = t >> 6, row = t / G, g = t % G;
  for (int k = 0; k < n; k++) {
    const float cur = column_value<G>(v, k, lane);
    float mag = -1.f;
    int crow = INT_MAX;
    if (g == 0 && row >= k && row < n) {
      const float a = fabsf(cur);
      // NaN never displaces the incumbent; a NaN incumbent is never displaced
      mag = (a == a) ? a : (row == k ? INFINITY : -1.f);
      crow = row;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const float om = __shfl_xor(mag, off);
      const int orr = __shfl_xor(crow, off);
      if (better(om, orr, mag, crow)) { mag = om; crow = orr; }
    }
    if (NW > 1) {
      if (lane == 0) { S.cmag[wave] = mag; S.crow[wave] = crow; }
      __syncthreads();
      mag = lane < NW ? S.cmag[lane] : -1.f;
      crow = lane < NW ? S.crow[lane] : INT_MAX;
#pragma unroll
      for (int off = 8; off > 0; off >>= 1) {
        const float om = __shfl_xor(mag, off);
        const int orr = __shfl_xor(crow, off);
        if (better(om, orr, mag, crow)) { mag = om; crow = orr; }
      }
    }
    const int pr = __builtin_amdgcn_readfirstlane(crow);
    if (row == pr) {
      store_lds(v, S.P + 64 * g);
      const int lim = k - 64 * g;  // positions <= lim are columns <= k
#pragma unroll
      for (int c = 0; c < 64; c += 4) {
        float4 u;
        u.x = (c + 0 > lim) ? v[c + 0] : 0.f;
        u.y = (c + 1 > lim) ? v[c + 1] : 0.f;
        u.z = (c + 2 > lim) ? v[c + 2] : 0.f;
        u.w = (c + 3 > lim) ? v[c + 3] : 0.f;
        *(float4*)(S.U + 64 * g + c) = u;
      }
      if (g == 0) S.pcur = cur;
    }
    if (pr != k && row == k) {
      store_lds(v, S.K + 64 * g);
      if (g == 0) S.kcur = cur;
    }
    __syncthreads();
    if (pr != k) {
      if (row == k) load_lds(v, S.P + 64 * g);
      else if (row == pr) load_lds(v, S.K + 64 * g);
    }
    const float praw = S.pcur;
    const float pivot = guard_f(praw);
    if (t == 0) {
      S.dpiv[k] = pivot;
      if (pr != k) {
        const int sw = S.perm[k];
        S.perm[k] = S.perm[pr];
        S.perm[pr] = sw;
      }
      if (praw == 0.f && S.info == 0) S.info = k + 1;
    }
    // the multiplier of each row below k from its column value after the
    // interchange: row pr now holds the old row k
    const bool below = row > k && row < n;
    const float a_ik = (row == pr && pr != k) ? S.kcur : cur;
    if (below) axpy_lds(v, a_ik / pivot, S.U + 64 * g);
  }
}

// Cholesky of the team's problem (lower triangle read): right-looking, the
// diagonal guard_nonzero'ed before the square root so a negative pivot gives
// NaN as _array/cholesky.mojo:26-73 does.  Column k's raw values stay in v
// (updates only touch columns > k); L = raw / ljj is formed at write-out
// (S.dpiv holds ljj).  S.info gets the first non-positive pivot (1-based).
template <int G>
__device__ __forceinline__ void chol_team(float (&v)[64], int n, BatchShared<G>& S, int t) {
  constexpr int NMAX = Team<G>::NMAX;
  const int lane = t & 63, row = t / G, g = t % G;
  for (int k = 0; k < n; k++) {
    float* col = S.col[k & 1];
    const float cur = column_value<G>(v, k, lane);
    if (g == 0 && row < NMAX) {
      col[row] = (row > k && row < n) ? cur : 0.f;
      if (row == k) S.pcur = cur;
    }
    __syncthreads();
    const float d = S.pcur;
    const float ljj = sqrtf(guard_f(d));
    if (t == 0) {
      S.dpiv[k] = ljj;
      if (!(d > 0.f) && S.info == 0) S.info = k + 1;
    }
    if (row > k && row < n) {
      const float sc = (cur / ljj) / ljj;
      const float* cv = col + 64 * g;
#pragma unroll
      for (int c = 0; c < 64; c += 8) {
#pragma unroll
        for (int i = c; i < c + 8; i++) v[i] -= sc * cv[i];
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
}

// op(F) x = xs in place (LDS), F the team's FP32 factor (row-major n x n).
// LOWER walks top-down, else bottom-up; TRANS reads F transposed (L^T).
// Diagonal blocks of 64 are solved by wavefront 0 with every index static;
// the other rows are updated blockwise by the whole team.
template <int G, bool LOWER, bool UNIT, bool TRANS>
__device__ __forceinline__ void tri_solve_team(const float* __restrict__ F, int n, BatchShared<G>& S, int t) {
  const int lane = t & 63, wave = t >> 6, row = t / G, g = t % G;
  const int nb = (n + 63) / 64;
  for (int bb = 0; bb < nb; bb++) {
    const int b = LOWER ? bb : nb - 1 - bb;
    const int r0 = 64 * b, bs = min(64, n - r0);
    if (wave == 0) {
      float a[64];
      const __amdgpu_buffer_rsrc_t rf = buf_rsrc(F + (int64_t)r0 * n + r0);
#pragma unroll
      for (int c = 0; c < 64; c++) {
        // op(F)[r0 + lane][r0 + c]
        const uint32_t off = TRANS ? (uint32_t)(c * n + lane) : (uint32_t)(lane * n + c);
        a[c] = buf_load<float>(rf, ((lane < bs) & (c < bs)) ? 4u * off : kBufOob);
      }
      float acc = lane < bs ? S.xs[r0 + lane] : 0.f;
      // an opaque copy of the lane index: otherwise the 128 lane comparisons
      // below are hoisted out of the block loop and spilled
      int ln = lane;
      asm volatile("" : "+v"(ln));
      if (LOWER) {
#pragma unroll
        for (int c = 0; c < 64; c++) {
          if (c < bs) {
            float xc = readlane_f(acc, c);
            if (!UNIT) xc = xc / guard_f(readlane_f(a[c], c));
            if (ln == c) acc = xc;
            else if (ln > c) acc -= a[c] * xc;
          }
        }
      } else {
#pragma unroll
        for (int c = 63; c >= 0; c--) {
          if (c < bs) {
            float xc = readlane_f(acc, c);
            if (!UNIT) xc = xc / guard_f(readlane_f(a[c], c));
            if (ln == c) acc = xc;
            else if (ln < c) acc -= a[c] * xc;
          }
        }
      }
      if (lane < bs) S.xs[r0 + lane] = acc;
    }
    __syncthreads();
    const bool todo = row < n && (LOWER ? row >= r0 + bs : row < r0);
    float part = 0.f;
    if (todo) {
      constexpr int CW = 64 / G;  // block columns per lane
      const __amdgpu_buffer_rsrc_t rf = buf_rsrc(F);
#pragma unroll
      for (int i0 = 0; i0 < CW; i0 += 8) {
#pragma unroll
        for (int i = i0; i < i0 + 8; i++) {
          const int c = g * CW + i;
          const uint32_t off = TRANS ? (uint32_t)((r0 + c) * n + row) : (uint32_t)(row * n + r0 + c);
          part += buf_load<float>(rf, c < bs ? 4u * off : kBufOob) * S.xs[r0 + (c < bs ? c : 0)];
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    part = group_sum_f<G>(part);
    if (todo && g == 0) S.xs[row] -= part;
    __syncthreads();
  }
}

// x64 += (or =) op^-1 of the current xs (the FP32 correction)
template <int G, bool CHOL>
__device__ __forceinline__ void solve_team(const float* __restrict__ F, int n, BatchShared<G>& S, int t) {
  if (CHOL) {
    tri_solve_team<G, true, false, false>(F, n, S, t);
    tri_solve_team<G, false, false, true>(F, n, S, t);
  } else {
    tri_solve_team<G, true, true, false>(F, n, S, t);
    tri_solve_team<G, false, false, false>(F, n, S, t);
  }
}

// r64 = b - A x64 (FP64), and the backward-error norms.  A Cholesky problem
// is the symmetric matrix its lower triangle defines.
template <int G, bool CHOL>
__device__ __forceinline__ double residual_team(const double* __restrict__ A, const double* __restrict__ b, int n,
                                bool first, double& anorm, BatchShared<G>& S, int t) {
  const int row = t / G, g = t % G;
  if (t < 4) S.norms[t] = 0ull;
  __syncthreads();
  double part = 0.0, apart = 0.0;
  if (row < n) {
    const __amdgpu_buffer_rsrc_t ra = buf_rsrc(A);
    const uint32_t rowoff = (uint32_t)row * (uint32_t)n;
#pragma unroll 16
    for (int i = 0; i < 64; i++) {
      const int c = 64 * g + i;
      double a = buf_load<double>(ra, c < n ? 8u * (rowoff + (uint32_t)c) : kBufOob);
      if (CHOL && c > row && c < n) a = A[(int64_t)c * n + row];
      part += a * (c < n ? S.x64[c] : 0.0);
      apart += fabs(a);
    }
  }
  part = group_sum_d<G>(part);
  apart = group_sum_d<G>(apart);
  if (row < n && g == 0) {
    const double r = b[row] - part;
    S.r64[row] = r;
    atomicMax(&S.norms[0], (unsigned long long)__double_as_longlong(fabs(r)));
    atomicMax(&S.norms[1], (unsigned long long)__double_as_longlong(fabs(S.x64[row])));
    if (first) atomicMax(&S.norms[2], (unsigned long long)__double_as_longlong(apart));
  }
  __syncthreads();
  if (first) anorm = __longlong_as_double((long long)S.norms[2]);
  const double rmax = __longlong_as_double((long long)S.norms[0]);
  const double xmax = __longlong_as_double((long long)S.norms[1]);
  const double den = anorm * xmax;
  return den > 0.0 ? rmax / den : (rmax == 0.0 ? 0.0 : INFINITY);
}

template <int G, bool CHOL>
__global__ void __launch_bounds__(Team<G>::T) batched_solve_kernel(BatchArgs args, int cls) {
  __shared__ BatchShared<G> S;
  const int t = threadIdx.x;
  const int row = t / G;
  const int begin = args.cls[cls], count = args.cls[kClasses + cls];
  for (;;) {
    if (t == 0) S.slot = atomicAdd(&args.head[cls], 1);
    __syncthreads();
    const int slot = S.slot;
    if (slot >= count) break;
    const int p = args.order[begin + slot];
    const int n = args.n[p];
    const double* A = args.A + args.offA[p];
    const double* b = args.B + args.offB[p];
    float* F = args.F + args.offA[p];
    if (t == 0) S.info = 0;
    if (row < Team<G>::NMAX && (t % G) == 0) S.perm[row] = row;

    // load (FP64 -> FP32 registers); a Cholesky problem reads its lower triangle.
    // rel = row - 64 g, made opaque per problem: otherwise the 2 x 64
    // comparisons against it are hoisted out of the problem loop and spilled.
    int rel = row - 64 * (t % G);
    asm volatile("" : "+v"(rel));
    float v[64];
    {
      const int g = t % G;
      // one descriptor per problem (uniform); the byte offsets stay < 512 KiB
      const __amdgpu_buffer_rsrc_t ra = buf_rsrc(A);
      const uint32_t rowoff = (uint32_t)(row < n ? row : 0) * (uint32_t)n;
      // sixteen loads in flight at a time: all 64 FP64 loads at once would
      // need 128 registers on top of the 64 they land in
#pragma unroll
      for (int i0 = 0; i0 < 64; i0 += 16) {
#pragma unroll
        for (int i = i0; i < i0 + 16; i++) {
          const int c = 64 * g + i;
          const bool ok = (row < n) & (c < n) & (!CHOL | (i <= rel));
          v[i] = (float)buf_load<double>(ra, ok ? 8u * (rowoff + (uint32_t)c) : kBufOob);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    __syncthreads();
    if (CHOL)
      chol_team<G>(v, n, S, t);
    else
      lu_team<G>(v, n, S, t);
    __syncthreads();  // the last column's pivot (S.dpiv) is read below
    // the FP32 factor: multipliers (raw / pivot) below the diagonal, U (LU)
    // or ljj on it (Cholesky), zero above it for Cholesky
    if (row < n) {
      const int g = t % G;
#pragma unroll
      for (int i0 = 0; i0 < 64; i0 += 8) {
#pragma unroll
        for (int i = i0; i < i0 + 8; i++) {
          const int c = 64 * g + i;
          float f = v[i];
          if (i < rel) f = f / S.dpiv[c];
          else if (CHOL) f = (i == rel) ? S.dpiv[c] : 0.f;
          if (c < n) F[(int64_t)row * n + c] = f;
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    __syncthreads();

    // x0 = op^-1 P b in FP32, then FP64 refinement; one solve site (the
    // triangular solves are large once unrolled)
    double anorm = 0.0, eta = 0.0, prev = INFINITY;
    int it = 0;
    for (bool first = true;; first = false) {
      if (row < n && (t % G) == 0) {
        const int src = CHOL ? row : S.perm[row];
        S.xs[row] = first ? (float)b[src] : (float)S.r64[src];
      }
      __syncthreads();
      solve_team<G, CHOL>(F, n, S, t);
      if (row < n && (t % G) == 0) S.x64[row] = first ? (double)S.xs[row] : S.x64[row] + (double)S.xs[row];
      __syncthreads();
      eta = residual_team<G, CHOL>(A, b, n, first, anorm, S, t);
      if (!first) it++;
      if (!(eta > args.tol) || it >= args.max_iter || !(eta < 0.5 * prev)) break;
      prev = eta;
    }
    if (row < n && (t % G) == 0) args.X[args.offB[p] + row] = S.x64[row];
    if (t == 0) {
      args.info[p] = S.info;
      args.iters[p] = it;
      args.berr[p] = eta;
    }
    __syncthreads();  // S is reused by the next problem
  }
}

// ---- size-class queue: counting sort by n, largest first ---------------------
// A problem is taken when 1 <= n <= 256, kind is 0 or 1 and its matrix and
// vectors lie inside the arrays the caller sized (so no kernel can touch
// memory outside them); any other is reported with info = -1.
__device__ __forceinline__ bool batch_ok(int nn, int kd, int64_t oa, int64_t ob, int64_t a_elems, int64_t b_elems) {
  return nn >= 1 && nn <= kBatchMaxN && kd <= 1 && oa >= 0 && oa + (int64_t)nn * nn <= a_elems && ob >= 0 &&
         ob + nn <= b_elems;
}

__global__ void batch_hist_kernel(int count, const int32_t* __restrict__ n, const uint8_t* __restrict__ kind,
                                  const int64_t* __restrict__ offA, const int64_t* __restrict__ offB, int64_t a_elems,
                                  int64_t b_elems, int32_t* __restrict__ hist, int32_t* __restrict__ info,
                                  double* __restrict__ berr, int32_t* __restrict__ iters) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int nn = n[i];
  const bool ok = batch_ok(nn, kind[i], offA[i], offB[i], a_elems, b_elems);
  if (ok) {
    atomicAdd(&hist[kind[i] * (kBatchMaxN + 1) + nn], 1);
  } else {
    info[i] = -1;  // not a problem this kernel takes: reported, not solved
    iters[i] = 0;
    berr[i] = NAN;
  }
}

__global__ void batch_scan_kernel(const int32_t* __restrict__ hist, int32_t* __restrict__ cursor,
                                  int32_t* __restrict__ cls, int32_t* __restrict__ head) {
  if (threadIdx.x != 0) return;
  int run = 0;
  int cbeg[kClasses], ccnt[kClasses];
  for (int c = 0; c < kClasses; c++) cbeg[c] = ccnt[c] = 0;
  for (int kd = 0; kd < 2; kd++) {
    for (int nn = kBatchMaxN; nn >= 1; nn--) {
      const int c = kd * kSizeClasses + class_of(nn);
      const int bin = kd * (kBatchMaxN + 1) + nn;
      if (ccnt[c] == 0) cbeg[c] = run;
      cursor[bin] = run;
      ccnt[c] += hist[bin];
      run += hist[bin];
    }
  }
  for (int c = 0; c < kClasses; c++) {
    cls[c] = cbeg[c];
    cls[kClasses + c] = ccnt[c];
    head[c] = 0;
  }
}

__global__ void batch_scatter_kernel(int count, const int32_t* __restrict__ n, const uint8_t* __restrict__ kind,
                                     const int64_t* __restrict__ offA, const int64_t* __restrict__ offB,
                                     int64_t a_elems, int64_t b_elems, int32_t* __restrict__ cursor,
                                     int32_t* __restrict__ order) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int nn = n[i];
  if (batch_ok(nn, kind[i], offA[i], offB[i], a_elems, b_elems))
    order[atomicAdd(&cursor[kind[i] * (kBatchMaxN + 1) + nn], 1)] = i;
}

__global__ void zero_i32_kernel(int32_t* p, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0;
}

template <int G, bool CHOL>
int launch_class(nmx_ctx* ctx, const BatchArgs& a, int cls, int per_cu) {
  const int grid = ctx->num_cus * per_cu;
  batched_solve_kernel<G, CHOL><<<grid, Team<G>::T, 0, ctx->stream>>>(a, cls);
  NMX_HIP(ctx, hipGetLastError());
  return NMX_OK;
}

}  // namespace

NMX_API void nmx_batch_cfg_default(nmx_batch_cfg* cfg) {
  if (!cfg) return;
  cfg->max_iter = 10;
  cfg->tol = 1e-15;
  cfg->teams_per_cu[0] = 1;
  cfg->teams_per_cu[1] = 4;
  cfg->teams_per_cu[2] = 16;
}

NMX_API int nmx_batched_gesv(nmx_ctx* ctx, int count, const int32_t* n, const int64_t* offA, const int64_t* offB,
                             const uint8_t* kind, const double* A, int64_t a_elems, const double* B, int64_t b_elems,
                             double* X, int32_t* info, int32_t* iters, double* berr, const nmx_batch_cfg* cfg_in) {
  if (!ctx) return -1;
  if (count < 0) return -2;
  if (count == 0) return NMX_OK;
  if (!n) return -3;
  if (!offA) return -4;
  if (!offB) return -5;
  if (!kind) return -6;
  if (!A) return -7;
  if (a_elems <= 0) return -8;
  if (!B) return -9;
  if (b_elems <= 0) return -10;
  if (!X) return -11;
  if (!info) return -12;
  if (!iters) return -13;
  if (!berr) return -14;
  nmx_batch_cfg cfg;
  if (cfg_in) cfg = *cfg_in; else nmx_batch_cfg_default(&cfg);
  if (cfg.max_iter < 0 || !(cfg.tol >= 0.0)) return -15;
  for (int c = 0; c < kSizeClasses; c++)
    if (cfg.teams_per_cu[c] <= 0 || cfg.teams_per_cu[c] > 64) return -15;

  // workspace: FP32 factors, then the queue
  const size_t f_bytes = ((sizeof(float) * (size_t)a_elems + 255) / 256) * 256;
  const size_t q_ints = 2 * (size_t)kBins + 2 * kClasses + kClasses + (size_t)count;
  void* ws = nullptr;
  NMX_TRY(nmx_workspace(ctx, f_bytes + sizeof(int32_t) * q_ints, &ws));
  float* F = (float*)ws;
  int32_t* hist = (int32_t*)((char*)ws + f_bytes);
  int32_t* cursor = hist + kBins;
  int32_t* cls = cursor + kBins;
  int32_t* head = cls + 2 * kClasses;
  int32_t* order = head + kClasses;

  {
    ProfScope ps(ctx, K_OTHER, 0.0, 0.0);
    zero_i32_kernel<<<ceil_div(kBins, 256), 256, 0, ctx->stream>>>(hist, kBins);
    batch_hist_kernel<<<ceil_div(count, 256), 256, 0, ctx->stream>>>(count, n, kind, offA, offB, a_elems, b_elems,
                                                                       hist, info, berr, iters);
    batch_scan_kernel<<<1, 64, 0, ctx->stream>>>(hist, cursor, cls, head);
    batch_scatter_kernel<<<ceil_div(count, 256), 256, 0, ctx->stream>>>(count, n, kind, offA, offB, a_elems,
                                                                          b_elems, cursor, order);
    NMX_HIP(ctx, hipGetLastError());
  }
  BatchArgs a{count, n, offA, offB, kind, A, B, X, info, iters, berr, F, order, cls, head, cfg.max_iter, cfg.tol};
  {
    ProfScope ps(ctx, K_BATCH, 0.0, 8.0 * (double)a_elems);
    // largest problems first; LU then Cholesky within a size class
    NMX_TRY((launch_class<4, false>(ctx, a, 0, cfg.teams_per_cu[0])));
    NMX_TRY((launch_class<4, true>(ctx, a, kSizeClasses + 0, cfg.teams_per_cu[0])));
    NMX_TRY((launch_class<2, false>(ctx, a, 1, cfg.teams_per_cu[1])));
    NMX_TRY((launch_class<2, true>(ctx, a, kSizeClasses + 1, cfg.teams_per_cu[1])));
    NMX_TRY((launch_class<1, false>(ctx, a, 2, cfg.teams_per_cu[2])));
    NMX_TRY((launch_class<1, true>(ctx, a, kSizeClasses + 2, cfg.teams_per_cu[2])));
  }
  return NMX_OK;
}
