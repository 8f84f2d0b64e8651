// Batched uneven-size solves: thousands of independent systems of order
// 1..256 in one call, FP32 factorization, FP64 residual and refinement.
//
// numax runs its small-matrix solvers one problem per GPU thread through
// `map_blocks` (numax/core/functional.mojo:691-834) over the register-tier
// algorithms -- pivoted LU (numax/linalg/_array/lu.mojo:165-216) and
// Cholesky (numax/linalg/_array/cholesky.mojo:26-101) -- with one compile-time
// n per launch and n <= 64 (numax/linalg/common.mojo:30-35).  Here one
// wavefront owns one problem at a time and factors it blocked, right-looking,
// 32 columns per panel:
//
//   - the panel lives in registers (lane l holds rows k + l + 64 q), its
//     pivot search is a DPP key reduction, its pivot row is read lane-to-
//     scalar, so a column costs no barrier and no LDS round trip but the
//     interchange;
//   - the panel's interchanges reach the other columns as one gather of the
//     net row moves;
//   - U12 = L11^-1 A12 (LU) or L21 = A21 L11^-T (Cholesky) is one
//     triangular solve per lane;
//   - the trailing update A22 -= L21 U12 (or the lower triangle of
//     A22 -= L21 L21^T) is 32 x 32 tiles of v_mfma_f32_32x32x2f32.
//
// The FP32 working copy is a scratch matrix per problem (L2-resident while
// its wavefront works on it); the solves read it 64 rows at a time; the
// residual b - A x is FP64 from the original matrix.  Problems are sorted by
// kind and size class (rows per lane Q = 1..4) on the device and handed out
// largest first from per-class queues: the kernels are persistent and need
// no host round trip.
#include <limits.h>
#include <math.h>

#include <type_traits>
#include <vector>

#include "nmx_internal.h"

namespace {

constexpr int kBatchMaxN = 256;
constexpr int kSizeClasses = 4;             // n in (64 (c-1), 64 c]: Q = c rows per lane
constexpr int kClasses = 2 * kSizeClasses;  // (LU, Cholesky) x size class
constexpr int kBins = 2 * (kBatchMaxN + 1);  // histogram bin = kind * 257 + n
constexpr int kNB = 32;                     // panel width
constexpr int kWaves = 4;                   // independent wavefronts per workgroup
constexpr float kFloorF = 1e-30f;           // numax/linalg/common.mojo:27-29

__host__ __device__ constexpr int class_of(int n) { return (n - 1) >> 6; }

typedef float floatx16 __attribute__((ext_vector_type(16)));

// guard_nonzero (numax/core/numeric.mojo:346-364): |x| raised to the floor,
// sign kept, +floor at +0, NaN passes through.
__device__ __forceinline__ float guard_f(float x) {
  const float m = fabsf(x);
  return (m >= kFloorF || m != m) ? x : copysignf(kFloorF, x);
}

// Orders a wavefront's LDS and global accesses around cross-lane hand-offs
// (no other wavefront shares the data, so no barrier is needed or allowed:
// the wavefronts of a workgroup work on different problems).
__device__ __forceinline__ void wave_sync() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup"); }

// max over the wavefront of a non-negative double, as its bit pattern
__device__ __forceinline__ double wave_dmax(double x) {
  const uint64_t b = (uint64_t)__double_as_longlong(x);
  const uint32_t hi = wave_umax((uint32_t)(b >> 32));
  const uint32_t lo = wave_umax((uint32_t)(b >> 32) == hi ? (uint32_t)b : 0u);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
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
  float* F;              // FP32 working copies
  const int64_t* offF;   // problem p's copy at F + offF[p] (a prefix sum of n^2)
  const int32_t* order;  // problem indices, by class, largest n first
  const int32_t* cls;    // [0..8) class begin, [8..16) class count
  int32_t* head;         // per-class queue heads
  int max_iter;
  double tol;
};

// A wavefront's LDS, sized for its class's largest order NMAX = 64 Q
template <int NMAX>
struct WaveLds {
  float xs[NMAX];         // FP32 solve vector; the residual is written here in solve order
  double x64[NMAX];       // solution
  int perm[NMAX];         // original row now at row i (LU)
  int ipiv[NMAX];         // interchanges, LAPACK order
  float rowbuf[2][kNB];   // interchange staging
  int mv_dst[2 * kNB], mv_src[2 * kNB];
};
// the rolled solves (n <= 128) stage each 64 x 64 diagonal block in LDS;
// at 20 KB or less a wavefront, two workgroups of four fit a CU's 160 KB
template <int NMAX>
struct WaveLdsBlk : WaveLds<NMAX> {
  float blk[64][65];
};
static_assert(sizeof(WaveLdsBlk<128>) * kWaves <= 160 * 1024 / 2, "two workgroups a CU");
// the unrolled classes (n > 128) need LDS only for the staging tile
constexpr int kTileLd = 33;  // staging tile row stride (words): rows and columns both conflict-free
template <int NMAX>
struct WaveLdsTile : WaveLds<NMAX> {
  float tt[32 * kTileLd];
};
// a 32 x 33 staging tile: its own array, or the start of the solve block
template <int NMAX>
__device__ __forceinline__ float* tile_of(WaveLdsBlk<NMAX>& S) { return &S.blk[0][0]; }
template <int NMAX>
__device__ __forceinline__ float* tile_of(WaveLdsTile<NMAX>& S) { return S.tt; }
static_assert(sizeof(WaveLdsBlk<64>::blk) >= 32 * kTileLd * sizeof(float), "staging tile fits the solve block");

// Orders a wavefront's accesses to its own staging tile: the LDS executes a
// wavefront's instructions in order, so waiting for its LDS counter (and
// keeping the compiler from moving accesses across) is enough; unlike
// wave_sync it does not wait for global loads and stores in flight.
__device__ __forceinline__ void lds_sync() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

// Every row held in registers is rotated left by one column after each
// step of a rolled column loop, so the current column is always register 0
// and all indices stay static: fully unrolled, these loops would be hundreds
// of KB of straight-line code that each wavefront streams through the
// instruction cache once per panel.
template <int N>
__device__ __forceinline__ void rot1(float (&v)[N]) {
  const float f = v[0];
#pragma unroll
  for (int c = 0; c + 1 < N; c++) v[c] = v[c + 1];
  v[N - 1] = f;
}

// ---- panel: pivoted LU of F[k:n, k:k+nb] in registers ------------------------
// PivotedLU's arithmetic (_array/lu.mojo:165-216): first maximum of |a_ik|
// with the diagonal as the incumbent (a NaN incumbent is kept, a NaN never
// displaces one), full-row interchange, pivot guard_nonzero(a_kk),
// multipliers by true division.
template <int Q, class L>
__device__ __forceinline__ void panel_lu(float* __restrict__ F, int n, int k, int nb, L& S, int lane,
                                         int& info) {
  const __amdgpu_buffer_rsrc_t rf = buf_rsrc(F);
  float pv[Q][kNB];
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const int row = k + lane + 64 * q;
    const uint32_t ro = (uint32_t)(row < n ? row : 0) * (uint32_t)n + (uint32_t)k;
#pragma unroll
    for (int c = 0; c < kNB; c++) pv[q][c] = buf_load<float>(rf, ((row < n) & (c < nb)) ? 4u * (ro + c) : kBufOob);
  }
#pragma unroll 1
  for (int j = 0; j < nb; j++) {
    const int kj = k + j;
    uint32_t key = 0;
    int crow = INT_MAX;
#pragma unroll
    for (int q = 0; q < Q; q++) {
      const int row = k + lane + 64 * q;
      if (row >= kj && row < n) {
        const float a = fabsf(pv[q][0]);
        const uint32_t kk = (a == a) ? __float_as_uint(a) + 1u : (row == kj ? 0xFFFFFFFFu : 0u);
        if (kk > key) { key = kk; crow = row; }  // rows ascend with q: ties keep the first
      }
    }
    const uint32_t K = wave_umax(key);
    const int pr = (int)wave_umin(key == K ? (uint32_t)crow : 0xFFFFFFFFu);
    if (pr != kj) {
      // row kj sits in lane j (slot 0), row pr in lane lp, slot qp
      const int lp = (pr - k) & 63, qp = (pr - k) >> 6;
      if (lane == j) {
#pragma unroll
        for (int c = 0; c < kNB; c++) S.rowbuf[0][c] = pv[0][c];
      }
      // one block per slot, each marked distinct: merged into one block the
      // accesses would become pv[qp][c], a dynamic index
#pragma unroll
      for (int q = 0; q < Q; q++)
        if (q == qp && lane == lp) {
          asm volatile("; slot %0" ::"n"(q));
#pragma unroll
          for (int c = 0; c < kNB; c++) S.rowbuf[1][c] = pv[q][c];
        }
      wave_sync();
      if (lane == j) {
#pragma unroll
        for (int c = 0; c < kNB; c++) pv[0][c] = S.rowbuf[1][c];
      }
#pragma unroll
      for (int q = 0; q < Q; q++)
        if (q == qp && lane == lp) {
          asm volatile("; slot %0" ::"n"(q));
#pragma unroll
          for (int c = 0; c < kNB; c++) pv[q][c] = S.rowbuf[0][c];
        }
      if (lane == 0) {
        const int sw = S.perm[kj];
        S.perm[kj] = S.perm[pr];
        S.perm[pr] = sw;
      }
      wave_sync();
    }
    if (lane == 0) S.ipiv[kj] = pr;
    // the pivot row (lane j) into scalar registers; positions >= nb - j are
    // finished columns rotated to the end and must not be updated
    const float praw = readlane_f(pv[0][0], j);
    if (praw == 0.f && info == 0) info = kj + 1;
    const float pivot = guard_f(praw);
    const int rem = nb - j;
    float u[kNB];
#pragma unroll
    for (int c = 1; c < kNB; c++) u[c] = c < rem ? readlane_f(pv[0][c], j) : 0.f;
#pragma unroll
    for (int q = 0; q < Q; q++) {
      const int row = k + lane + 64 * q;
      if (row > kj && row < n) {
        const float l = pv[q][0] / pivot;
        pv[q][0] = l;
#pragma unroll
        for (int c = 1; c < kNB; c++) pv[q][c] -= l * u[c];
      }
      rot1(pv[q]);
    }
  }
#pragma unroll 1
  for (int j = nb; j < kNB; j++) {
#pragma unroll
    for (int q = 0; q < Q; q++) rot1(pv[q]);
  }
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const int row = k + lane + 64 * q;
    if (row < n) {
#pragma unroll
      for (int c = 0; c < kNB; c++)
        if (c < nb) F[(int64_t)row * n + k + c] = pv[q][c];
    }
  }
}

// The panel's interchanges ipiv[k..k+nb) over the columns outside it: the
// net row moves (<= 2 nb rows) are worked out once by the wavefront, then
// each lane moves whole columns (all loads before the stores).
template <int Q, class L>
__device__ __forceinline__ void swap_outside(float* __restrict__ F, int n, int k, int nb, L& S, int lane) {
  int myrow = lane < nb ? k + lane : -1;
  int mine = myrow;
  int nr = nb;
  const int qv = lane < nb ? S.ipiv[k + lane] : 0;  // the pivots once, one per lane
  for (int j = 0; j < nb; j++) {
    const int qrow = __builtin_amdgcn_readlane(qv, j);
    const unsigned long long hit = __ballot(lane < nr && myrow == qrow);
    int pos;
    if (hit) {
      pos = __ffsll((long long)hit) - 1;
    } else {
      pos = nr;
      if (lane == nr) { myrow = qrow; mine = qrow; }
      nr++;
    }
    const int cj = __shfl(mine, j);
    const int cp = __shfl(mine, pos);
    if (lane == j) mine = cp;
    if (lane == pos) mine = cj;
  }
  const bool moved = lane < nr && mine != myrow;
  const unsigned long long mm = __ballot(moved);
  if (moved) {
    const int at = __popcll(mm & ((1ull << lane) - 1ull));
    S.mv_dst[at] = myrow;
    S.mv_src[at] = mine;
  }
  const int moves = __popcll(mm);
  wave_sync();
  if (moves == 0) return;
  // lane e holds move e; each move's rows reach every lane by v_readlane (an
  // LDS read per move put one round trip ahead of each load).  All lanes stay
  // active: a column outside the range loads column 0 and stores nothing.
  static_assert(2 * kNB == 64, "one move per lane");
  const int my_src = S.mv_src[lane < moves ? lane : moves - 1];
  const int my_dst = S.mv_dst[lane < moves ? lane : moves - 1];
#pragma unroll 1
  for (int s = 0; s < Q; s++) {
    const int c = lane + 64 * s;
    const bool ok = c < n && (c < k || c >= k + nb);
    const int cl = ok ? c : 0;
    float vals[2 * kNB];
#pragma unroll
    for (int e = 0; e < 2 * kNB; e++) vals[e] = F[(int64_t)__builtin_amdgcn_readlane(my_src, e) * n + cl];
#pragma unroll
    for (int e = 0; e < 2 * kNB; e++)
      if (ok && e < moves) F[(int64_t)__builtin_amdgcn_readlane(my_dst, e) * n + cl] = vals[e];
  }
  wave_sync();
}

// U12 := L11^-1 A12 over columns k+nb..n, right-looking over the rows of
// L11 with the lane's columns (x) and L11's rows (lane i < nb holds row i)
// both rotated, so L[jj + c][jj] is lane jj + c's register 0.
template <int Q>
__device__ __forceinline__ void trsm_u12(float* __restrict__ F, int n, int k, int nb, float* __restrict__ tt,
                                         int lane) {
  const __amdgpu_buffer_rsrc_t rf = buf_rsrc(F);
  float l11[kNB];
  {
    const uint32_t ro = (uint32_t)(k + (lane < nb ? lane : 0)) * (uint32_t)n + (uint32_t)k;
#pragma unroll
    for (int c = 0; c < kNB; c++) l11[c] = buf_load<float>(rf, ((lane < nb) & (c < nb)) ? 4u * (ro + c) : kBufOob);
  }
  float x[Q][kNB];
#pragma unroll
  for (int s = 0; s < Q; s++) {
    const int c = k + nb + lane + 64 * s;
#pragma unroll
    for (int i = 0; i < kNB; i++)
      x[s][i] = buf_load<float>(rf, ((c < n) & (i < nb)) ? 4u * ((uint32_t)(k + i) * (uint32_t)n + (uint32_t)c) : kBufOob);
  }
#pragma unroll 1
  for (int jj = 0; jj < nb; jj++) {
    const int rem = nb - jj;
#pragma unroll
    for (int c = 1; c < kNB; c++) {
      const float l = c < rem ? readlane_f(l11[0], jj + c) : 0.f;
#pragma unroll
      for (int s = 0; s < Q; s++) x[s][c] -= l * x[s][0];
    }
#pragma unroll
    for (int s = 0; s < Q; s++) rot1(x[s]);
    rot1(l11);
  }
#pragma unroll 1
  for (int jj = nb; jj < kNB; jj++) {
#pragma unroll
    for (int s = 0; s < Q; s++) rot1(x[s]);
  }
#pragma unroll
  for (int s = 0; s < Q; s++) {
    const int c = k + nb + lane + 64 * s;
    if (c < n) {
#pragma unroll
      for (int i = 0; i < kNB; i++)
        if (i < nb) F[(int64_t)(k + i) * n + c] = x[s][i];
    }
  }
  wave_sync();
}

// ---- Cholesky panel: the diagonal block in lanes 0..nb, then L21 ------------
// _array/cholesky.mojo:26-73's arithmetic per column: the diagonal
// guard_nonzero'ed before the square root (a negative pivot gives NaN), the
// column divided by ljj; right-looking here, rows rotated as in the LU panel.
template <int Q, class L>
__device__ __forceinline__ void panel_chol(float* __restrict__ F, int n, int k, int nb, L& S, int lane,
                                           int& info) {
  const __amdgpu_buffer_rsrc_t rf = buf_rsrc(F);
  float dv[kNB];
  {
    const bool own = lane < nb;
    const uint32_t ro = (uint32_t)(k + (own ? lane : 0)) * (uint32_t)n + (uint32_t)k;
#pragma unroll
    for (int c = 0; c < kNB; c++) dv[c] = buf_load<float>(rf, (own & (c < nb)) ? 4u * (ro + c) : kBufOob);
  }
#pragma unroll 1
  for (int j = 0; j < nb; j++) {
    const float d = readlane_f(dv[0], j);
    if (!(d > 0.f) && info == 0) info = k + j + 1;
    const float ljj = sqrtf(guard_f(d));
    if (lane == j) dv[0] = ljj;
    else if (lane > j) dv[0] = dv[0] / ljj;
    const int rem = nb - j;
#pragma unroll
    for (int c = 1; c < kNB; c++) {
      const float lc = c < rem ? readlane_f(dv[0], j + c) : 0.f;
      if (lane > j) dv[c] -= dv[0] * lc;
    }
    rot1(dv);
  }
#pragma unroll 1
  for (int j = nb; j < kNB; j++) rot1(dv);
  if (lane < nb) {
#pragma unroll
    for (int c = 0; c < kNB; c++)
      if (c < nb) F[(int64_t)(k + lane) * n + k + c] = c <= lane ? dv[c] : 0.f;
  }
  // L21 := A21 L11^-T, right-looking over the columns of L11 with x and
  // L11's rows rotated: L[jj + c][jj] is lane jj + c's register 0
  float x[Q][kNB];
#pragma unroll
  for (int s = 0; s < Q; s++) {
    const int row = k + nb + lane + 64 * s;
#pragma unroll
    for (int c = 0; c < kNB; c++)
      x[s][c] = buf_load<float>(rf, ((row < n) & (c < nb)) ? 4u * ((uint32_t)row * (uint32_t)n + (uint32_t)(k + c))
                                                          : kBufOob);
  }
#pragma unroll 1
  for (int jj = 0; jj < nb; jj++) {
    const float djj = readlane_f(dv[0], jj);
#pragma unroll
    for (int s = 0; s < Q; s++) x[s][0] = x[s][0] / djj;
    const int rem = nb - jj;
#pragma unroll
    for (int c = 1; c < kNB; c++) {
      const float l = c < rem ? readlane_f(dv[0], jj + c) : 0.f;
#pragma unroll
      for (int s = 0; s < Q; s++) x[s][c] -= x[s][0] * l;
    }
#pragma unroll
    for (int s = 0; s < Q; s++) rot1(x[s]);
    rot1(dv);
  }
#pragma unroll 1
  for (int jj = nb; jj < kNB; jj++) {
#pragma unroll
    for (int s = 0; s < Q; s++) rot1(x[s]);
  }
#pragma unroll
  for (int s = 0; s < Q; s++) {
    const int row = k + nb + lane + 64 * s;
    if (row < n) {
#pragma unroll
      for (int c = 0; c < kNB; c++)
        if (c < nb) F[(int64_t)row * n + k + c] = x[s][c];
    }
  }
  wave_sync();
}

// ---- fully unrolled panel, U12 and solve variants (n > 128) -------------------
// Every column index static, the pivot row and L11 exchanged through scalar
// reads; the classes with n > 128 run these.
// PivotedLU's arithmetic (_array/lu.mojo:165-216): first maximum of |a_ik|
// with the diagonal as the incumbent (a NaN incumbent is kept, a NaN never
// displaces one), full-row interchange, pivot guard_nonzero(a_kk),
// multipliers by true division.  Column J of the panel is a template
// instance, so every register index is static however large the unrolled
// panel gets (a partially unrolled loop would index pv dynamically and move
// it to scratch memory).
template <int Q, int J, class L>
__device__ __forceinline__ void lu_column(float (&pv)[Q][kNB], int n, int k, L& S, int lane, int& info) {
  const int kj = k + J;
  uint32_t key = 0;
  int crow = INT_MAX;
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const int row = k + lane + 64 * q;
    if (row >= kj && row < n) {
      const float a = fabsf(pv[q][J]);
      const uint32_t kk = (a == a) ? __float_as_uint(a) + 1u : (row == kj ? 0xFFFFFFFFu : 0u);
      if (kk > key) { key = kk; crow = row; }  // rows ascend with q: ties keep the first
    }
  }
  const uint32_t K = wave_umax(key);
  const int pr = (int)wave_umin(key == K ? (uint32_t)crow : 0xFFFFFFFFu);
  if (pr != kj) {
    // row kj sits in lane J (q = 0), row pr in lane lp, slot qp
    const int lp = (pr - k) & 63, qp = (pr - k) >> 6;
    if (lane == J) {
#pragma unroll
      for (int c = 0; c < kNB; c++) S.rowbuf[0][c] = pv[0][c];
    }
    // one block per slot, each marked distinct: merged into one block the
    // accesses would become pv[qp][c], a dynamic index
#pragma unroll
    for (int q = 0; q < Q; q++)
      if (q == qp && lane == lp) {
        asm volatile("; slot %0" ::"n"(q));
#pragma unroll
        for (int c = 0; c < kNB; c++) S.rowbuf[1][c] = pv[q][c];
      }
    wave_sync();
    if (lane == J) {
#pragma unroll
      for (int c = 0; c < kNB; c++) pv[0][c] = S.rowbuf[1][c];
    }
#pragma unroll
    for (int q = 0; q < Q; q++)
      if (q == qp && lane == lp) {
        asm volatile("; slot %0" ::"n"(q));
#pragma unroll
        for (int c = 0; c < kNB; c++) pv[q][c] = S.rowbuf[0][c];
      }
    if (lane == 0) {
      const int sw = S.perm[kj];
      S.perm[kj] = S.perm[pr];
      S.perm[pr] = sw;
    }
    wave_sync();
  }
  if (lane == 0) S.ipiv[kj] = pr;
  // the pivot row, now in lane J, read into scalar registers
  const float praw = readlane_f(pv[0][J], J);
  if (praw == 0.f && info == 0) info = kj + 1;
  const float pivot = guard_f(praw);
  float u[kNB];
#pragma unroll
  for (int c = J + 1; c < kNB; c++) u[c] = readlane_f(pv[0][c], J);
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const int row = k + lane + 64 * q;
    if (row > kj && row < n) {
      const float l = pv[q][J] / pivot;
      pv[q][J] = l;
#pragma unroll
      for (int c = J + 1; c < kNB; c++) pv[q][c] -= l * u[c];
    }
  }
}

template <int Q, int J, class L>
__device__ __forceinline__ void lu_columns(float (&pv)[Q][kNB], int n, int k, int nb, L& S, int lane,
                                           int& info) {
  if constexpr (J < kNB) {
    if (J < nb) {
      lu_column<Q, J>(pv, n, k, S, lane, info);
      lu_columns<Q, J + 1>(pv, n, k, nb, S, lane, info);
    }
  }
}

template <int Q, class L>
__device__ __forceinline__ void panel_lu_unrolled(float* __restrict__ F, int n, int k, int nb, L& S, int lane,
                                         int& info) {
  const __amdgpu_buffer_rsrc_t rf = buf_rsrc(F);
  float pv[Q][kNB];
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const int row = k + lane + 64 * q;
    const uint32_t ro = (uint32_t)(row < n ? row : 0) * (uint32_t)n + (uint32_t)k;
#pragma unroll
    for (int c = 0; c < kNB; c++) pv[q][c] = buf_load<float>(rf, ((row < n) & (c < nb)) ? 4u * (ro + c) : kBufOob);
  }
  lu_columns<Q, 0>(pv, n, k, nb, S, lane, info);
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const int row = k + lane + 64 * q;
    if (row < n) {
#pragma unroll
      for (int c = 0; c < kNB; c++)
        if (c < nb) F[(int64_t)row * n + k + c] = pv[q][c];
    }
  }
}

// U12 := L11^-1 A12 over columns k+nb..n.  Lane i < nb holds row i of the
// unit-lower L11, so each entry reaches every lane as a scalar read (an LDS
// copy read by all lanes gets all 496 reads hoisted ahead of the arithmetic
// and spills), and is applied to the lane's Q columns.
template <int Q>
__device__ __forceinline__ void trsm_u12_unrolled(float* __restrict__ F, int n, int k, int nb,
                                                  float* __restrict__ tt, int lane) {
  const __amdgpu_buffer_rsrc_t rf = buf_rsrc(F);
  float l11[kNB];
  {
    const uint32_t ro = (uint32_t)(k + (lane < nb ? lane : 0)) * (uint32_t)n + (uint32_t)k;
#pragma unroll
    for (int c = 0; c < kNB; c++) l11[c] = buf_load<float>(rf, ((lane < nb) & (c < nb)) ? 4u * (ro + c) : kBufOob);
  }
  float x[Q][kNB];
#pragma unroll
  for (int s = 0; s < Q; s++) {
    const int c = k + nb + lane + 64 * s;
#pragma unroll
    for (int i = 0; i < kNB; i++)
      x[s][i] = buf_load<float>(rf, ((c < n) & (i < nb)) ? 4u * ((uint32_t)(k + i) * (uint32_t)n + (uint32_t)c) : kBufOob);
  }
#pragma unroll
  for (int i = 1; i < kNB; i++) {
#pragma unroll
    for (int jj = 0; jj < i; jj++) {
      const float l = readlane_f(l11[jj], i);
#pragma unroll
      for (int s = 0; s < Q; s++) x[s][i] -= l * x[s][jj];
    }
  }
#pragma unroll
  for (int s = 0; s < Q; s++) {
    const int c = k + nb + lane + 64 * s;
    if (c < n) {
#pragma unroll
      for (int i = 0; i < kNB; i++)
        if (i < nb) F[(int64_t)(k + i) * n + c] = x[s][i];
    }
  }
  wave_sync();
}

// ---- Cholesky panel: the diagonal block in lanes 0..nb, then L21 ------------
// _array/cholesky.mojo:26-73's arithmetic per column: the diagonal
// guard_nonzero'ed before the square root (a negative pivot gives NaN), the
// column divided by ljj; right-looking here.
template <int Q, class L>
__device__ __forceinline__ void panel_chol_unrolled(float* __restrict__ F, int n, int k, int nb, L& S, int lane,
                                           int& info) {
  const __amdgpu_buffer_rsrc_t rf = buf_rsrc(F);
  float dv[kNB];
  {
    const bool own = lane < nb;
    const uint32_t ro = (uint32_t)(k + (own ? lane : 0)) * (uint32_t)n + (uint32_t)k;
#pragma unroll
    for (int c = 0; c < kNB; c++) dv[c] = buf_load<float>(rf, (own & (c < nb)) ? 4u * (ro + c) : kBufOob);
  }
  // a guard, not a break: with an early exit the loop is left rolled and
  // dv indexed at run time (a copy of all 32 registers per inner step)
#pragma unroll
  for (int j = 0; j < kNB; j++) {
    if (j < nb) {
      const float d = readlane_f(dv[j], j);
      if (!(d > 0.f) && info == 0) info = k + j + 1;
      const float ljj = sqrtf(guard_f(d));
      if (lane == j) dv[j] = ljj;
      else if (lane > j) dv[j] = dv[j] / ljj;
#pragma unroll
      for (int c = j + 1; c < kNB; c++) {
        const float lc = readlane_f(dv[j], c);
        if (lane > j) dv[c] -= dv[j] * lc;
      }
    }
  }
  if (lane < nb) {
#pragma unroll
    for (int c = 0; c < kNB; c++)
      if (c < nb) F[(int64_t)(k + lane) * n + k + c] = c <= lane ? dv[c] : 0.f;
  }
  // L21 := A21 L11^-T, one row per lane and slot; L11[c][jj] reaches every
  // lane as a scalar read of lane c's register jj
  float x[Q][kNB];
#pragma unroll
  for (int s = 0; s < Q; s++) {
    const int row = k + nb + lane + 64 * s;
#pragma unroll
    for (int c = 0; c < kNB; c++)
      x[s][c] = buf_load<float>(rf, ((row < n) & (c < nb)) ? 4u * ((uint32_t)row * (uint32_t)n + (uint32_t)(k + c))
                                                          : kBufOob);
  }
#pragma unroll
  for (int c = 0; c < kNB; c++) {
#pragma unroll
    for (int jj = 0; jj < c; jj++) {
      const float l = readlane_f(dv[jj], c);
#pragma unroll
      for (int s = 0; s < Q; s++) x[s][c] -= x[s][jj] * l;
    }
    const float d = readlane_f(dv[c], c);
#pragma unroll
    for (int s = 0; s < Q; s++) x[s][c] = x[s][c] / d;
  }
#pragma unroll
  for (int s = 0; s < Q; s++) {
    const int row = k + nb + lane + 64 * s;
    if (row < n) {
#pragma unroll
      for (int c = 0; c < kNB; c++)
        if (c < nb) F[(int64_t)row * n + k + c] = x[s][c];
    }
  }
  wave_sync();
}

// ---- trailing update on MFMA -----------------------------------------------
// C -= L21 U12 (LU) or the lower tiles of C -= L21 L21^T (Cholesky), 32 x 32
// tiles of v_mfma_f32_32x32x2f32 (A element: lane l holds A[l%32][l/32]; B:
// B[l/32][l%32]; D item i: row 8 (i/4) + 4 (l/32) + i%4, column l%32).
// The A operand (a 32-row tile of L21, and the B operand of a Cholesky
// update, which is the same layout of another L21 tile) wants a row per lane
// pair: it is loaded as 128-byte row segments and rearranged in the staging
// tile.
__device__ __forceinline__ void l21_raw(__amdgpu_buffer_rsrc_t rf, int n, int k, int nb, int row0, int lane,
                                        float (&raw)[16]) {
  const int cl = lane & 31, rh = lane >> 5;
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const int r = row0 + 2 * i + rh;
    raw[i] = buf_load<float>(rf, ((r < n) & (cl < nb)) ? 4u * ((uint32_t)r * (uint32_t)n + (uint32_t)(k + cl))
                                                       : kBufOob);
  }
}
// raw (row 2 i + l / 32, column l % 32 of the tile) to the MFMA layout
// (row l % 32, column 2 s + l / 32)
__device__ __forceinline__ void l21_mfma(float* __restrict__ tt, int lane, const float (&raw)[16], float (&a)[16]) {
  const int cl = lane & 31, rh = lane >> 5;
#pragma unroll
  for (int i = 0; i < 16; i++) tt[(2 * i + rh) * kTileLd + cl] = raw[i];
  lds_sync();
#pragma unroll
  for (int s = 0; s < 16; s++) a[s] = tt[cl * kTileLd + 2 * s + rh];
  lds_sync();
}

template <bool CHOL>
__device__ __forceinline__ void trailing_update(float* __restrict__ F, int n, int k, int nb, float* __restrict__ tt,
                                                int lane) {
  const int r0 = k + nb;
  const int m2 = n - r0;
  if (m2 <= 0) return;
  const int tiles = (m2 + 31) >> 5;
  const __amdgpu_buffer_rsrc_t rf = buf_rsrc(F);
  const int l32 = lane & 31, kh = lane >> 5;
#pragma unroll 1
  for (int tj = 0; tj < tiles; tj++) {
    const int col = r0 + 32 * tj + l32;
    float b[16];
    if (CHOL) {
      float raw[16];
      l21_raw(rf, n, k, nb, r0 + 32 * tj, lane, raw);
      l21_mfma(tt, lane, raw, b);
    } else {
#pragma unroll
      for (int s = 0; s < 16; s++) {
        const int kk = 2 * s + kh;
        b[s] = buf_load<float>(rf, ((kk < nb) & (col < n)) ? 4u * ((uint32_t)(k + kk) * (uint32_t)n + (uint32_t)col)
                                                           : kBufOob);
      }
    }
    auto load_c = [&](int ti, floatx16& c) {
#pragma unroll
      for (int i = 0; i < 16; i++) {
        const int row = r0 + 32 * ti + 8 * (i >> 2) + 4 * kh + (i & 3);
        c[i] = buf_load<float>(rf, ((row < n) & (col < n) & (ti < tiles))
                                       ? 4u * ((uint32_t)row * (uint32_t)n + (uint32_t)col)
                                       : kBufOob);
      }
    };
    const int t0 = CHOL ? tj : 0;
    float a[16];
    floatx16 acc, cn;
    {
      float raw[16];
      l21_raw(rf, n, k, nb, r0 + 32 * t0, lane, raw);
      l21_mfma(tt, lane, raw, a);
    }
    load_c(t0, acc);
    // tile ti + 1's operands are loaded while tile ti is on the matrix cores
#pragma unroll 1
    for (int ti = t0; ti < tiles; ti++) {
      float rawn[16];
      l21_raw(rf, n, k, nb, r0 + 32 * (ti + 1), lane, rawn);
      load_c(ti + 1, cn);
#pragma unroll
      for (int s = 0; s < 16; s++) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(-a[s], b[s], acc, 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 16; i++) {
        const int row = r0 + 32 * ti + 8 * (i >> 2) + 4 * kh + (i & 3);
        if (row < n && col < n) F[(int64_t)row * n + col] = acc[i];
      }
      l21_mfma(tt, lane, rawn, a);
      acc = cn;
    }
  }
  wave_sync();
}

// op(F) x = xs in place, 64 rows at a time: the diagonal block is staged
// in LDS and swept by a rolled loop (lane c's value reaches the others as a
// scalar read), then the other rows are updated by one dot product per lane
// and slot.  TRANS reads F transposed (L^T).
template <int Q, bool LOWER, bool UNIT, bool TRANS, class L>
__device__ __forceinline__ void tri_solve(const float* __restrict__ F, int n, L& S, int lane) {
  const __amdgpu_buffer_rsrc_t rf = buf_rsrc(F);
  const int nbk = (n + 63) >> 6;
#pragma unroll 1
  for (int bb = 0; bb < nbk; bb++) {
    const int b = LOWER ? bb : nbk - 1 - bb;
    const int r0 = 64 * b, bs = min(64, n - r0);
    // stage op(F)[r0 + i][r0 + c] at blk[i][c]; F's rows are read along the
    // lanes (coalesced) and TRANS stores them as columns of blk
#pragma unroll 1
    for (int i0 = 0; i0 < 64; i0 += 16) {
      float v[16];
#pragma unroll
      for (int u = 0; u < 16; u++) {
        const int i = i0 + u;
        const uint32_t off = (uint32_t)(r0 + i) * (uint32_t)n + (uint32_t)(r0 + lane);
        v[u] = buf_load<float>(rf, ((lane < bs) & (i < bs)) ? 4u * off : kBufOob);
      }
#pragma unroll
      for (int u = 0; u < 16; u++) {
        if (TRANS) S.blk[lane][i0 + u] = v[u];
        else S.blk[i0 + u][lane] = v[u];
      }
    }
    wave_sync();
    float acc = lane < bs ? S.xs[r0 + lane] : 0.f;
#pragma unroll 1
    for (int t = 0; t < bs; t++) {
      const int c = LOWER ? t : bs - 1 - t;
      float xc = readlane_f(acc, c);
      if (!UNIT) xc = xc / guard_f(S.blk[c][c]);
      if (lane == c) acc = xc;
      else if (LOWER ? lane > c : lane < c) acc -= S.blk[lane][c] * xc;
    }
    if (lane < bs) S.xs[r0 + lane] = acc;
    wave_sync();
    {
      float part[Q];
      bool todo[Q];
#pragma unroll
      for (int s = 0; s < Q; s++) {
        const int row = lane + 64 * s;
        todo[s] = row < n && (LOWER ? row >= r0 + bs : row < r0);
        part[s] = 0.f;
      }
#pragma unroll 1
      for (int c0 = 0; c0 < bs; c0 += 16) {
        float fv[Q][16];
#pragma unroll
        for (int s = 0; s < Q; s++) {
          const int row = lane + 64 * s;
#pragma unroll
          for (int u = 0; u < 16; u++) {
            const int c = c0 + u;
            const uint32_t off = TRANS ? (uint32_t)(r0 + c) * (uint32_t)n + (uint32_t)row
                                       : (uint32_t)row * (uint32_t)n + (uint32_t)(r0 + c);
            fv[s][u] = buf_load<float>(rf, (todo[s] & (c < bs)) ? 4u * off : kBufOob);
          }
        }
#pragma unroll
        for (int u = 0; u < 16; u++) {
          const float xc = S.xs[r0 + (c0 + u < bs ? c0 + u : 0)];
#pragma unroll
          for (int s = 0; s < Q; s++) part[s] += fv[s][u] * xc;
        }
      }
#pragma unroll
      for (int s = 0; s < Q; s++)
        if (todo[s]) S.xs[lane + 64 * s] -= part[s];
    }
    wave_sync();
  }
}

// op(F) x = xs in place, 64 rows at a time: the diagonal block with every
// index static (values exchanged through scalar reads), then the other rows
// by one dot product per lane and slot.  TRANS reads F transposed (L^T).
template <int Q, bool LOWER, bool UNIT, bool TRANS, class L>
__device__ __forceinline__ void tri_solve_unrolled(const float* __restrict__ F, int n, L& S, int lane) {
  const __amdgpu_buffer_rsrc_t rf = buf_rsrc(F);
  const int nbk = (n + 63) >> 6;
#pragma unroll 1
  for (int bb = 0; bb < nbk; bb++) {
    const int b = LOWER ? bb : nbk - 1 - bb;
    const int r0 = 64 * b, bs = min(64, n - r0);
    {
      float a[64];
#pragma unroll
      for (int c = 0; c < 64; c++) {
        const uint32_t off = TRANS ? (uint32_t)(r0 + c) * (uint32_t)n + (uint32_t)(r0 + lane)
                                   : (uint32_t)(r0 + lane) * (uint32_t)n + (uint32_t)(r0 + c);
        a[c] = buf_load<float>(rf, ((lane < bs) & (c < bs)) ? 4u * off : kBufOob);
      }
      float acc = lane < bs ? S.xs[r0 + lane] : 0.f;
      // an opaque lane index keeps the 128 comparisons below from being hoisted
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
    wave_sync();
    {
      float part[Q];
      bool todo[Q];
#pragma unroll
      for (int s = 0; s < Q; s++) {
        const int row = lane + 64 * s;
        todo[s] = row < n && (LOWER ? row >= r0 + bs : row < r0);
        part[s] = 0.f;
      }
#pragma unroll 1
      for (int c0 = 0; c0 < bs; c0 += 16) {
        float fv[Q][16];
#pragma unroll
        for (int s = 0; s < Q; s++) {
          const int row = lane + 64 * s;
#pragma unroll
          for (int u = 0; u < 16; u++) {
            const int c = c0 + u;
            const uint32_t off = TRANS ? (uint32_t)(r0 + c) * (uint32_t)n + (uint32_t)row
                                       : (uint32_t)row * (uint32_t)n + (uint32_t)(r0 + c);
            fv[s][u] = buf_load<float>(rf, (todo[s] & (c < bs)) ? 4u * off : kBufOob);
          }
        }
#pragma unroll
        for (int u = 0; u < 16; u++) {
          const float xc = S.xs[r0 + (c0 + u < bs ? c0 + u : 0)];
#pragma unroll
          for (int s = 0; s < Q; s++) part[s] += fv[s][u] * xc;
        }
      }
#pragma unroll
      for (int s = 0; s < Q; s++)
        if (todo[s]) S.xs[lane + 64 * s] -= part[s];
    }
    wave_sync();
  }
}

// After the factorization, F is rewritten for the solves, which read it
// only as op(F) = F^T (a load instruction then covers consecutive words
// instead of one row per lane): an LU factor L\U is transposed in place; a
// Cholesky factor, whose strict upper triangle is unused, gets L^T there
// (F = L + strict_upper(L^T), so op(F)'s lower triangle is L and its upper
// L^T).  32 x 32 tiles go through a 32 x 33 LDS tile: lane l covers row
// 2 i + l / 32 and column l % 32 of a tile, i = 0..15, for both the load and
// the store, so every global access is two 128-byte rows.
__device__ __forceinline__ void tile_load(__amdgpu_buffer_rsrc_t rf, int n, int I, int J, int lane, float (&v)[16]) {
  const int cl = lane & 31, rh = lane >> 5;
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const int r = 32 * I + 2 * i + rh, c = 32 * J + cl;
    v[i] = buf_load<float>(rf, ((r < n) & (c < n)) ? 4u * ((uint32_t)r * (uint32_t)n + (uint32_t)c) : kBufOob);
  }
}
// v becomes its transpose (v[i] at row 2 i + l / 32, column l % 32 then
// holds the entry at row l % 32, column 2 i + l / 32)
__device__ __forceinline__ void tile_xpose(float* __restrict__ tt, int lane, float (&v)[16]) {
  const int cl = lane & 31, rh = lane >> 5;
#pragma unroll
  for (int i = 0; i < 16; i++) tt[(2 * i + rh) * 33 + cl] = v[i];
  wave_sync();
#pragma unroll
  for (int i = 0; i < 16; i++) v[i] = tt[cl * 33 + 2 * i + rh];
  wave_sync();
}
__device__ __forceinline__ void tile_store(__amdgpu_buffer_rsrc_t rf, int n, int I, int J, int lane,
                                           const float (&v)[16], bool upper_only) {
  const int cl = lane & 31, rh = lane >> 5;
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const int r = 32 * I + 2 * i + rh, c = 32 * J + cl;
    const bool ok = (r < n) & (c < n) & (!upper_only | (c > r));
    buf_store(rf, ok ? 4u * ((uint32_t)r * (uint32_t)n + (uint32_t)c) : kBufOob, v[i]);
  }
}

template <bool CHOL>
__device__ __forceinline__ void factor_for_solves(float* __restrict__ F, int n, float* __restrict__ tt, int lane) {
  const __amdgpu_buffer_rsrc_t rf = buf_rsrc(F);
  const int nt = (n + 31) >> 5;
#pragma unroll 1
  for (int I = 0; I < nt; I++) {
    if (CHOL) {
#pragma unroll 1
      for (int J = 0; J <= I; J++) {  // tile (I, J) of L, transposed, into (J, I); above the diagonal only
        float v[16];
        tile_load(rf, n, I, J, lane, v);
        tile_xpose(tt, lane, v);
        tile_store(rf, n, J, I, lane, v, I == J);
      }
    } else {
      {
        float v[16];
        tile_load(rf, n, I, I, lane, v);
        tile_xpose(tt, lane, v);
        tile_store(rf, n, I, I, lane, v, false);
      }
#pragma unroll 1
      for (int J = I + 1; J < nt; J++) {  // swap (I, J) and (J, I), each transposed
        float v[16], w[16];
        tile_load(rf, n, I, J, lane, v);
        tile_load(rf, n, J, I, lane, w);
        tile_xpose(tt, lane, v);
        tile_xpose(tt, lane, w);
        tile_store(rf, n, J, I, lane, v, false);
        tile_store(rf, n, I, J, lane, w, false);
      }
    }
  }
  wave_sync();
}

// The classes with n <= 128 run the rolled phases (compact code, so the
// instruction cache holds a wavefront's loop); the classes with n > 128 run
// the fully unrolled ones.
template <int Q>
constexpr bool kRolled = Q <= 2;

template <int Q, bool CHOL, class L>
__device__ __forceinline__ void solve_wave(const float* __restrict__ F, int n, L& S, int lane) {
  // F holds the factor transposed (LU) or mirrored (Cholesky), so both
  // triangles are read as op(F) = F^T: consecutive lanes, consecutive words
  if constexpr (kRolled<Q>) {
    tri_solve<Q, true, !CHOL, true>(F, n, S, lane);
    tri_solve<Q, false, false, true>(F, n, S, lane);
  } else {
    tri_solve_unrolled<Q, true, !CHOL, true>(F, n, S, lane);
    tri_solve_unrolled<Q, false, false, true>(F, n, S, lane);
  }
}

// Sums of v[i] over the 16 lanes of each row of 16 lanes, scattered: lane c
// of a row returns the sum for i = c.  Four halving exchanges (lane c keeps
// the half of its values whose index has c's bit and sends the other half to
// lane c ^ bit), 15 shuffles in all instead of 16 reductions.
__device__ __forceinline__ double reduce_scatter16(const double (&v)[16], int lane) {
  const int c = lane & 15;
  double w8[8], w4[4], w2[2];
#pragma unroll
  for (int t = 0; t < 8; t++) {
    const bool hi = c & 8;
    w8[t] = (hi ? v[t + 8] : v[t]) + __shfl_xor(hi ? v[t] : v[t + 8], 8);
  }
#pragma unroll
  for (int t = 0; t < 4; t++) {
    const bool hi = c & 4;
    w4[t] = (hi ? w8[t + 4] : w8[t]) + __shfl_xor(hi ? w8[t] : w8[t + 4], 4);
  }
#pragma unroll
  for (int t = 0; t < 2; t++) {
    const bool hi = c & 2;
    w2[t] = (hi ? w4[t + 2] : w4[t]) + __shfl_xor(hi ? w4[t] : w4[t + 2], 2);
  }
  const bool hi = c & 1;
  return (hi ? w2[1] : w2[0]) + __shfl_xor(hi ? w2[0] : w2[1], 1);
}

// r = b - A x64 in FP64 (a Cholesky problem is the symmetric matrix its
// lower triangle defines), rounded to FP32 into xs in the order the next
// solve takes its right-hand side (row perm[p] at p); returns
// ||r||_inf / (||A||_inf ||x||_inf).  ||A||_inf is summed on the first call.
// Positions go 64 at a time; lane l covers positions g + 4 i + l / 16
// (i = 0..15) and columns c0 + l % 16 of each 16-column chunk, so a load
// instruction reads four 128-byte row segments, and the 16 partial sums per
// lane are combined once per group by reduce_scatter16.  The strict upper
// triangle of a Cholesky matrix is read as columns of its lower triangle:
// lane l then covers positions g + 16 i + l % 16 (i = 0..3) and rows
// j0 + l / 16 of A, four 128-byte column segments a load.
template <int Q, bool CHOL, bool FIRST, class L>
__device__ __forceinline__ double residual_wave(const double* __restrict__ A, const double* __restrict__ b, int n,
                                                double& anorm, L& S, int lane) {
  constexpr bool first = FIRST;
  const __amdgpu_buffer_rsrc_t ra = buf_rsrc(A);
  const int cl = lane & 15, rq = lane >> 4;
  double rmax = 0.0, amax = 0.0;
#pragma unroll 1
  for (int g = 0; g < n; g += 64) {
    int rowi[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int p = g + 4 * i + rq;
      rowi[i] = p < n ? (CHOL ? p : S.perm[p]) : -1;
    }
    double acc[16], as[16];
#pragma unroll
    for (int i = 0; i < 16; i++) acc[i] = as[i] = 0.0;
    const int cend = CHOL ? min(n, g + 64) : n;  // Cholesky: columns j <= row here
#pragma unroll 1
    for (int c0 = 0; c0 < cend; c0 += 16) {
      const int j = c0 + cl;
      const double xj = S.x64[j < n ? j : 0];
      double av[16];
#pragma unroll
      for (int i = 0; i < 16; i++) {
        const int row = rowi[i];
        const bool ok = (row >= 0) & (j < n) & (!CHOL | (j <= row));
        av[i] = buf_load<double>(ra, ok ? 8u * ((uint32_t)row * (uint32_t)n + (uint32_t)j) : kBufOob);
      }
#pragma unroll
      for (int i = 0; i < 16; i++) {
        acc[i] += av[i] * xj;
        if (first) as[i] += fabs(av[i]);
      }
    }
    double sum = reduce_scatter16(acc, lane);
    double asum = 0.0;
    if constexpr (first) asum = reduce_scatter16(as, lane);
    const int p = g + 4 * cl + rq;  // this lane's position after the scatter
    if (CHOL) {
      const int pr = lane & 15, jq = lane >> 4;
      double up[4] = {0.0, 0.0, 0.0, 0.0}, ua[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
      for (int j0 = g; j0 < n; j0 += 4) {
        const int jj = j0 + jq;
        const double xj = S.x64[jj < n ? jj : 0];
        double av[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int pp = g + 16 * i + pr;
          const bool ok = (pp < n) & (jj < n) & (jj > pp);
          av[i] = buf_load<double>(ra, ok ? 8u * ((uint32_t)jj * (uint32_t)n + (uint32_t)pp) : kBufOob);
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
          up[i] += av[i] * xj;
          if (first) ua[i] += fabs(av[i]);
        }
      }
#pragma unroll
      for (int i = 0; i < 4; i++) {
        up[i] += __shfl_xor(up[i], 16);
        up[i] += __shfl_xor(up[i], 32);
        if (first) {
          ua[i] += __shfl_xor(ua[i], 16);
          ua[i] += __shfl_xor(ua[i], 32);
        }
      }
      // position p = g + 16 (m / 16) + m % 16 with m = 4 cl + rq: entry m / 16 of lane m % 16
      const int m = 4 * cl + rq, src = m & 15, e = m >> 4;
      double u = 0.0, a = 0.0;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const double ui = __shfl(up[i], src);
        const double ai = first ? __shfl(ua[i], src) : 0.0;
        if (e == i) { u = ui; a = ai; }
      }
      sum += u;
      asum += a;
    }
    if (p < n) {
      const double r = b[CHOL ? p : S.perm[p]] - sum;
      S.xs[p] = (float)r;
      rmax = fmax(rmax, fabs(r));
      amax = fmax(amax, asum);
    }
  }
  double xmax = 0.0;
#pragma unroll 1
  for (int s = 0; s < Q; s++) {
    const int i = lane + 64 * s;
    if (i < n) xmax = fmax(xmax, fabs(S.x64[i]));
  }
  rmax = wave_dmax(rmax);
  xmax = wave_dmax(xmax);
  if (first) anorm = wave_dmax(amax);
  wave_sync();
  const double den = anorm * xmax;
  return den > 0.0 ? rmax / den : (rmax == 0.0 ? 0.0 : INFINITY);
}

// Two wavefronts per SIMD where the registers allow it: left to itself the
// compiler takes 256 VGPRs and 16-28 AGPRs, which leaves room for one, and
// a lone wavefront waits out every load of its chain of dependent phases.
// The unrolled LU panel (n > 128) held to 256 registers spills about ten
// words to scratch memory per column step, which costs more than the second
// wavefront gains (14.3 ms against 15.2 for the 20,000-system batch), so
// those two classes keep one.
template <int Q, bool CHOL>
constexpr int kWavesPerSimd = (CHOL || kRolled<Q>) ? 2 : 1;
template <int Q, bool CHOL>
__global__ void __launch_bounds__(64 * kWaves) __attribute__((amdgpu_waves_per_eu(kWavesPerSimd<Q, CHOL>)))
batched_wave_kernel(BatchArgs args, int cls) {
  using Lds = typename std::conditional<kRolled<Q>, WaveLdsBlk<64 * Q>, WaveLdsTile<64 * Q>>::type;
  __shared__ Lds lds[kWaves];
  const int lane = threadIdx.x & 63;
  // the wavefront index as a scalar, so every LDS address below is one base
  // register plus an immediate offset
  Lds& S = lds[__builtin_amdgcn_readfirstlane(threadIdx.x >> 6)];
  const int begin = args.cls[cls], count = args.cls[kClasses + cls];
  for (;;) {
    int slot = 0;
    if (lane == 0) slot = atomicAdd(&args.head[cls], 1);
    slot = __shfl(slot, 0);
    if (slot >= count) break;
    const int p = args.order[begin + slot];
    const int n = args.n[p];
    const double* A = args.A + args.offA[p];
    const double* b = args.B + args.offB[p];
    float* F = args.F + args.offF[p];

    // FP32 working copy (a Cholesky problem keeps its lower triangle);
    // 32 loads in flight per lane, then the stores
    {
      const __amdgpu_buffer_rsrc_t ra = buf_rsrc(A);
      constexpr int RB = 32 / Q;  // rows per batch
#pragma unroll 1
      for (int i0 = 0; i0 < n; i0 += RB) {
        double v[RB][Q];
#pragma unroll
        for (int r = 0; r < RB; r++) {
          const int i = i0 + r;
#pragma unroll
          for (int s = 0; s < Q; s++) {
            const int j = lane + 64 * s;
            const bool keep = (i < n) & (j < n) & (!CHOL | (j <= i));
            v[r][s] = buf_load<double>(ra, keep ? 8u * ((uint32_t)i * (uint32_t)n + (uint32_t)j) : kBufOob);
          }
        }
#pragma unroll
        for (int r = 0; r < RB; r++) {
          const int i = i0 + r;
#pragma unroll
          for (int s = 0; s < Q; s++) {
            const int j = lane + 64 * s;
            if ((i < n) & (j < n)) F[(int64_t)i * n + j] = (float)v[r][s];
          }
        }
      }
      for (int s = 0; s < Q; s++) {
        const int i = lane + 64 * s;
        if (i < n) S.perm[i] = i;
      }
    }
    wave_sync();

    int info = 0;
#pragma unroll 1
    for (int k = 0; k < n; k += kNB) {
      const int nb = min(kNB, n - k);
      if (CHOL) {
        if constexpr (kRolled<Q>) panel_chol<Q>(F, n, k, nb, S, lane, info);
        else panel_chol_unrolled<Q>(F, n, k, nb, S, lane, info);
      } else {
        if constexpr (kRolled<Q>) panel_lu<Q>(F, n, k, nb, S, lane, info);
        else panel_lu_unrolled<Q>(F, n, k, nb, S, lane, info);
        wave_sync();
        swap_outside<Q>(F, n, k, nb, S, lane);
        if (k + nb < n) {
          if constexpr (kRolled<Q>) trsm_u12<Q>(F, n, k, nb, tile_of(S), lane);
          else trsm_u12_unrolled<Q>(F, n, k, nb, tile_of(S), lane);
        }
      }
      trailing_update<CHOL>(F, n, k, nb, tile_of(S), lane);
    }
    factor_for_solves<CHOL>(F, n, tile_of(S), lane);

    // x0 = op^-1 P b in FP32, then FP64 refinement (one solve site)
    double anorm = 0.0, eta = 0.0, prev = INFINITY;
    int it = 0;
    for (bool first = true;; first = false) {
      if (first) {  // later right-hand sides are the residuals residual_wave left in xs
#pragma unroll 1
        for (int s = 0; s < Q; s++) {
          const int row = lane + 64 * s;
          if (row < n) S.xs[row] = (float)b[CHOL ? row : S.perm[row]];
        }
        wave_sync();
      }
      solve_wave<Q, CHOL>(F, n, S, lane);
#pragma unroll 1
      for (int s = 0; s < Q; s++) {
        const int row = lane + 64 * s;
        if (row < n) S.x64[row] = first ? (double)S.xs[row] : S.x64[row] + (double)S.xs[row];
      }
      wave_sync();
      eta = first ? residual_wave<Q, CHOL, true>(A, b, n, anorm, S, lane)
                  : residual_wave<Q, CHOL, false>(A, b, n, anorm, S, lane);
      if (!first) it++;
      if (!(eta > args.tol) || it >= args.max_iter || !(eta < 0.5 * prev)) break;
      prev = eta;
    }
#pragma unroll 1
    for (int s = 0; s < Q; s++) {
      const int row = lane + 64 * s;
      if (row < n) args.X[args.offB[p] + row] = S.x64[row];
    }
    if (lane == 0) {
      args.info[p] = info;
      args.iters[p] = it;
      args.berr[p] = eta;
    }
    wave_sync();
  }
}

// ---- size-class queue: counting sort by (kind, n), largest first -------------
// A problem is taken when 1 <= n <= 256, kind is 0 or 1 and its matrix and
// vectors lie inside the arrays the caller sized (so no kernel can touch
// memory outside them); any other is reported with info = -1.
__device__ __forceinline__ bool batch_ok(int nn, int kd, int64_t oa, int64_t ob, int64_t a_elems, int64_t b_elems) {
  return nn >= 1 && nn <= kBatchMaxN && kd <= 1 && oa >= 0 && oa + (int64_t)nn * nn <= a_elems && ob >= 0 &&
         ob + nn <= b_elems;
}

// Offsets of the FP32 working copies: an exclusive prefix sum of n^2 over
// the problems in index order, so problems that share (or overlap in) A
// still factor into separate scratch.  One workgroup; a problem whose copy
// would pass f_elems (only possible when problems overlap in A) gets -1 and
// is skipped like an invalid one.
constexpr int kScanThreads = 1024;
__global__ void __launch_bounds__(kScanThreads) batch_offf_kernel(int count, const int32_t* __restrict__ n,
                                                                  const uint8_t* __restrict__ kind,
                                                                  const int64_t* __restrict__ offA,
                                                                  const int64_t* __restrict__ offB, int64_t a_elems,
                                                                  int64_t b_elems, int64_t f_elems,
                                                                  int64_t* __restrict__ offF) {
  __shared__ int64_t part[kScanThreads];
  const int t = threadIdx.x;
  const int chunk = (count + kScanThreads - 1) / kScanThreads;
  const int lo = min(count, t * chunk), hi = min(count, lo + chunk);
  int64_t sum = 0;
  for (int i = lo; i < hi; i++) {
    const int nn = n[i];
    if (batch_ok(nn, kind[i], offA[i], offB[i], a_elems, b_elems)) sum += (int64_t)nn * nn;
  }
  part[t] = sum;
  __syncthreads();
  for (int d = 1; d < kScanThreads; d <<= 1) {  // Hillis-Steele inclusive scan
    const int64_t v = t >= d ? part[t - d] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int64_t run = part[t] - sum;
  for (int i = lo; i < hi; i++) {
    const int nn = n[i];
    if (batch_ok(nn, kind[i], offA[i], offB[i], a_elems, b_elems)) {
      const int64_t need = (int64_t)nn * nn;
      offF[i] = run + need <= f_elems ? run : -1;
      run += need;
    } else {
      offF[i] = -1;
    }
  }
}

__global__ void batch_hist_kernel(int count, const int32_t* __restrict__ n, const uint8_t* __restrict__ kind,
                                  const int64_t* __restrict__ offF, int32_t* __restrict__ hist,
                                  int32_t* __restrict__ info, double* __restrict__ berr, int32_t* __restrict__ iters) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  if (offF[i] >= 0) {
    atomicAdd(&hist[kind[i] * (kBatchMaxN + 1) + n[i]], 1);
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
                                     const int64_t* __restrict__ offF, int32_t* __restrict__ cursor,
                                     int32_t* __restrict__ order) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  if (offF[i] >= 0) order[atomicAdd(&cursor[kind[i] * (kBatchMaxN + 1) + n[i]], 1)] = i;
}

__global__ void zero_i32_kernel(int32_t* p, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0;
}

template <int Q, bool CHOL>
int launch_class(nmx_ctx* ctx, const BatchArgs& a, int per_cu) {
  const int cls = (CHOL ? kSizeClasses : 0) + (Q - 1);
  batched_wave_kernel<Q, CHOL><<<ctx->num_cus * per_cu, 64 * kWaves, 0, ctx->stream>>>(a, cls);
  NMX_HIP(ctx, hipGetLastError());
  return NMX_OK;
}

}  // namespace

NMX_API void nmx_batch_cfg_default(nmx_batch_cfg* cfg) {
  if (!cfg) return;
  cfg->max_iter = 10;
  cfg->tol = 1e-15;
  cfg->teams_per_cu[0] = 2;  // workgroups of four independent wavefronts per CU, n > 128
  cfg->teams_per_cu[1] = 2;  // 64 < n <= 128
  cfg->teams_per_cu[2] = 4;  // n <= 64
}

// Arguments of the batched solve, checked once for both entry points.
static int batched_args(nmx_ctx* ctx, int count, const int32_t* n, const int64_t* offA, const int64_t* offB,
                        const uint8_t* kind, const double* A, int64_t a_elems, const double* B, int64_t b_elems,
                        double* X, int32_t* info, int32_t* iters, double* berr, const nmx_batch_cfg* cfg_in,
                        nmx_batch_cfg* cfg) {
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
  if (cfg_in) *cfg = *cfg_in; else nmx_batch_cfg_default(cfg);
  if (cfg->max_iter < 0 || !(cfg->tol >= 0.0)) return -15;
  for (int c = 0; c < 3; c++)
    if (cfg->teams_per_cu[c] <= 0 || cfg->teams_per_cu[c] > 64) return -15;
  return NMX_OK;
}

// Scratch of one call: FP32 working copies (as many elements as A), their
// offsets, then the size-class queue.
struct BatchWs {
  size_t f_bytes, o_bytes, q_ints;
  size_t total() const { return f_bytes + o_bytes + sizeof(int32_t) * q_ints; }
};
static BatchWs batched_ws(int count, int64_t a_elems) {
  BatchWs w;
  w.f_bytes = ((sizeof(float) * (size_t)a_elems + 255) / 256) * 256;
  w.o_bytes = ((sizeof(int64_t) * (size_t)count + 255) / 256) * 256;
  w.q_ints = 2 * (size_t)kBins + 2 * kClasses + kClasses + (size_t)count;
  return w;
}

// The solve itself, in the scratch `ws` (batched_ws bytes) the caller found;
// `two_streams` lets the Cholesky classes run on the context's second stream.
static int batched_run(nmx_ctx* ctx, int count, const int32_t* n, const int64_t* offA, const int64_t* offB,
                       const uint8_t* kind, const double* A, int64_t a_elems, const double* B, int64_t b_elems,
                       double* X, int32_t* info, int32_t* iters, double* berr, const nmx_batch_cfg& cfg, void* ws,
                       bool two_streams) {
  const BatchWs wl = batched_ws(count, a_elems);
  float* F = (float*)ws;
  int64_t* offF = (int64_t*)((char*)ws + wl.f_bytes);
  int32_t* hist = (int32_t*)((char*)ws + wl.f_bytes + wl.o_bytes);
  int32_t* cursor = hist + kBins;
  int32_t* cls = cursor + kBins;
  int32_t* head = cls + 2 * kClasses;
  int32_t* order = head + kClasses;

  {
    ProfScope ps(ctx, K_OTHER, 0.0, 0.0);
    zero_i32_kernel<<<ceil_div(kBins, 256), 256, 0, ctx->stream>>>(hist, kBins);
    batch_offf_kernel<<<1, kScanThreads, 0, ctx->stream>>>(count, n, kind, offA, offB, a_elems, b_elems, a_elems,
                                                           offF);
    batch_hist_kernel<<<ceil_div(count, 256), 256, 0, ctx->stream>>>(count, n, kind, offF, hist, info, berr, iters);
    batch_scan_kernel<<<1, 64, 0, ctx->stream>>>(hist, cursor, cls, head);
    batch_scatter_kernel<<<ceil_div(count, 256), 256, 0, ctx->stream>>>(count, n, kind, offF, cursor, order);
    NMX_HIP(ctx, hipGetLastError());
  }
  BatchArgs a{count, n, offA, offB, kind, A, B, X, info, iters, berr, F, offF, order, cls, head, cfg.max_iter, cfg.tol};
  {
    ProfScope ps(ctx, K_BATCH, 0.0, 8.0 * (double)a_elems);
    // Largest problems first.  Each class is one persistent launch, and
    // launches on one stream run one after another, so a class's last
    // problems would leave most of the chip idle; with two streams the
    // Cholesky classes fill the slots the LU classes free (and the reverse).
    // Problems are independent, so the results do not depend on the split.
    const bool split = two_streams && (ctx->aux || !ctx->capturing) && nmx_aux_stream(ctx) == NMX_OK;
    if (split) {
      NMX_HIP(ctx, hipEventRecord(ctx->ahead_ev, ctx->stream));
      NMX_HIP(ctx, hipStreamWaitEvent(ctx->aux, ctx->ahead_ev, 0));
    }
    NMX_TRY((launch_class<4, false>(ctx, a, cfg.teams_per_cu[0])));
    {
      StreamSwap on(ctx, split ? ctx->aux : ctx->stream);
      NMX_TRY((launch_class<4, true>(ctx, a, cfg.teams_per_cu[0])));
    }
    NMX_TRY((launch_class<3, false>(ctx, a, cfg.teams_per_cu[0])));
    {
      StreamSwap on(ctx, split ? ctx->aux : ctx->stream);
      NMX_TRY((launch_class<3, true>(ctx, a, cfg.teams_per_cu[0])));
    }
    NMX_TRY((launch_class<2, false>(ctx, a, cfg.teams_per_cu[1])));
    {
      StreamSwap on(ctx, split ? ctx->aux : ctx->stream);
      NMX_TRY((launch_class<2, true>(ctx, a, cfg.teams_per_cu[1])));
    }
    NMX_TRY((launch_class<1, false>(ctx, a, cfg.teams_per_cu[2])));
    {
      StreamSwap on(ctx, split ? ctx->aux : ctx->stream);
      NMX_TRY((launch_class<1, true>(ctx, a, cfg.teams_per_cu[2])));
    }
    if (split) {
      NMX_HIP(ctx, hipEventRecord(ctx->behind_ev, ctx->aux));
      NMX_HIP(ctx, hipStreamWaitEvent(ctx->stream, ctx->behind_ev, 0));
    }
  }
  return NMX_OK;
}

NMX_API int nmx_batched_gesv(nmx_ctx* ctx, int count, const int32_t* n, const int64_t* offA, const int64_t* offB,
                             const uint8_t* kind, const double* A, int64_t a_elems, const double* B, int64_t b_elems,
                             double* X, int32_t* info, int32_t* iters, double* berr, const nmx_batch_cfg* cfg_in) {
  nmx_batch_cfg cfg;
  NMX_TRY(batched_args(ctx, count, n, offA, offB, kind, A, a_elems, B, b_elems, X, info, iters, berr, cfg_in, &cfg));
  if (count == 0) return NMX_OK;
  void* ws = nullptr;
  NMX_TRY(nmx_workspace(ctx, batched_ws(count, a_elems).total(), &ws));
  return batched_run(ctx, count, n, offA, offB, kind, A, a_elems, B, b_elems, X, info, iters, berr, cfg, ws, true);
}

NMX_API int nmx_batched_gesv_noalloc(nmx_ctx* ctx, int count, const int32_t* n, const int64_t* offA,
                                     const int64_t* offB, const uint8_t* kind, const double* A, int64_t a_elems,
                                     const double* B, int64_t b_elems, double* X, int32_t* info, int32_t* iters,
                                     double* berr, const nmx_batch_cfg* cfg_in) {
  nmx_batch_cfg cfg;
  NMX_TRY(batched_args(ctx, count, n, offA, offB, kind, A, a_elems, B, b_elems, X, info, iters, berr, cfg_in, &cfg));
  if (count == 0) return NMX_OK;
  // the context's device copy of the static scratch array: never grown
  if (!ctx->static_scratch || batched_ws(count, a_elems).total() > ctx->static_scratch_bytes)
    return NMX_ERR_UNSUPPORTED;
  // one stream: creating the second one on first use would allocate
  return batched_run(ctx, count, n, offA, offB, kind, A, a_elems, B, b_elems, X, info, iters, berr, cfg,
                     ctx->static_scratch, false);
}

// ---- autotuner: resident workgroups per CU, per size class ---------------------
// For each of the three launch classes (n > 128, 64 < n <= 128, n <= 64) a
// synthetic batch of `samples` problems of that class (n uniform in the
// class, half LU with a diagonal shift, half Cholesky on a diagonally
// dominant symmetric matrix) is solved once per candidate value and timed
// with events; the fastest value goes into cfg->teams_per_cu.  The other
// fields of *cfg are kept (tol, max_iter shape the refinement being timed).
NMX_API int nmx_batch_autotune(nmx_ctx* ctx, int samples, nmx_batch_cfg* cfg) {
  if (!ctx) return -1;
  if (samples < 1 || samples > 1000000) return -2;
  if (!cfg) return -3;
  if (ctx->capturing) return NMX_ERR_UNSUPPORTED;
  static const int kCand[] = {1, 2, 3, 4, 6, 8};
  const int lo[3] = {129, 65, 1}, hi[3] = {256, 128, 64};
  uint64_t state = 0x9E3779B97F4A7C15ull;
  auto next = [&state]() {  // splitmix64
    uint64_t z = (state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  auto unif = [&next]() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0; };
  int status = NMX_OK;
  for (int cls = 0; cls < 3 && status == NMX_OK; cls++) {
    std::vector<int32_t> n(samples);
    std::vector<uint8_t> kind(samples);
    std::vector<int64_t> oa(samples), ob(samples);
    int64_t ae = 0, be = 0;
    for (int p = 0; p < samples; p++) {
      n[p] = lo[cls] + (int)(next() % (uint64_t)(hi[cls] - lo[cls] + 1));
      kind[p] = (uint8_t)(p & 1);
      oa[p] = ae;
      ob[p] = be;
      ae += (int64_t)n[p] * n[p];
      be += n[p];
    }
    std::vector<double> a(ae), b(be);
    for (int p = 0; p < samples; p++) {
      const int m = n[p];
      double* ap = a.data() + oa[p];
      for (int i = 0; i < m; i++)
        for (int j = 0; j < m; j++) ap[(int64_t)i * m + j] = unif();
      if (kind[p])
        for (int i = 0; i < m; i++)
          for (int j = 0; j < i; j++) ap[(int64_t)j * m + i] = ap[(int64_t)i * m + j];
      for (int i = 0; i < m; i++) ap[(int64_t)i * m + i] += kind[p] ? (double)m : 0.1 * m;
      for (int i = 0; i < m; i++) b[ob[p] + i] = unif();
    }
    void* mem = nullptr;
    const size_t bytes = 4 * (size_t)samples * 3 + 8 * (size_t)samples * 3 + (size_t)samples + 8 * (size_t)(ae + 2 * be) +
                         1024;
    if (hipMalloc(&mem, bytes) != hipSuccess) return NMX_ERR_ALLOC;
    char* q = (char*)mem;
    auto carve_b = [&q](size_t sz) {
      char* r = q;
      q += (sz + 255) & ~(size_t)255;
      return (void*)r;
    };
    int32_t* dn = (int32_t*)carve_b(4 * (size_t)samples);
    int32_t* dinfo = (int32_t*)carve_b(4 * (size_t)samples);
    int32_t* diters = (int32_t*)carve_b(4 * (size_t)samples);
    int64_t* doa = (int64_t*)carve_b(8 * (size_t)samples);
    int64_t* dob = (int64_t*)carve_b(8 * (size_t)samples);
    double* dberr = (double*)carve_b(8 * (size_t)samples);
    uint8_t* dk = (uint8_t*)carve_b((size_t)samples);
    double* dA = (double*)carve_b(8 * (size_t)ae);
    double* dB = (double*)carve_b(8 * (size_t)be);
    double* dX = (double*)carve_b(8 * (size_t)be);
    hipStream_t s = ctx->stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipMemcpyAsync(dn, n.data(), 4 * (size_t)samples, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(doa, oa.data(), 8 * (size_t)samples, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(dob, ob.data(), 8 * (size_t)samples, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(dk, kind.data(), (size_t)samples, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(dA, a.data(), 8 * (size_t)ae, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(dB, b.data(), 8 * (size_t)be, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
      status = NMX_ERR_HIP;
    float best_ms = INFINITY;
    int best = cfg->teams_per_cu[cls];
    for (int ci = 0; ci < (int)(sizeof(kCand) / sizeof(kCand[0])) && status == NMX_OK; ci++) {
      nmx_batch_cfg c = *cfg;
      c.teams_per_cu[cls] = kCand[ci];
      float tot = 0.f;
      for (int rep = 0; rep < 3 && status == NMX_OK; rep++) {  // the first run is a warm-up
        if (hipEventRecord(e0, s) != hipSuccess) status = NMX_ERR_HIP;
        if (status == NMX_OK)
          status = nmx_batched_gesv(ctx, samples, dn, doa, dob, dk, dA, ae, dB, be, dX, dinfo, diters, dberr, &c);
        if (status == NMX_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess))
          status = NMX_ERR_HIP;
        float ms = 0.f;
        if (status == NMX_OK && rep > 0) {
          hipEventElapsedTime(&ms, e0, e1);
          tot += ms;
        }
      }
      if (status == NMX_OK && tot < best_ms) {
        best_ms = tot;
        best = kCand[ci];
      }
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    hipStreamSynchronize(s);
    hipFree(mem);
    if (status == NMX_OK) cfg->teams_per_cu[cls] = best;
  }
  if (status != NMX_OK) return status;
  return nmx_ctx_check(ctx);
}
