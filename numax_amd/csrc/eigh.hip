// Batched symmetric eigensolver for small matrices (n <= 64), FP32: the
// register-tier `eigh` that numax's `map` runs one matrix per GPU thread
// (numax/linalg/_array/eigen.mojo:316-392, rotation :34-75), one wavefront
// segment per matrix here.
//
// numax sweeps the pairs (p, q) in cyclic-by-row order, one rotation at a
// time.  Rotations on disjoint pairs commute, so here every step rotates
// NP/2 disjoint neighbour pairs at once -- (0,1),(2,3),... on even steps,
// (1,2),(3,4),... on odd ones -- and each rotation also exchanges its two
// indices.  With the exchange every index travels one way through the
// ordering, so NP steps meet every pair exactly once (the odd-even
// transposition network), and all register indices stay static.  The
// rotation itself is numax's (_jacobi_rotation): t = sign(d) 2 a_pq /
// guard(|d| + sqrt(d^2 + 4 a_pq^2)), c = 1 / sqrt(1 + t^2), s = t c, with
// d = a_qq - a_pp, and the same fixed number of sweeps (12 by default, no
// convergence test).  The diagonal is carried separately (a_pp - t a_pq,
// a_qq + t a_pq), which is where the eigenvalues are read.  A different
// rotation order changes rounding, not the converged result: eigenvalues are
// compared sorted against the FP64 oracle.
//
// Layout: lane L of a segment of NP lanes holds column L of A and of V in
// registers (NP = 32 packs two matrices per wavefront, NP = 64 one); rows
// are register indices.  A matrix of order n < NP is padded with zero rows
// and columns, which only ever rotate by exact exchanges (a_pq = 0).
#include <math.h>

#include "nmx_internal.h"

namespace {

constexpr float kFloorF = 1e-30f;  // numax/linalg/common.mojo:27-29

__device__ __forceinline__ float guard_f(float x) {
  const float m = fabsf(x);
  return (m >= kFloorF || m != m) ? x : copysignf(kFloorF, x);
}

// (m ? a : b) for a wave-uniform 64-bit mask, as an opaque v_cndmask (a plain
// ?: between two register-array elements can be folded into an indexed load,
// which would move the array to scratch memory).
__device__ __forceinline__ float sel_m(uint64_t m, float a, float b) {
  float r;
  asm("v_cndmask_b32_e64 %0, %1, %2, %3" : "=v"(r) : "v"(b), "v"(a), "s"(m));
  return r;
}

// reg[idx(lane)] where idx's bit b is set exactly on the lanes of mask m[b]:
// a select tree whose masks are constants of the lane pattern.
template <int NP>
__device__ __forceinline__ float pick_lane(const float (&r)[NP], const uint64_t (&m)[6]) {
  float a[32], b[16], c[8], d[4], e[2];
  if constexpr (NP == 64) {
#pragma unroll
    for (int i = 0; i < 32; i++) a[i] = sel_m(m[0], r[2 * i + 1], r[2 * i]);
#pragma unroll
    for (int i = 0; i < 16; i++) b[i] = sel_m(m[1], a[2 * i + 1], a[2 * i]);
#pragma unroll
    for (int i = 0; i < 8; i++) c[i] = sel_m(m[2], b[2 * i + 1], b[2 * i]);
#pragma unroll
    for (int i = 0; i < 4; i++) d[i] = sel_m(m[3], c[2 * i + 1], c[2 * i]);
#pragma unroll
    for (int i = 0; i < 2; i++) e[i] = sel_m(m[4], d[2 * i + 1], d[2 * i]);
    return sel_m(m[5], e[1], e[0]);
  } else {
#pragma unroll
    for (int i = 0; i < 16; i++) b[i] = sel_m(m[0], r[2 * i + 1], r[2 * i]);
#pragma unroll
    for (int i = 0; i < 8; i++) c[i] = sel_m(m[1], b[2 * i + 1], b[2 * i]);
#pragma unroll
    for (int i = 0; i < 4; i++) d[i] = sel_m(m[2], c[2 * i + 1], c[2 * i]);
#pragma unroll
    for (int i = 0; i < 2; i++) e[i] = sel_m(m[3], d[2 * i + 1], d[2 * i]);
    (void)a;
    return sel_m(m[4], e[1], e[0]);
  }
}

// One step of the ordering: rotate the pairs (OFF, OFF+1), (OFF+2, OFF+3),
// ... of every segment, exchanging the two indices of each pair.
template <int NP, int OFF>
__device__ __forceinline__ void jacobi_step(float (&a)[NP], float (&v)[NP], float& d, int& player, int lane, int L,
                                            int base, const uint64_t (&mnext)[6]) {
  const bool isP = (((L - OFF) & 1) == 0) && (L >= OFF) && (L + 1 < NP);
  const bool isQ = (((L - OFF) & 1) == 1) && (L >= 1 + OFF);
  const int partner = isP ? lane + 1 : (isQ ? lane - 1 : lane);
  // the pair's rotation, computed on its first lane (_array/eigen.mojo:34-75)
  const float apq_own = pick_lane<NP>(a, mnext);  // a[L+1][L]: a_pq on a first lane
  const float d_next = __shfl(d, lane + 1 < 64 ? lane + 1 : lane);
  float c, s, t;
  {
    const float twice = 2.f * apq_own;
    const float den = d_next - d;
    const float radius = sqrtf(den * den + twice * twice);
    const float safe = guard_f(fabsf(den) + radius);
    t = copysignf(1.f, den) * twice / safe;
    c = 1.f / sqrtf(1.f + t * t);
    s = t * c;
  }
  const float c_o = __shfl(c, partner), s_o = __shfl(s, partner), t_o = __shfl(t, partner);
  const float apq_o = __shfl(apq_own, partner), d_o = __shfl(d, partner);
  if (isQ) {
    c = c_o;
    s = s_o;
    t = t_o;
  }
  const float apq = isQ ? apq_o : apq_own;
  // diagonal after the rotation and the exchange
  const float d_new = isP ? d_next + t * apq : (isQ ? d_o - t * apq : d);
  // rows: register pairs (r, r+1), r = OFF, OFF+2, ... (every column)
#pragma unroll
  for (int r = OFF; r + 1 < NP; r += 2) {
    const float cr = __shfl(c, base + r), sr = __shfl(s, base + r);
    const float x = a[r], y = a[r + 1];
    a[r] = sr * x + cr * y;
    a[r + 1] = cr * x - sr * y;
  }
  // columns of A and V (with the exchange): new = alpha own + beta partner's
  const float alpha = isP ? s : (isQ ? -s : 1.f);
  const float beta = (isP || isQ) ? c : 0.f;
#pragma unroll
  for (int i = 0; i < NP; i++) {
    const float oa = __shfl(a[i], partner);
    const float ov = __shfl(v[i], partner);
    a[i] = alpha * a[i] + beta * oa;
    v[i] = alpha * v[i] + beta * ov;
  }
  d = d_new;
  player = __shfl(player, partner);
}

struct EighArgs {
  int count;
  const int32_t* n;
  const int64_t* offA;  // matrix (and vectors) element offset
  const int64_t* offW;  // eigenvalue offset
  const float* A;
  float* W;
  float* V;             // nullable
  int sweeps;
  const int32_t* order;
  const int32_t* cls;   // [0] begin / [1] count for NP = 32, [2] / [3] for NP = 64
  int32_t* head;
};

template <int NP>
__global__ void __launch_bounds__(256) syevj_kernel(EighArgs args, int cls) {
  constexpr int PPW = 64 / NP;  // matrices per wavefront
  const int lane = threadIdx.x & 63;
  const int L = lane & (NP - 1);       // slot within the segment
  const int base = lane & ~(NP - 1);   // first lane of the segment
  const int seg = lane / NP;
  const int begin = args.cls[2 * cls], count = args.cls[2 * cls + 1];

  // masks of the lanes whose (L + 1) mod NP has bit b set (for a_{L,L+1})
  // and whose L has bit b set (for the initial diagonal)
  uint64_t mnext[6], mself[6];
#pragma unroll
  for (int b = 0; b < 6; b++) {
    mnext[b] = __ballot((((L + 1) & (NP - 1)) >> b) & 1);
    mself[b] = __ballot((L >> b) & 1);
  }

  for (;;) {
    int slot0 = 0;
    if (lane == 0) slot0 = atomicAdd(&args.head[cls], PPW);
    slot0 = __shfl(slot0, 0);
    if (slot0 >= count) break;
    const int slot = slot0 + seg;
    const bool live = slot < count;
    const int p = live ? args.order[begin + slot] : 0;
    const int n = live ? args.n[p] : 0;
    const float* A = args.A + (live ? args.offA[p] : 0);
    (void)seg;

    float a[NP], v[NP];
#pragma unroll
    for (int i = 0; i < NP; i++) {
      a[i] = (i < n && L < n) ? A[(int64_t)i * n + L] : 0.f;
      v[i] = (i == L) ? 1.f : 0.f;
    }
    float d = pick_lane<NP>(a, mself);
    int player = L;

    for (int sw = 0; sw < args.sweeps; sw++) {
#pragma unroll 1
      for (int step = 0; step < NP; step += 2) {
        jacobi_step<NP, 0>(a, v, d, player, lane, L, base, mnext);
        jacobi_step<NP, 1>(a, v, d, player, lane, L, base, mnext);
      }
    }
    // padding players (>= n) only ever exchanged; each real player's value
    // and vector go to its own index
    if (live && player < n) {
      args.W[args.offW[p] + player] = d;
      if (args.V) {
        float* V = args.V + args.offA[p];
#pragma unroll
        for (int i = 0; i < NP; i++)
          if (i < n) V[(int64_t)i * n + player] = v[i];
      }
    }
  }
}

// A matrix is taken when 1 <= n <= 64 and its matrix and eigenvalues lie
// inside the arrays the caller sized; others are skipped.
__device__ __forceinline__ bool eigh_ok(int nn, int64_t oa, int64_t ow, int64_t a_elems, int64_t w_elems) {
  return nn >= 1 && nn <= 64 && oa >= 0 && oa + (int64_t)nn * nn <= a_elems && ow >= 0 && ow + nn <= w_elems;
}

__global__ void eigh_sort_kernel(int count, const int32_t* __restrict__ n, const int64_t* __restrict__ offA,
                                 const int64_t* __restrict__ offW, int64_t a_elems, int64_t w_elems,
                                 int32_t* __restrict__ cls, int32_t* __restrict__ order, int32_t* __restrict__ head) {
  // one workgroup: counts, then a stable scatter by class (n <= 32 first)
  __shared__ int cnt[2];
  const int t = threadIdx.x;
  if (t < 2) cnt[t] = 0;
  __syncthreads();
  for (int i = t; i < count; i += blockDim.x) {
    const int nn = n[i];
    if (eigh_ok(nn, offA[i], offW[i], a_elems, w_elems)) atomicAdd(&cnt[nn > 32], 1);
  }
  __syncthreads();
  if (t == 0) {
    cls[0] = 0;
    cls[1] = cnt[0];
    cls[2] = cnt[0];
    cls[3] = cnt[1];
    head[0] = 0;
    head[1] = 0;
    cnt[0] = 0;
    cnt[1] = 0;
  }
  __syncthreads();
  for (int i = t; i < count; i += blockDim.x) {
    const int nn = n[i];
    if (eigh_ok(nn, offA[i], offW[i], a_elems, w_elems)) {
      const int c = nn > 32;
      order[cls[2 * c] + atomicAdd(&cnt[c], 1)] = i;
    }
  }
}

}  // namespace

NMX_API int nmx_batched_syevj_f32(nmx_ctx* ctx, int count, const int32_t* n, const int64_t* offA,
                                  const int64_t* offW, const float* A, int64_t a_elems, float* W, int64_t w_elems,
                                  float* V, int sweeps) {
  if (!ctx) return -1;
  if (count < 0) return -2;
  if (count == 0) return NMX_OK;
  if (!n) return -3;
  if (!offA) return -4;
  if (!offW) return -5;
  if (!A) return -6;
  if (a_elems <= 0) return -7;
  if (!W) return -8;
  if (w_elems <= 0) return -9;
  if (sweeps < 0) return -11;
  void* ws = nullptr;
  NMX_TRY(nmx_workspace(ctx, sizeof(int32_t) * (8 + (size_t)count), &ws));
  int32_t* cls = (int32_t*)ws;
  int32_t* head = cls + 4;
  int32_t* order = cls + 8;
  {
    ProfScope ps(ctx, K_OTHER, 0.0, 0.0);
    eigh_sort_kernel<<<1, 1024, 0, ctx->stream>>>(count, n, offA, offW, a_elems, w_elems, cls, order, head);
    NMX_HIP(ctx, hipGetLastError());
  }
  EighArgs a{count, n, offA, offW, A, W, V, sweeps, order, cls, head};
  {
    ProfScope ps(ctx, K_EIGH, 0.0, 0.0);
    const int grid = ctx->num_cus * 4;
    syevj_kernel<64><<<grid, 256, 0, ctx->stream>>>(a, 1);
    syevj_kernel<32><<<grid, 256, 0, ctx->stream>>>(a, 0);
    NMX_HIP(ctx, hipGetLastError());
  }
  return NMX_OK;
}
