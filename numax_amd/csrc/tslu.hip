// Tournament-pivoted leaf (TSLU, the panel of communication-avoiding LU:
// Grigori, Demmel & Xiang, "CALU: a communication optimal LU factorization
// algorithm", SIMAX 32(4), 2011) for the mixed-precision factorization.
//
// The exact leaf (getf2_leaf) makes one pivot decision per column across
// every workgroup that holds a part of the panel: 16384 decisions in
// sequence at n = 16384, each a cross-workgroup exchange.  The
// mixed-precision factor only has to be a good preconditioner, not numax's
// pivot sequence, so its leaf picks all w pivot rows of a leaf at once:
//
//   1. every wavefront runs partial pivoting on its own 256-row chunk of the
//      m x w panel (rows in registers, no exchange) and proposes the w rows
//      it chose;
//   2. the proposals are merged eight sets at a time -- the last wavefront
//      of a group to finish (an atomic counter, no waiting) reloads the
//      original values of the group's 256 candidate rows and runs partial
//      pivoting on them -- until one set is left;
//   3. the final merge's pivoting gives the pivot order and, from the
//      selected rows' eliminated values, L11 \ U11; it also works out the
//      LAPACK interchanges that bring the selected rows to the top;
//   4. the apply kernel writes L11 \ U11 on top and L21 = A21 U11^-1 below,
//      one row per thread.
//
// Rows are distinct candidates and numbers stay FP32.  What differs from
// exact partial pivoting is only which rows become pivots: the tournament's
// winner of a column is the largest entry among the candidates, not among
// all rows.  The growth this allows is bounded like GEPP's in practice
// (CALU paper, section 4).  The plain FP64 path never uses this leaf.
#include <utility>

#include "nmx_internal.h"

namespace {

constexpr int kTsW = kLeafW;             // leaf columns at most (32)
constexpr int kTsR = 4;                  // rows per lane
constexpr int kTsRows = 64 * kTsR;       // rows one wavefront holds: a chunk or a merge (256)
constexpr int kTsFan = kTsRows / kTsW;   // candidate sets one merge takes (8)
constexpr float kTsFloor = 1e-30f;       // an exact-zero pivot is divided as this (numax/linalg/common.mojo:27-29)

// Pivot key of a live row: magnitude above NaN above nothing, ties to the
// lower slot.  The low 8 bits of |x|'s pattern are replaced by the slot
// (2^-15 relative resolution: near-equal candidates count as equal).
__device__ __forceinline__ uint32_t ts_key(float x, int slot) {
  const uint32_t b = __float_as_uint(x) & 0x7FFFFFFFu;
  const uint32_t mag = b > 0x7F800000u ? 1u : (b >> 8) + 2u;
  return (mag << 8) | (uint32_t)(255 - slot);
}

// Partial pivoting over the wavefront's rows: lane `lane`, register row r is
// slot r * 64 + lane, original row rid[r] (-1: none).  Selects up to w rows;
// sel (lane j < w) receives the row chosen at step j, -1 if none was left.
// The selected row's values at selection (its multipliers, then its U row)
// are written to rows[j] (LDS) -- the final merge keeps them as L11 \ U11.
// FINAL also simulates the LAPACK interchanges and finds the first exactly
// zero pivot.
struct FinalState {
  int tc;        // lane t < w: original row now at top position r0 + t
  int bp, bc;    // a position below the top whose content changed, and that content
  int q;         // lane j: ipiv[r0 + j]
  int first_zero;
};

// One column step of the wavefront's partial pivoting; J is a compile-time
// column, so every register index is static (a loop the compiler might not
// unroll would put the row tile in scratch memory).
struct GeppState {
  int live;   // bit r: register row r is a candidate still
  int sel;    // lane j: row chosen at step j
  bool done;  // uniform: w columns taken, or no candidate left in this set
};

template <bool FINAL, int J>
__device__ __forceinline__ void gepp_step(float (&v)[kTsR][kTsW], const int (&rid)[kTsR], int w, int lane,
                                          float (*rows)[kTsW], int r0, GeppState& g, FinalState& fs) {
  uint32_t k = 0;
#pragma unroll
  for (int r = 0; r < kTsR; r++) k = max(k, ((g.live >> r) & 1) ? ts_key(v[r][J], r * 64 + lane) : 0u);
  const uint32_t kmax = wave_umax(k);
  g.done = g.done || J >= w || kmax == 0;
  if (g.done) return;
  const int slot = 255 - (int)(kmax & 0xFFu);
  const int wl = slot & 63, wr = slot >> 6;
  // the pivot lane hands its row over through LDS and retires it
  if (lane == wl) {
#pragma unroll
    for (int r = 0; r < kTsR; r++) {
      if (r == wr) {
#pragma unroll
        for (int c = 0; c < kTsW; c++) rows[J][c] = v[r][c];
        // ends each row's branch differently, so the branches are not merged
        // into one store from a pointer chosen by wr (which would put v in
        // scratch memory)
        asm volatile("; pivot row %0" ::"n"(r));
      }
    }
    g.live &= ~(1 << wr);
  }
  const int prow_id = wr == 0   ? __builtin_amdgcn_readlane(rid[0], wl)
                      : wr == 1 ? __builtin_amdgcn_readlane(rid[1], wl)
                      : wr == 2 ? __builtin_amdgcn_readlane(rid[2], wl)
                                : __builtin_amdgcn_readlane(rid[3], wl);
  if (lane == J) g.sel = prow_id;
  // the pivot lane's LDS stores land before any lane reads the row
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  const float p = rows[J][J];
  const float rinv = 1.0f / (p == 0.0f ? kTsFloor : p);
  float l[kTsR];
#pragma unroll
  for (int r = 0; r < kTsR; r++) l[r] = ((g.live >> r) & 1) ? v[r][J] * rinv : 0.0f;
#pragma unroll
  for (int c = J + 1; c < kTsW; c++) {
    const float u = rows[J][c];
#pragma unroll
    for (int r = 0; r < kTsR; r++) v[r][c] = fmaf(-l[r], u, v[r][c]);
  }
  if constexpr (FINAL) {
    if (p == 0.0f && fs.first_zero < 0) fs.first_zero = J;
    // LAPACK interchange J: bring row prow_id to position r0 + J
    const unsigned long long tm = __ballot(lane >= J && lane < w && fs.tc == prow_id);
    const unsigned long long bm = __ballot(fs.bp >= 0 && fs.bc == prow_id);
    const int cj = __builtin_amdgcn_readlane(fs.tc, J);
    int q;
    if (tm) {
      const int t = __ffsll((long long)tm) - 1;
      q = r0 + t;
      if (lane == t) fs.tc = cj;
    } else if (bm) {
      const int t = __ffsll((long long)bm) - 1;
      q = __builtin_amdgcn_readlane(fs.bp, t);
      if (lane == t) fs.bc = cj;
    } else {
      q = prow_id;
      if (lane == J) {
        fs.bp = q;
        fs.bc = cj;
      }
    }
    if (lane == J) {
      fs.tc = prow_id;
      fs.q = q;
    }
  }
}

template <bool FINAL, int... Js>
__device__ __forceinline__ int gepp_columns(std::integer_sequence<int, Js...>, float (&v)[kTsR][kTsW],
                                            const int (&rid)[kTsR], int w, int lane, float (*rows)[kTsW], int r0,
                                            FinalState& fs) {
  GeppState g{0, -1, false};
#pragma unroll
  for (int r = 0; r < kTsR; r++) g.live |= (rid[r] >= 0) << r;
  (gepp_step<FINAL, Js>(v, rid, w, lane, rows, r0, g, fs), ...);
  return g.sel;
}

template <bool FINAL>
__device__ __forceinline__ int wave_gepp(float (&v)[kTsR][kTsW], const int (&rid)[kTsR], int w, int lane,
                                         float (*rows)[kTsW], int r0, FinalState& fs) {
  return gepp_columns<FINAL>(std::make_integer_sequence<int, kTsW>(), v, rid, w, lane, rows, r0, fs);
}

struct TsluOut {
  float lu[kTsW][kTsW];   // L11 \ U11 of the selected rows, in pivot order
  float top[kTsW][kTsW];  // original values of the panel's first w rows
  int32_t move_pos[kTsW]; // rows below the top that receive a displaced top row
  int32_t move_src[kTsW]; // ... which one (0 .. w-1)
  int32_t nmoves;
};

// sets (or merge groups) of the next level up
__host__ __device__ __forceinline__ int next_count(int s) { return (s + kTsFan - 1) / kTsFan; }

__global__ void __launch_bounds__(64) tslu_tournament_kernel(const float* __restrict__ A, int lda, int n, int r0,
                                                            int c0, int w, int32_t* __restrict__ counters,
                                                            int32_t* __restrict__ cand, TsluOut* __restrict__ out,
                                                            int32_t* __restrict__ ipiv, int32_t* __restrict__ info) {
  __shared__ float rows[kTsW][kTsW];
  const int lane = threadIdx.x;
  float v[kTsR][kTsW];
  int rid[kTsR];
  FinalState fs{r0 + lane, -1, -1, 0, -1};

  // the final merge also keeps the original first w rows (its apply step
  // overwrites them) and writes everything the apply kernel needs
  auto finish = [&]() {
    if (lane < w) {
      const float* src = A + (int64_t)(r0 + lane) * lda + c0;
#pragma unroll
      for (int c = 0; c < kTsW; c++) out->top[lane][c] = c < w ? src[c] : 0.0f;
#pragma unroll
      for (int c = 0; c < kTsW; c++) out->lu[lane][c] = rows[lane][c];
      ipiv[r0 + lane] = fs.q;
    }
    const unsigned long long mv = __ballot(fs.bp >= 0);
    if (fs.bp >= 0) {
      const int idx = __popcll(mv & ((1ull << lane) - 1ull));
      out->move_pos[idx] = fs.bp;
      out->move_src[idx] = fs.bc - r0;
    }
    if (lane == 0) {
      out->nmoves = __popcll(mv);
      if (fs.first_zero >= 0 && *info == 0) *info = c0 + fs.first_zero + 1;
    }
  };

  // Level 0 is this wavefront's chunk of rows; every later level is the
  // merge of up to kTsFan sets of the level below, done by the last member
  // of the group to arrive.  `cnt` sets make up the current level, which
  // starts at set index `base` (counters of its groups at `ctr`); this
  // wavefront's rows produce its set `my`.
  int cnt = (n - r0 + kTsRows - 1) / kTsRows, base = 0, ctr = 0, my = blockIdx.x;
#pragma unroll
  for (int r = 0; r < kTsR; r++) {
    const int row = r0 + my * kTsRows + r * 64 + lane;
    rid[r] = row < n ? row : -1;
  }
  for (;;) {
#pragma unroll
    for (int r = 0; r < kTsR; r++) {
      const float* src = A + (int64_t)(rid[r] < 0 ? r0 : rid[r]) * lda + c0;
#pragma unroll
      for (int c = 0; c < kTsW; c++) v[r][c] = (rid[r] >= 0 && c < w) ? src[c] : 0.0f;
    }
    if (cnt == 1) {
      wave_gepp<true>(v, rid, w, lane, rows, r0, fs);
      finish();
      return;
    }
    const int sel = wave_gepp<false>(v, rid, w, lane, rows, r0, fs);
    if (lane < kTsW) cand[(int64_t)(base + my) * kTsW + lane] = sel;
    const int groups = next_count(cnt);
    const int grp = my / kTsFan;
    const int members = min(kTsFan, cnt - grp * kTsFan);
    __syncthreads();  // this wavefront's candidate stores are issued before its arrival
    int old = 0;
    if (lane == 0) old = __hip_atomic_fetch_add(&counters[ctr + grp], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    old = __builtin_amdgcn_readfirstlane(old);
    if (old != members - 1) return;  // another member merges this group
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    // the group's candidates: slot q is set q / 32, entry q % 32
    const int32_t* gs = cand + (int64_t)(base + grp * kTsFan) * kTsW;
#pragma unroll
    for (int r = 0; r < kTsR; r++) {
      const int q = r * 64 + lane;
      rid[r] = q / kTsW < members ? __hip_atomic_load(gs + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : -1;
    }
    base += cnt;
    ctr += groups;
    cnt = groups;
    my = grp;
  }
}

// Writes L11 \ U11 over the first w rows and L21 = A21 U11^-1 below them
// (a displaced top row's values come from the saved copy), one row per
// thread; block 0 also clears the merge counters for the next leaf.
constexpr int kApplyT = 256;
// x := x U^-1 (U upper, 1 / diagonal in dinv), column K at a time
template <int K>
__device__ __forceinline__ void subst_step(float (&x)[kTsW], const float (*U)[kTsW], const float* dinv) {
  x[K] *= dinv[K];
#pragma unroll
  for (int c = K + 1; c < kTsW; c++) x[c] = fmaf(-x[K], U[K][c], x[c]);
  __builtin_amdgcn_sched_barrier(0);  // row K of U is read at step K, not all of U up front
}
template <int... Ks>
__device__ __forceinline__ void subst_columns(std::integer_sequence<int, Ks...>, float (&x)[kTsW],
                                              const float (*U)[kTsW], const float* dinv) {
  (subst_step<Ks>(x, U, dinv), ...);
}
__global__ void __launch_bounds__(kApplyT) tslu_apply_kernel(float* __restrict__ A, int lda, int n, int r0, int c0,
                                                            int w, const TsluOut* __restrict__ out,
                                                            int32_t* __restrict__ counters, int ncounters) {
  __shared__ __attribute__((aligned(16))) float U[kTsW][kTsW];
  __shared__ float dinv[kTsW];
  __shared__ int mpos[kTsW], msrc[kTsW];
  __shared__ int nm;
  const int t = threadIdx.x;
  for (int e = t; e < kTsW * kTsW; e += kApplyT) U[e / kTsW][e % kTsW] = out->lu[e / kTsW][e % kTsW];
  if (t < kTsW) {
    const float d = t < w ? out->lu[t][t] : 1.0f;
    dinv[t] = 1.0f / (d == 0.0f ? kTsFloor : d);
    mpos[t] = out->move_pos[t];
    msrc[t] = out->move_src[t];
  }
  if (t == 0) nm = out->nmoves;
  if (blockIdx.x == 0)
    for (int e = t; e < ncounters; e += kApplyT) counters[e] = 0;
  __syncthreads();
  const int i0 = r0 + blockIdx.x * kApplyT;
  const int i = i0 + t;
  // this block's rows through one descriptor: columns past w and rows past
  // n read as zero and their stores are dropped
  const __amdgpu_buffer_rsrc_t rs = buf_rsrc(A + (int64_t)i0 * lda + c0);
  const bool own = i < n;
  auto off = [&](int c) -> uint32_t {
    return (own && c < w) ? ((uint32_t)t * (uint32_t)lda + (uint32_t)c) * 4u : kBufOob;
  };
  float x[kTsW];
  if (i < r0 + w) {
#pragma unroll
    for (int c = 0; c < kTsW; c++) buf_store(rs, off(c), U[i - r0][c]);
    return;
  }
#pragma unroll
  for (int c = 0; c < kTsW; c++) x[c] = buf_load<float>(rs, off(c));
  int src = -1;
  for (int mi = 0; mi < nm; mi++)
    if (mpos[mi] == i) src = msrc[mi];
  if (src >= 0) {  // a displaced top row: its original values were saved
#pragma unroll
    for (int c = 0; c < kTsW; c++) x[c] = c < w ? out->top[src][c] : 0.0f;
  }
  subst_columns(std::make_integer_sequence<int, kTsW>(), x, U, dinv);
#pragma unroll
  for (int c = 0; c < kTsW; c++) buf_store(rs, off(c), x[c]);
}

}  // namespace

// Scratch of the tournament: merge counters (zero between leaves), the
// candidate sets of every level, the final merge's output.
constexpr int kTsluMaxCounters = 4096;
size_t tslu_scratch_bytes(int n) {
  const int G = ceil_div(n, kTsRows);
  size_t sets = 0;
  for (int s = G; s >= 1; s = s == 1 ? 0 : next_count(s)) sets += (size_t)s;
  return 4 * (size_t)kTsluMaxCounters + 4 * sets * kTsW + sizeof(TsluOut) + 1024;
}

int tslu_leaf_launch(nmx_ctx* ctx, void* scratch, size_t scratch_bytes, int n, float* A, int lda, int r0, int c0,
                     int w, int32_t* ipiv, int32_t* info) {
  const int m = n - r0;
  if (m <= 0 || w <= 0) return NMX_OK;
  if (w > kTsW || m < w) return NMX_ERR_UNSUPPORTED;
  if (tslu_scratch_bytes(n) > scratch_bytes) return NMX_ERR_UNSUPPORTED;
  const int G = ceil_div(m, kTsRows);
  int ncounters = 0;
  for (int s = G; s > 1; s = next_count(s)) ncounters += next_count(s);
  if (ncounters > kTsluMaxCounters) return NMX_ERR_UNSUPPORTED;
  int32_t* counters = (int32_t*)scratch;
  TsluOut* out = (TsluOut*)((char*)scratch + 4 * (size_t)kTsluMaxCounters);
  int32_t* cand = (int32_t*)((char*)out + ((sizeof(TsluOut) + 255) & ~(size_t)255));
  {
    ProfScope ps(ctx, K_LEAF, (double)m * w * w, 4.0 * (double)m * w);
    tslu_tournament_kernel<<<G, 64, 0, ctx->stream>>>(A, lda, n, r0, c0, w, counters, cand, out, ipiv, info);
    NMX_HIP(ctx, hipGetLastError());
  }
  {
    ProfScope ps(ctx, K_MP_PANEL, (double)m * w * w, 8.0 * (double)m * w);
    tslu_apply_kernel<<<ceil_div(m, kApplyT), kApplyT, 0, ctx->stream>>>(A, lda, n, r0, c0, w, out, counters,
                                                                       ncounters);
    NMX_HIP(ctx, hipGetLastError());
  }
  return laswp_launch<float>(ctx, A, lda, 0, n, c0, c0 + w, ipiv, r0, r0 + w);
}

// Clears the merge counters once, when the scratch is allocated.
__global__ void tslu_zero_kernel(int32_t* p, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0;
}
int tslu_scratch_init(nmx_ctx* ctx, void* scratch) {
  tslu_zero_kernel<<<ceil_div(kTsluMaxCounters, 256), 256, 0, ctx->stream>>>((int32_t*)scratch, kTsluMaxCounters);
  NMX_HIP(ctx, hipGetLastError());
  return NMX_OK;
}
