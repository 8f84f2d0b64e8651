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
//   3. the final merge's pivoting gives the pivot order and U11 (the pivot
//      rows' eliminated values); it also works out the LAPACK interchanges
//      that bring the selected rows to the top;
//   4. the apply kernel writes L11 \ U11 on top and L21 = A21 U11^-1 below,
//      one row per thread, L11 recomputed the same way from the selected
//      rows' original values.
//
// Rows are distinct candidates and numbers stay FP32.  What differs from
// exact partial pivoting is only which rows become pivots: the tournament's
// winner of a column is the largest entry among the candidates, not among
// all rows.  The growth this allows is bounded like GEPP's in practice
// (CALU paper, section 4).  The plain FP64 path never uses this leaf.
//
// One wavefront issues an FP32 FMA about every 6.4 cycles on MI355X, a DPP
// wave maximum takes ~120 cycles and an LDS round trip ~300: the column step
// keeps the pivot key to one instruction per row and broadcasts the pivot
// row by v_readlane (no LDS round trip).  Rows are loaded coalesced (eight
// rows per instruction, 16 bytes a lane) and transposed through LDS: a load
// that gathers 64 rows takes about 30x longer.
#include <utility>

#include "nmx_internal.h"

namespace {

constexpr int kTsW = kLeafW;             // leaf columns at most (32)
constexpr int kTsR = 4;                  // rows per lane
constexpr int kTsRows = 64 * kTsR;       // rows one wavefront holds: a chunk or a merge (256)
constexpr int kTsFan = kTsRows / kTsW;   // candidate sets one merge takes (8)
constexpr float kTsFloor = 1e-30f;       // an exact-zero pivot is divided as this (numax/linalg/common.mojo:27-29)
constexpr int kStgLd = kTsW + 4;         // LDS staging row stride (16-byte aligned, fewer bank conflicts)

// Pivot keys: (|x|'s pattern with its low 9 bits cleared) | (256 - slot) for
// a live row, 0 for none -- one v_and_or per row.  Ties (and values within
// 2^-14 of each other) go to the lower slot; a NaN wins (the solve then
// fails its refinement, which reports it).
constexpr uint32_t kKeyMask = 0x7FFFFE00u;

struct FinalState {
  int tc;        // lane t < w: original row now at top position r0 + t
  int bp, bc;    // a position below the top whose content changed, and that content
  int q;         // lane j: ipiv[r0 + j]
  int first_zero;
};

struct GeppState {
  uint32_t km[kTsR], kc[kTsR];  // key mask / slot code of register row r (0 once retired)
  int sel;                      // lane j: row chosen at step j
  bool done;                    // uniform: w columns taken, or no candidate left in this set
};

// One step of partial pivoting over the wavefront's rows for compile-time
// column J (so every register index is static): lane `lane`, register row r
// is slot r * 64 + lane, original row rid[r] (-1: none).  FINAL (a uniform
// flag) also stores U's row J in rows[J] and simulates the LAPACK
// interchanges.
template <int J>
__device__ __forceinline__ void gepp_step(float (&v)[kTsR][kTsW], const int (&rid)[kTsR], int w, int lane,
                                          float (*rows)[kTsW], int r0, bool final, GeppState& g, FinalState& fs) {
  uint32_t k = 0;
#pragma unroll
  for (int r = 0; r < kTsR; r++) k = max(k, (__float_as_uint(v[r][J]) & g.km[r]) | g.kc[r]);
  const uint32_t kmax = wave_umax(k);
  g.done = g.done || J >= w || kmax == 0;
  if (g.done) return;
  const int slot = 256 - (int)(kmax & 0x1FFu);
  const int wl = slot & 63, wr = slot >> 6;
  // the pivot row into scalar registers (columns J..), read from lane wl
  float su[kTsW];
  int prow_id = 0;
#pragma unroll
  for (int r = 0; r < kTsR; r++) {
    if (r == wr) {
#pragma unroll
      for (int c = J; c < kTsW; c++) su[c] = readlane_f(v[r][c], wl);
      prow_id = __builtin_amdgcn_readlane(rid[r], wl);
      // ends each row's branch differently, so the branches are not merged
      // into one read from a register row chosen by wr (which would put v in
      // scratch memory)
      asm volatile("; pivot row %0" ::"n"(r));
    }
  }
  if (lane == wl) {
#pragma unroll
    for (int r = 0; r < kTsR; r++)
      if (r == wr) g.km[r] = g.kc[r] = 0u;  // retired
  }
  if (lane == J) g.sel = prow_id;
  const float p = su[J];
  const float rinv = __builtin_amdgcn_rcpf(p == 0.0f ? kTsFloor : p);
  float l[kTsR];
#pragma unroll
  for (int r = 0; r < kTsR; r++) l[r] = g.km[r] ? v[r][J] * rinv : 0.0f;
#pragma unroll
  for (int c = J + 1; c < kTsW; c++) {
#pragma unroll
    for (int r = 0; r < kTsR; r++) v[r][c] = fmaf(-l[r], su[c], v[r][c]);
  }
  if (final) {  // uniform
    if (lane == 0) {
#pragma unroll
      for (int c = J; c < kTsW; c++) rows[J][c] = su[c];
    }
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

template <int... Js>
__device__ __forceinline__ int gepp_columns(std::integer_sequence<int, Js...>, float (&v)[kTsR][kTsW],
                                            const int (&rid)[kTsR], int w, int lane, float (*rows)[kTsW], int r0,
                                            bool final, FinalState& fs) {
  GeppState g;
#pragma unroll
  for (int r = 0; r < kTsR; r++) {
    const bool live = rid[r] >= 0;
    g.km[r] = live ? kKeyMask : 0u;
    g.kc[r] = live ? (uint32_t)(256 - (r * 64 + lane)) : 0u;
  }
  g.sel = -1;
  g.done = false;
  (gepp_step<Js>(v, rid, w, lane, rows, r0, final, g, fs), ...);
  return g.sel;
}

// Partial pivoting of the wavefront's rows: sel (lane j < w) receives the
// row chosen at step j (-1: none left).  The multipliers are not kept: the
// final merge's L11 is recomputed from the selected rows' original values,
// like L21.  The pivot reciprocal is v_rcp (the multipliers here only steer
// the choice of rows).
__device__ __forceinline__ int wave_gepp(float (&v)[kTsR][kTsW], const int (&rid)[kTsR], int w, int lane,
                                         float (*rows)[kTsW], int r0, bool final, FinalState& fs) {
  return gepp_columns(std::make_integer_sequence<int, kTsW>(), v, rid, w, lane, rows, r0, final, fs);
}

// The wavefront's rows (slot r * 64 + lane holds row rid[r]; srow[g] is the
// row of slot g * 8 + lane / 8), columns c0 .. c0+w, into v (zeros for
// absent rows and columns past w).  VEC: eight rows per instruction as
// 16-byte pieces (lda, c0 and w multiples of 4, A 16-byte aligned), sixteen
// instructions in flight, transposed through the LDS staging area;
// otherwise one row per lane, every load in bounds (an absent row reads row
// r0, a column past w reads column w-1) and the unwanted values zeroed after.
template <bool VEC>
__device__ __forceinline__ void load_rows(float (&v)[kTsR][kTsW], const int (&rid)[kTsR], const int (&srow)[32],
                                          const float* A, int lda, int r0, int c0, int w, int lane,
                                          float (*stg)[kStgLd]) {
  if constexpr (VEC) {
    const int q = lane >> 3, seg = lane & 7;
    const bool segok = seg * 4 < w;
#pragma unroll
    for (int b = 0; b < 2; b++) {
      float4 x[16];
#pragma unroll
      for (int k = 0; k < 16; k++) {
        const int row = srow[b * 16 + k];
        x[k] = *((const float4*)(A + (int64_t)(row < 0 ? r0 : row) * lda + c0) + (segok ? seg : 0));
      }
#pragma unroll
      for (int k = 0; k < 16; k++) {
        const int slot = (b * 16 + k) * 8 + q;
        *(float4*)&stg[slot][seg * 4] = (srow[b * 16 + k] < 0 || !segok) ? make_float4(0.f, 0.f, 0.f, 0.f) : x[k];
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
#pragma unroll
    for (int r = 0; r < kTsR; r++) {
#pragma unroll
      for (int c = 0; c < kTsW; c += 4) {
        const float4 x = *(const float4*)&stg[r * 64 + lane][c];
        v[r][c] = x.x;
        v[r][c + 1] = x.y;
        v[r][c + 2] = x.z;
        v[r][c + 3] = x.w;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");  // the staging area is free again
  } else {
#pragma unroll
    for (int r = 0; r < kTsR; r++) {
      const float* src = A + (int64_t)(rid[r] < 0 ? r0 : rid[r]) * lda + c0;
#pragma unroll
      for (int c = 0; c < kTsW; c++) {
        const float x = src[min(c, w - 1)];
        v[r][c] = (rid[r] >= 0 && c < w) ? x : 0.0f;
      }
    }
  }
}

struct TsluOut {
  float u[kTsW][kTsW];     // U11 (row j from column j on, zeros before), in pivot order
  int32_t sel[kTsW];       // original row of pivot j
  int32_t move_pos[kTsW];  // rows below the top that receive a displaced top row
  int32_t move_src[kTsW];  // ... which one (0 .. w-1)
  int32_t nmoves;
};

// sets (or merge groups) of the next level up
__host__ __device__ __forceinline__ int next_count(int s) { return (s + kTsFan - 1) / kTsFan; }

template <bool VEC>
__global__ void __launch_bounds__(64) tslu_tournament_kernel(const float* __restrict__ A, int lda, int n, int r0,
                                                            int c0, int w, int32_t* __restrict__ counters,
                                                            int32_t* __restrict__ cand, TsluOut* __restrict__ out,
                                                            int32_t* __restrict__ ipiv, int32_t* __restrict__ info) {
  __shared__ __attribute__((aligned(16))) float rows[kTsW][kTsW];
  __shared__ __attribute__((aligned(16))) float stg[VEC ? kTsRows : 1][kStgLd];
  const int lane = threadIdx.x;
  float v[kTsR][kTsW];
  int rid[kTsR];
  int srow[32];  // the row of slot g * 8 + lane / 8 (VEC loads)
  FinalState fs{r0 + lane, -1, -1, 0, -1};

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
#pragma unroll
  for (int g = 0; g < 32; g++) {
    const int row = r0 + my * kTsRows + g * 8 + (lane >> 3);
    srow[g] = row < n ? row : -1;
  }
  for (;;) {
    load_rows<VEC>(v, rid, srow, A, lda, r0, c0, w, lane, stg);
    const bool final = cnt == 1;
    const int sel = wave_gepp(v, rid, w, lane, rows, r0, final, fs);
    if (final) {
      // U11, the pivots, the interchanges and the moves for the apply kernel
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
#pragma unroll
      for (int e = lane; e < kTsW * kTsW; e += 64) {
        const int j = e / kTsW, c = e % kTsW;
        out->u[j][c] = (c >= j && c < w && j < w) ? rows[j][c] : 0.0f;
      }
      if (lane < w) {
        out->sel[lane] = sel;
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
      return;
    }
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
    if constexpr (VEC) {
#pragma unroll
      for (int g = 0; g < 32; g++) {
        const int q = g * 8 + (lane >> 3);
        srow[g] = q / kTsW < members ? __hip_atomic_load(gs + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : -1;
      }
    }
    base += cnt;
    ctr += groups;
    cnt = groups;
    my = grp;
  }
}

// x := x U^-1 (U upper, 1 / diagonal in dinv), column K at a time
constexpr int kApplyT = 256;
template <int K>
__device__ __forceinline__ void subst_step(float (&x)[kTsW], const float (*U)[kTsW], const float* dinv) {
  x[K] *= dinv[K];
#pragma unroll
  for (int c = K + 1; c < kTsW; c++) x[c] = fmaf(-x[K], U[K][c], x[c]);
}
template <int... Ks>
__device__ __forceinline__ void subst_columns(std::integer_sequence<int, Ks...>, float (&x)[kTsW],
                                              const float (*U)[kTsW], const float* dinv) {
  (subst_step<Ks>(x, U, dinv), ...);
}

// Workgroup 0 writes every row whose contents move: the w pivot positions
// (L11 from the selected rows' original values, U11 from the merge) and the
// positions below that receive a displaced top row (its L21).  It reads
// only rows it writes itself, before writing any, so no copy is needed.
// Every other workgroup takes 256 rows of L21 = A21 U11^-1 in place (rows
// that receive a displaced top row left to workgroup 0), staged through LDS
// so loads and stores are coalesced (VEC).  Workgroup 0 also clears the
// merge counters for the next leaf.
template <bool VEC>
__global__ void __launch_bounds__(kApplyT) tslu_apply_kernel(float* __restrict__ A, int lda, int n, int r0, int c0,
                                                            int w, const TsluOut* __restrict__ out,
                                                            int32_t* __restrict__ counters, int ncounters) {
  __shared__ __attribute__((aligned(16))) float U[kTsW][kTsW];
  __shared__ float dinv[kTsW];
  __shared__ int mpos[kTsW], msrc[kTsW], psel[kTsW];
  __shared__ int nm;
  __shared__ uint32_t movedbits[kApplyT / 32];  // this workgroup's rows that workgroup 0 writes
  __shared__ __attribute__((aligned(16))) float stg[VEC ? kApplyT : 1][kStgLd];
  const int t = threadIdx.x;
  if (t < kApplyT / 32) movedbits[t] = 0u;
  for (int e = t; e < kTsW * kTsW; e += kApplyT) U[e / kTsW][e % kTsW] = out->u[e / kTsW][e % kTsW];
  if (t < kTsW) {
    const float d = t < w ? out->u[t][t] : 1.0f;
    dinv[t] = 1.0f / (d == 0.0f ? kTsFloor : d);
    mpos[t] = out->move_pos[t];
    msrc[t] = out->move_src[t];
    psel[t] = out->sel[t];
  }
  if (t == 0) nm = out->nmoves;
  const bool special = blockIdx.x == 0;
  const int i0 = r0 + w + (blockIdx.x - 1) * kApplyT;
  __syncthreads();
  if (!special && t < nm) {
    const int rel = mpos[t] - i0;
    if (rel >= 0 && rel < kApplyT) atomicOr(&movedbits[rel >> 5], 1u << (rel & 31));
  }
  __syncthreads();
  // this thread's row: its destination and where its values come from
  float x[kTsW];
  int dst = -1;
  const int seg = t & 7;
  const bool segok = seg * 4 < w;
  if (special) {
    for (int e = t; e < ncounters; e += kApplyT) counters[e] = 0;
    // thread t < w: pivot position r0 + t; w <= t < w + nm: moved position
    int src = -1;
    if (t < w) {
      dst = r0 + t;
      src = psel[t];
    } else if (t < w + nm) {
      dst = mpos[t - w];
      src = r0 + msrc[t - w];
    }
    const float* sp = A + (int64_t)(src < 0 ? r0 : src) * lda + c0;
#pragma unroll
    for (int c = 0; c < kTsW; c++) {
      const float y = sp[min(c, w - 1)];
      x[c] = (src >= 0 && c < w) ? y : 0.0f;
    }
    __syncthreads();  // every row this workgroup reads is read before any is written
  } else if constexpr (VEC) {
    // 256 rows x w columns: eight rows per wavefront instruction, 16 bytes a lane
    float4 v4[kApplyT / 32];  // all eight loads in flight, then the stores
#pragma unroll
    for (int g = 0; g < kApplyT / 32; g++) {
      const int gi = min(i0 + g * 32 + (t >> 3), n - 1);
      v4[g] = *((const float4*)(A + (int64_t)gi * lda + c0) + (segok ? seg : 0));
    }
#pragma unroll
    for (int g = 0; g < kApplyT / 32; g++) {
      const int row = g * 32 + (t >> 3);
      *(float4*)&stg[row][seg * 4] = (i0 + row >= n || !segok) ? make_float4(0.f, 0.f, 0.f, 0.f) : v4[g];
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < kTsW; c += 4) {
      const float4 v4 = *(const float4*)&stg[t][c];
      x[c] = v4.x;
      x[c + 1] = v4.y;
      x[c + 2] = v4.z;
      x[c + 3] = v4.w;
    }
    dst = i0 + t;
  } else {
    dst = i0 + t;
    const float* sp = A + (int64_t)min(dst, n - 1) * lda + c0;
#pragma unroll
    for (int c = 0; c < kTsW; c++) {
      const float y = sp[min(c, w - 1)];
      x[c] = c < w ? y : 0.0f;
    }
  }
  subst_columns(std::make_integer_sequence<int, kTsW>(), x, U, dinv);
  if (special || !VEC) {
    const bool moved = !special && ((movedbits[t >> 5] >> (t & 31)) & 1u);  // a moved row is workgroup 0's
    if (dst < 0 || dst >= n || moved) return;
    float* d = A + (int64_t)dst * lda + c0;
    const int top = dst - r0;  // < w only for the pivot positions
#pragma unroll
    for (int c = 0; c < kTsW; c++)
      if (c < w) d[c] = (top < w && c >= top) ? U[top][c] : x[c];
    return;
  }
  if constexpr (VEC) {
#pragma unroll
    for (int c = 0; c < kTsW; c += 4) *(float4*)&stg[t][c] = make_float4(x[c], x[c + 1], x[c + 2], x[c + 3]);
    __syncthreads();
#pragma unroll
    for (int g = 0; g < kApplyT / 32; g++) {
      const int row = g * 32 + (t >> 3);
      const bool skip = (movedbits[row >> 5] >> (row & 31)) & 1u;  // a moved row: workgroup 0 writes it
      if (i0 + row < n && segok && !skip)
        *((float4*)(A + (int64_t)(i0 + row) * lda + c0) + seg) = *(const float4*)&stg[row][seg * 4];
    }
  }
}

}  // namespace

// Scratch of the tournament: merge counters (zero between leaves), the
// final merge's output, the candidate sets of every level.
constexpr int kTsluMaxCounters = 4096;
size_t tslu_scratch_bytes(int n) {
  const int G = ceil_div(n, kTsRows);
  size_t sets = 0;
  for (int s = G; s >= 1; s = s == 1 ? 0 : next_count(s)) sets += (size_t)s;
  return 4 * (size_t)kTsluMaxCounters + ((sizeof(TsluOut) + 255) & ~(size_t)255) + 4 * sets * kTsW + 256;
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
  // 16-byte row pieces: every row start and the leaf's first column aligned
  const bool vec = lda % 4 == 0 && c0 % 4 == 0 && w % 4 == 0 && ((uintptr_t)A & 15) == 0;
  {
    ProfScope ps(ctx, K_LEAF, (double)m * w * w, 4.0 * (double)m * w);
    if (vec)
      tslu_tournament_kernel<true><<<G, 64, 0, ctx->stream>>>(A, lda, n, r0, c0, w, counters, cand, out, ipiv, info);
    else
      tslu_tournament_kernel<false><<<G, 64, 0, ctx->stream>>>(A, lda, n, r0, c0, w, counters, cand, out, ipiv, info);
    NMX_HIP(ctx, hipGetLastError());
  }
  {
    ProfScope ps(ctx, K_MP_PANEL, (double)m * w * w, 8.0 * (double)m * w);
    const int blocks = 1 + ceil_div(m - w, kApplyT);
    if (vec)
      tslu_apply_kernel<true><<<blocks, kApplyT, 0, ctx->stream>>>(A, lda, n, r0, c0, w, out, counters, ncounters);
    else
      tslu_apply_kernel<false><<<blocks, kApplyT, 0, ctx->stream>>>(A, lda, n, r0, c0, w, out, counters, ncounters);
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
