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
//      rows' original values; the same launch moves the rows in the panel's
//      other columns and forms U12 = L11^-1 A12 right of the leaf, so one
//      GEMM (A22 -= L21 U12) finishes the leaf.  The panel is a flat run of
//      such leaves, and the columns outside it see the interchanges once,
//      after the panel.
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
  int selslot;                  // lane j: its slot (register row r, lane l: r * 64 + l)
  bool done;                    // uniform: w columns taken, or no candidate left in this set
};

template <int J>
__device__ __forceinline__ uint32_t column_kmax(const float (&v)[kTsR][kTsW], const GeppState& g) {
  uint32_t k = 0;
#pragma unroll
  for (int r = 0; r < kTsR; r++) k = max(k, (__float_as_uint(v[r][J]) & g.km[r]) | g.kc[r]);
  return wave_umax(k);
}

// One step of partial pivoting over the wavefront's rows for compile-time
// column J (so every register index is static): lane `lane`, register row r
// is slot r * 64 + lane, original row rid[r] (-1: none).  FINAL (a uniform
// flag) also stores U's row J in rows[J] and simulates the LAPACK
// interchanges.
template <int J>
__device__ __forceinline__ void gepp_step(float (&v)[kTsR][kTsW], const int (&rid)[kTsR], int w, int lane,
                                          float (*rows)[kTsW], int r0, bool final, GeppState& g, FinalState& fs) {
  const uint32_t kmax = column_kmax<J>(v, g);
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
  if (lane == J) {
    g.sel = prow_id;
    g.selslot = slot;
  }
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
                                            bool final, FinalState& fs, int& selslot) {
  GeppState g;
#pragma unroll
  for (int r = 0; r < kTsR; r++) {
    const bool live = rid[r] >= 0;
    g.km[r] = live ? kKeyMask : 0u;
    g.kc[r] = live ? (uint32_t)(256 - (r * 64 + lane)) : 0u;
  }
  g.sel = -1;
  g.selslot = -1;
  g.done = false;
  (gepp_step<Js>(v, rid, w, lane, rows, r0, final, g, fs), ...);
  selslot = g.selslot;
  return g.sel;
}

// Partial pivoting of the wavefront's rows: sel (lane j < w) receives the
// row chosen at step j (-1: none left).  The multipliers are not kept: the
// final merge's L11 is recomputed from the selected rows' original values,
// like L21.  The pivot reciprocal is v_rcp (the multipliers here only steer
// the choice of rows).
__device__ __forceinline__ int wave_gepp(float (&v)[kTsR][kTsW], const int (&rid)[kTsR], int w, int lane,
                                         float (*rows)[kTsW], int r0, bool final, FinalState& fs, int& selslot) {
  return gepp_columns(std::make_integer_sequence<int, kTsW>(), v, rid, w, lane, rows, r0, final, fs, selslot);
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
  float asel[kTsW][kTsW];  // the pivot rows' original leaf values, in pivot order (L11 = asel U11^-1)
  int32_t sel[kTsW];       // original row of pivot j
  int32_t move_pos[kTsW];  // rows below the top that receive a displaced top row
  int32_t move_src[kTsW];  // ... which one (0 .. w-1)
  int32_t nmoves;
};

// sets (or merge groups) of the next level up
__host__ __device__ __forceinline__ int next_count(int s) { return (s + kTsFan - 1) / kTsFan; }

// The tournament for one leaf (VEC: the final merge takes asel, the pivot
// rows' original values, from its own LDS staging area).
template <bool VEC>
__global__ void __launch_bounds__(64) tslu_tournament_kernel(const float* __restrict__ A, int lda, int n, int r0,
                                                            int c0, int w, int32_t* __restrict__ counters,
                                                            int32_t* __restrict__ cand, TsluOut* __restrict__ out, int32_t* __restrict__ ipiv,
                                                            int32_t* __restrict__ info) {
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
    int selslot;
    const int sel = wave_gepp(v, rid, w, lane, rows, r0, final, fs, selslot);
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
        if constexpr (VEC) {
          const float* sv = stg[selslot < 0 ? 0 : selslot];
#pragma unroll
          for (int c = 0; c < kTsW; c += 4) {
            const float4 x = *(const float4*)&sv[c];
            *(float4*)&out->asel[lane][c] = selslot < 0 ? make_float4(0.f, 0.f, 0.f, 0.f) : x;
          }
        } else {
          const float* sp = A + (int64_t)(sel < 0 ? r0 : sel) * lda + c0;
          float a[kTsW];
#pragma unroll
          for (int c = 0; c < kTsW; c++) a[c] = sp[min(c, w - 1)];
#pragma unroll
          for (int c = 0; c < kTsW; c++) out->asel[lane][c] = c < w ? a[c] : 0.0f;
        }
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

// Three kinds of workgroups, all independent of each other (each writes
// what no other reads):
//  * workgroup 0 writes every row whose contents move, in the leaf's
//    columns: the w pivot positions (L11 from the selected rows' original
//    values, U11 from the merge) and the positions below that receive a
//    displaced top row (its L21).  It reads only rows it writes itself,
//    before writing any.  It also clears the merge counters for the next leaf;
//  * workgroups 1 .. l21_blocks take 256 rows each of L21 = A21 U11^-1 in
//    place (rows that receive a displaced top row left to workgroup 0),
//    staged through LDS so loads and stores are coalesced (VEC);
//  * the rest take 32 of the panel's other columns (pc0 .. pc1 outside the
//    leaf) each: they apply the leaf's row moves there and, right of the
//    leaf, turn the new top rows into U12 = L11^-1 A12 (forward substitution
//    with L11 = asel U11^-1).  After this kernel the panel's trailing
//    columns need only A22 -= L21 U12.
template <bool VEC>
__global__ void __launch_bounds__(kApplyT) tslu_apply_kernel(float* __restrict__ A, int lda, int n, int r0, int c0,
                                                            int w, const TsluOut* __restrict__ out,
                                                            int32_t* __restrict__ counters, int ncounters,
                                                            int l21_blocks, int pc0, int pc1) {
  __shared__ __attribute__((aligned(16))) float U[kTsW][kTsW];
  __shared__ float dinv[kTsW];
  __shared__ int mpos[kTsW], msrc[kTsW], psel[kTsW];
  __shared__ int nm;
  __shared__ uint32_t movedbits[kApplyT / 32];  // this workgroup's rows that workgroup 0 writes
  __shared__ __attribute__((aligned(16))) float stg[VEC ? kApplyT : 1][kStgLd];
  __shared__ float L11[kTsW][kTsW + 1];         // strips: L11, then U12 in place of A12
  __shared__ float a12[kTsW][kTsW + 1];
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
  const bool strip = (int)blockIdx.x > l21_blocks;
  const int i0 = r0 + w + (blockIdx.x - 1) * kApplyT;
  __syncthreads();
  if (!special && !strip && t < nm) {
    const int rel = mpos[t] - i0;
    if (rel >= 0 && rel < kApplyT) atomicOr(&movedbits[rel >> 5], 1u << (rel & 31));
  }
  __syncthreads();
  // this thread's row: its destination and where its values come from
  float x[kTsW];
  int dst = -1;
  const int seg = t & 7;
  const bool segok = seg * 4 < w;
  if (strip) {
    // thread t < w: row t of L11 from the selected row's original values
#pragma unroll
    for (int c = 0; c < kTsW; c++) x[c] = t < w ? out->asel[t][c] : 0.0f;
  } else if (special) {
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
  if (strip) {
    // columns cs .. ce of the panel, left of the leaf or right of it
    const int s = (int)blockIdx.x - 1 - l21_blocks;
    const int nleft = (c0 - pc0 + kTsW - 1) / kTsW;
    const bool right = s >= nleft;
    const int cs = right ? c0 + w + (s - nleft) * kTsW : pc0 + s * kTsW;
    const int ncols = min(kTsW, (right ? pc1 : c0) - cs);
    const int moves = w + nm;
    if (t < w) {
#pragma unroll
      for (int j = 0; j < kTsW; j++) L11[t][j] = j < t ? x[j] : 0.0f;
    }
    // every moved row's values in these columns (move q < w: pivot position
    // r0 + q gets row psel[q]; q >= w: mpos[q - w] gets top row r0 + msrc[..])
    constexpr int kPer = 2 * kTsW * kTsW / kApplyT;
    float val[kPer];
#pragma unroll
    for (int e = 0; e < kPer; e++) {
      const int idx = t + kApplyT * e, q = idx / kTsW, c = idx % kTsW;
      const int src = q < w ? psel[q] : r0 + msrc[min(q - w, kTsW - 1)];
      val[e] = (q < moves && c < ncols) ? A[(int64_t)src * lda + cs + c] : 0.0f;
    }
    if (right) {
#pragma unroll
      for (int e = 0; e < kPer; e++) {
        const int idx = t + kApplyT * e, q = idx / kTsW, c = idx % kTsW;
        if (q < w) a12[q][c] = val[e];
      }
    }
    __syncthreads();  // every row this workgroup moves is read (and L11 is set) before any is written
    if (right) {
      if (t < ncols) {  // U12 = L11^-1 A12, column t, in pivot order, in registers
        float y[kTsW];
#pragma unroll
        for (int r = 0; r < kTsW; r++) y[r] = a12[r][t];
#pragma unroll
        for (int r = 1; r < kTsW; r++) {
          if (r < w) {  // uniform
            float acc = y[r];
#pragma unroll
            for (int j = 0; j < r; j++) acc = fmaf(-L11[r][j], y[j], acc);
            y[r] = acc;
          }
        }
#pragma unroll
        for (int r = 1; r < kTsW; r++) a12[r][t] = y[r];
      }
      __syncthreads();
    }
#pragma unroll
    for (int e = 0; e < kPer; e++) {
      const int idx = t + kApplyT * e, q = idx / kTsW, c = idx % kTsW;
      if (q < moves && c < ncols) {
        const int d = q < w ? r0 + q : mpos[min(q - w, kTsW - 1)];
        A[(int64_t)d * lda + cs + c] = (right && q < w) ? a12[q][c] : val[e];
      }
    }
    return;
  }
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

// The panel's interchanges over the columns c_lo .. c_hi (outside the
// panel), once per panel: leaf by leaf in order, each thread moves one column -- all of a leaf's
// moved rows loaded (w pivot positions from the selected rows, nmoves
// displaced top rows into the vacated positions), then stored, so the
// result is the leaf's interchange sequence applied in order.  The moves are
// the tournament's own records, read once per workgroup into LDS.
constexpr int kTsluMaxLeaves = 32;  // leaves per panel (panel width <= 1024)
__global__ void __launch_bounds__(64) tslu_swap_outside_kernel(float* __restrict__ A, int lda, int k, int nb,
                                                               int c_lo, int c_hi, const TsluOut* __restrict__ outs,
                                                               int leaves) {
  __shared__ int msrc_s[kTsluMaxLeaves][2 * kTsW], mdst_s[kTsluMaxLeaves][2 * kTsW], mcnt_s[kTsluMaxLeaves];
  const int t = threadIdx.x;
  for (int e = t; e < leaves * 2 * kTsW; e += blockDim.x) {
    const int li = e / (2 * kTsW), q = e % (2 * kTsW);
    const TsluOut& o = outs[li];
    const int r0 = k + li * kTsW, w = min(kTsW, k + nb - r0), nm = o.nmoves;
    int src = -1, dst = -1;
    if (q < w) {
      src = o.sel[q];
      dst = r0 + q;
    } else if (q < w + nm) {
      src = r0 + o.move_src[q - w];
      dst = o.move_pos[q - w];
    }
    msrc_s[li][q] = src < 0 ? r0 : src;  // unused entries read a valid row and store nothing
    mdst_s[li][q] = dst;
    if (q == 0) mcnt_s[li] = w + nm;
  }
  __syncthreads();
  const int c = c_lo + (int)(blockIdx.x * blockDim.x) + t;
  if (c >= c_hi) return;
  for (int li = 0; li < leaves; li++) {
    float v[2 * kTsW];
#pragma unroll
    for (int q = 0; q < 2 * kTsW; q++) v[q] = A[(int64_t)msrc_s[li][q] * lda + c];
    const int cnt = mcnt_s[li];
#pragma unroll
    for (int q = 0; q < 2 * kTsW; q++)
      if (q < cnt) A[(int64_t)mdst_s[li][q] * lda + c] = v[q];
  }
}

}  // namespace

// Scratch of the tournament: merge counters (zero between leaves), two
// slots of final-merge output (a panel's interchange records stay readable
// while the next panel fills the other slot), the candidate sets of every
// level.
constexpr int kTsluMaxCounters = 4096;
constexpr size_t kOutsBytes = ((sizeof(TsluOut) * kTsluMaxLeaves + 255) & ~(size_t)255);
size_t tslu_scratch_bytes(int n) {
  const int G = ceil_div(n, kTsRows);
  size_t sets = 0;
  for (int s = G; s >= 1; s = s == 1 ? 0 : next_count(s)) sets += (size_t)s;
  return 4 * (size_t)kTsluMaxCounters + 2 * kOutsBytes + 4 * sets * kTsW + 256;
}

// One panel (columns k .. k+nb, rows k .. n) as a flat sequence of 32-column
// leaves: tournament, apply (which also moves the panel's other columns and
// forms U12), and one GEMM for the panel's columns right of the leaf; then
// (swap_outside) the panel's interchanges over every column outside it.
static TsluOut* outs_slot(void* scratch, int slot) {
  return (TsluOut*)((char*)scratch + 4 * (size_t)kTsluMaxCounters + (slot & 1) * kOutsBytes);
}

int tslu_swap_columns(nmx_ctx* ctx, void* scratch, float* A, int lda, int k, int nb, int c_lo, int c_hi, int slot) {
  if (c_hi <= c_lo || nb <= 0) return NMX_OK;
  if (c_lo < k + nb && c_hi > k) return NMX_ERR_UNSUPPORTED;  // the panel's own columns are moved by the apply
  const TsluOut* outs = outs_slot(scratch, slot);
  ProfScope ps(ctx, K_LASWP, 0.0, 16.0 * nb * (double)(c_hi - c_lo));
  tslu_swap_outside_kernel<<<ceil_div(c_hi - c_lo, 64), 64, 0, ctx->stream>>>(A, lda, k, nb, c_lo, c_hi, outs,
                                                                                ceil_div(nb, kTsW));
  NMX_HIP(ctx, hipGetLastError());
  return NMX_OK;
}

int tslu_panel_launch(nmx_ctx* ctx, void* scratch, size_t scratch_bytes, int n, float* A, int lda, int k, int nb,
                      int32_t* ipiv, int32_t* info, int swap_outside, int slot) {
  if (nb <= 0 || n - k <= 0) return NMX_OK;
  if (tslu_scratch_bytes(n) > scratch_bytes) return NMX_ERR_UNSUPPORTED;
  if (ceil_div(nb, kTsW) > kTsluMaxLeaves) return NMX_ERR_UNSUPPORTED;
  int32_t* counters = (int32_t*)scratch;
  TsluOut* outs = outs_slot(scratch, slot);
  int32_t* cand = (int32_t*)((char*)scratch + 4 * (size_t)kTsluMaxCounters + 2 * kOutsBytes);
  const int pc0 = k, pc1 = k + nb;
  for (int c0 = k; c0 < pc1; c0 += kTsW) {
    TsluOut* out = outs + (c0 - k) / kTsW;
    const int r0 = c0, w = std::min(kTsW, pc1 - c0), m = n - r0;
    if (m < w) return NMX_ERR_UNSUPPORTED;
    const int G = ceil_div(m, kTsRows);
    int ncounters = 0;
    for (int s = G; s > 1; s = next_count(s)) ncounters += next_count(s);
    if (ncounters > kTsluMaxCounters) return NMX_ERR_UNSUPPORTED;
    // 16-byte row pieces: every row start and the leaf's first column aligned
    const bool vec = lda % 4 == 0 && c0 % 4 == 0 && w % 4 == 0 && ((uintptr_t)A & 15) == 0;
    {
      ProfScope ps(ctx, K_LEAF, (double)m * w * w, 4.0 * (double)m * w);
      if (vec)
        tslu_tournament_kernel<true><<<G, 64, 0, ctx->stream>>>(A, lda, n, r0, c0, w, counters, cand, out, ipiv,
                                                                info);
      else
        tslu_tournament_kernel<false><<<G, 64, 0, ctx->stream>>>(A, lda, n, r0, c0, w, counters, cand, out, ipiv,
                                                                 info);
      NMX_HIP(ctx, hipGetLastError());
    }
    {
      const int other = (c0 - pc0) + (pc1 - c0 - w);
      ProfScope ps(ctx, K_MP_PANEL, (double)m * w * w, 8.0 * (double)m * w + 16.0 * w * (double)other);
      const int l21_blocks = ceil_div(m - w, kApplyT);
      const int strips = ceil_div(c0 - pc0, kTsW) + ceil_div(pc1 - c0 - w, kTsW);
      const int blocks = 1 + l21_blocks + strips;
      if (vec)
        tslu_apply_kernel<true><<<blocks, kApplyT, 0, ctx->stream>>>(A, lda, n, r0, c0, w, out, counters, ncounters,
                                                                     l21_blocks, pc0, pc1);
      else
        tslu_apply_kernel<false><<<blocks, kApplyT, 0, ctx->stream>>>(A, lda, n, r0, c0, w, out, counters, ncounters,
                                                                      l21_blocks, pc0, pc1);
      NMX_HIP(ctx, hipGetLastError());
    }
    const int rows = m - w, cols = pc1 - c0 - w;
    if (rows > 0 && cols > 0)
      NMX_TRY(gemm_launch<float>(ctx, 0, 0, rows, cols, w, -1.0f, A + (int64_t)(r0 + w) * lda + c0, lda,
                                 A + (int64_t)r0 * lda + c0 + w, lda, 1.0f, A + (int64_t)(r0 + w) * lda + c0 + w, lda,
                                 0));
  }
  if (swap_outside) {
    NMX_TRY(tslu_swap_columns(ctx, scratch, A, lda, k, nb, 0, k, slot));
    NMX_TRY(tslu_swap_columns(ctx, scratch, A, lda, k, nb, k + nb, n, slot));
  }
  return NMX_OK;
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
