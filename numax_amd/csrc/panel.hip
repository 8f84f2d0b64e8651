// Panel kernels of the blocked factorizations: the cooperative pivoting leaf,
// row interchanges, triangular solves, the Cholesky diagonal block, and the
// small gather / tril helpers.
//
// numax runs its panels as single-workgroup kernels over the whole remaining
// height (numax/linalg/panel.mojo:289-431) while the rest of the device idles;
// its own docstring names that as the ceiling (panel.mojo:41-49).  Here the
// leaf of the recursive panel is spread over up to 32 workgroups of 1024
// threads, one matrix row per thread held in registers, and the per-column
// pivot search is one exchange through a mailbox in device memory.
#include <limits.h>

#include "nmx_internal.h"

namespace {

constexpr double kPivotFloor = 1e-30;  // numax/linalg/common.mojo:27-29
constexpr long kSpinLimit = 1L << 23;  // bounded waits: a few seconds at worst

template <typename T>
__device__ __forceinline__ bool better(T ma, int ra, T mb, int rb) {
  return ma > mb || (ma == mb && ra < rb);
}

// ---- mailbox ----------------------------------------------------------------
// Every word a participant publishes is a 64-bit atomic holding 32 bits of
// payload and the 32-bit epoch of the column it belongs to, so a reader
// validates each word on its own and needs no fence and no second trip.  A
// double travels as two words.  Slots are double-buffered by epoch parity: a
// participant can only publish column j+2 after every participant has
// published j+1, which each does only after reading all of column j.
template <typename T>
struct MbLayout {
  static constexpr int VW = sizeof(T) / 4;       // words per value
  static constexpr int STRIDE = 1 + VW + kLeafW * VW;  // header, |pivot|, row values
  static constexpr int PARITY = kMaxParts * STRIDE + kLeafW * VW;  // + the diagonal row
};
static_assert(2 * MbLayout<double>::PARITY * 8 <= kMailboxBytes, "mailbox too small");

__device__ __forceinline__ uint64_t mb_pack(uint32_t payload, uint32_t epoch) {
  return ((uint64_t)payload << 32) | epoch;
}
__device__ __forceinline__ void mb_put(uint64_t* w, uint32_t payload, uint32_t epoch) {
  __hip_atomic_store(w, mb_pack(payload, epoch), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void mb_put_value(uint64_t* w, T v, uint32_t epoch) {
  if (sizeof(T) == 8) {
    const uint64_t bits = __double_as_longlong((double)v);
    mb_put(w, (uint32_t)bits, epoch);           // low half first
    mb_put(w + 1, (uint32_t)(bits >> 32), epoch);
  } else {
    mb_put(w, __float_as_uint((float)v), epoch);
  }
}

template <typename T, int N>
__device__ __forceinline__ void rotate_left(T (&v)[N]) {
  const T first = v[0];
#pragma unroll
  for (int c = 0; c + 1 < N; c++) v[c] = v[c + 1];
  v[N - 1] = first;
}

// Pivot keys: a candidate's magnitude as its (order-preserving) bit pattern
// plus one, 0 for "no candidate".  Equal keys are equal magnitudes, so the
// first maximum is the smallest row among the largest key.
template <typename T>
struct PivKey;
template <>
struct PivKey<float> {
  typedef uint32_t K;
  static __device__ __forceinline__ K of(float mag) { return __float_as_uint(mag) + 1u; }
  static __device__ __forceinline__ K wave_max(K k) { return wave_umax(k); }
  static constexpr int WORDS = 1;
  static __device__ __forceinline__ uint32_t word(K k, int i) { return k; }
  static __device__ __forceinline__ K from(const uint32_t* w) { return w[0]; }
};
template <>
struct PivKey<double> {
  typedef uint64_t K;
  static __device__ __forceinline__ K of(double mag) { return (uint64_t)__double_as_longlong(mag) + 1ull; }
  static __device__ __forceinline__ K wave_max(K k) {
    const uint32_t hi = wave_umax((uint32_t)(k >> 32));
    const uint32_t lo = wave_umax((uint32_t)(k >> 32) == hi ? (uint32_t)k : 0u);
    return ((uint64_t)hi << 32) | lo;
  }
  static constexpr int WORDS = 2;
  static __device__ __forceinline__ uint32_t word(K k, int i) { return i ? (uint32_t)(k >> 32) : (uint32_t)k; }
  static __device__ __forceinline__ K from(const uint32_t* w) { return ((uint64_t)w[1] << 32) | w[0]; }
};

// ---------------------------------------------------------------------------
// Cooperative leaf of the recursive LU panel.
//
// Factors columns c0..c0+w (w <= 32) over rows r0..n with partial pivoting,
// exactly getrf_panel's arithmetic (panel.mojo:289-431): first-maximum pivot
// (ties to the lower row; the diagonal row counts as magnitude >= 0, so an
// all-zero or all-NaN column keeps its diagonal; NaN never wins),
// interchange, an exact-zero pivot divided as 1e-30 with `info` recording the
// first, true division for the multipliers and the rank-1 update of the
// panel's own columns.  The leaf's interchanges are then replayed over every
// other column of the matrix right after (laswp_gather_kernel), which leaves
// the matrix numax's full-width swaps leave (panel.mojo:404-411).
//
// Workgroup p of `parts` owns rows r0 + 512 ROWS p + t + 512 i (i < ROWS) of
// its 512 threads, each row's 32 columns in registers.  The column loop stays
// rolled (unrolled it overflows the instruction cache); the rows are rotated
// left by one after each column so register indices stay static: at step j,
// v[.][0] is column j and v[.][1 .. w-j) are the columns still to update.
// Per column, in one workgroup: DPP reductions of the pivot key and row, each
// wavefront's winning row staged in LDS, one barrier.  Across workgroups
// (parts > 1, grid 8 * parts of which every eighth takes part, so all
// participants share one XCD and its L2): each part publishes its winner,
// its key and (part 0) the diagonal row as epoch-tagged 64-bit words, and
// every thread of every part spins on the words it stages, no fences: all of
// them in one round trip up to 16 parts; above that the keys and rows first,
// then the winner's row and the diagonal row.  Slots alternate by epoch
// parity: a part publishes column j+2 only after reading every part's key
// for column j+1, which each part publishes only after its reads of column j.
// ---------------------------------------------------------------------------
constexpr int kLeafT = 512;  // threads per workgroup
constexpr int kTwoPhaseParts = 16;  // above this many parts the exchange reads the winner's row only
constexpr int kLeafWaves = kLeafT / 64;

// ROWS rows per thread: 2 for a panel one workgroup holds (up to 1024
// rows), 1 when the panel is spread over several (half the per-column
// vector work per CU, twice the workgroups; up to 32 x 512 rows).
template <typename T, int ROWS>
__global__ void __launch_bounds__(kLeafT) getf2_leaf(T* __restrict__ A, int lda, int n, int r0, int c0, int w,
                                                     int32_t* __restrict__ ipiv, int32_t* __restrict__ info,
                                                     uint64_t* __restrict__ mb, uint32_t epoch0, int parts, int* err) {
  constexpr int W = kLeafW;
  constexpr int VW = MbLayout<T>::VW;
  constexpr int STRIDE = MbLayout<T>::STRIDE;
  typedef PivKey<T> PK;
  typedef typename PK::K Key;
  const bool multi = parts > 1;
  if (multi && (blockIdx.x % kXcds) != 0) return;
  const int p = multi ? (int)(blockIdx.x / kXcds) : 0;
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;

  // per-column staging, double-buffered by column parity
  __shared__ __attribute__((aligned(16))) T slot_v[2][kLeafWaves][W];  // each wavefront's winning row
  __shared__ Key slot_k[2][kLeafWaves];
  __shared__ int slot_r[2][kLeafWaves];
  __shared__ __attribute__((aligned(16))) T diag_v[2][W];               // the diagonal row
  // staged mailbox words, per part [values | key | row] (PSTRIDE even, so
  // every part's values stay 8-byte aligned), then the diagonal row
  constexpr int PSTRIDE = (W * VW + VW + 1 + 1) & ~1;
  __shared__ __attribute__((aligned(16))) uint32_t stage[2][kMaxParts * PSTRIDE + W * VW];
  __shared__ int aborted, info_set;

  if (t == 0) {
    aborted = 0;
    info_set = (*info != 0);
  }

  const int row0 = r0 + p * (kLeafT * ROWS) + t;  // row i = row0 + 512 i
  T v[ROWS][W];
  {
    // one descriptor at this workgroup's first row: offsets stay below 1024 rows
    const __amdgpu_buffer_rsrc_t rs = buf_rsrc(A + (int64_t)(r0 + p * (kLeafT * ROWS)) * lda + c0);
#pragma unroll
    for (int i = 0; i < ROWS; i++) {
      const bool own = row0 + kLeafT * i < n;
      const uint32_t rel = (uint32_t)(t + kLeafT * i) * (uint32_t)lda;
#pragma unroll
      for (int c = 0; c < W; c++)
        v[i][c] = buf_load<T>(rs, (own & (c < w)) ? (rel + c) * (uint32_t)sizeof(T) : kBufOob);
    }
  }
  __syncthreads();

  for (int j = 0; j < w; j++) {
    const int gr = r0 + j;
    const int gc = c0 + j;
    const int par = j & 1;
    const uint32_t epoch = epoch0 + (uint32_t)j + 1u;

    // 1. this thread's candidate: largest key, smallest row
    Key tk = 0;
    int trow = INT_MAX;
#pragma unroll
    for (int i = 0; i < ROWS; i++) {
      const int row = row0 + kLeafT * i;
      if (row >= gr && row < n) {
        T mag = fabs(v[i][0]);
        if (!(mag == mag)) mag = T(-1);           // NaN never wins ...
        if (row == gr && mag < T(0)) mag = T(0);  // ... the diagonal row starts at 0
        if (mag >= T(0)) {
          const Key k = PK::of(mag);
          if (k > tk) { tk = k; trow = row; }   // rows ascend with i: ties keep the first
        }
      }
    }
    // 2. the wavefront's: largest key, then the smallest row holding it
    const Key wk = PK::wave_max(tk);
    const int wrow = (int)wave_umin(tk == wk ? (uint32_t)trow : 0xFFFFFFFFu);
    // its row values (and the diagonal row's) into LDS
#pragma unroll
    for (int i = 0; i < ROWS; i++) {
      const int row = row0 + kLeafT * i;
      if (wk != 0 && row == wrow) {
#pragma unroll
        for (int c = 0; c < W; c++) slot_v[par][wave][c] = v[i][c];
      }
      if (row == gr) {
#pragma unroll
        for (int c = 0; c < W; c++) diag_v[par][c] = v[i][c];
      }
    }
    if (lane == 0) {
      slot_k[par][wave] = wk;
      slot_r[par][wave] = wrow;
    }
    __syncthreads();
    // 3. the workgroup's winner (every wavefront works it out): lane q reads
    // wavefront q's candidate, then the same key / smallest-row reductions
    Key ck = 0;
    int cr = INT_MAX;
    if (lane < kLeafWaves) {
      ck = slot_k[par][lane];
      cr = slot_r[par][lane];
    }
    const Key gk0 = PK::wave_max(ck);
    Key gk = gk0;
    int grow = INT_MAX, gw = 0;
    if (gk0 != 0) {  // uniform; rows are distinct, so one lane holds the winner
      grow = (int)wave_umin(ck == gk0 ? (uint32_t)cr : 0xFFFFFFFFu);
      gw = __ffsll((long long)__ballot(ck == gk0 && cr == grow)) - 1;
    }
    const T* urow;
    const T* drow;
    if (!multi) {
      urow = slot_v[par][gw];
      drow = diag_v[par];
    } else {
      // 4. publish this part's winner (or "none") and, from part 0, the diagonal row
      uint64_t* mbp = mb + (size_t)(epoch & 1u) * MbLayout<T>::PARITY;
      uint64_t* myslot = mbp + (size_t)p * STRIDE;
      if (t < W * VW) {
        const uint32_t* src = (const uint32_t*)slot_v[par][gw];
        mb_put(myslot + 1 + VW + t, gk != 0 ? src[t] : 0u, epoch);
      } else if (t < W * VW + VW) {
        mb_put(myslot + 1 + (t - W * VW), PK::word(gk, t - W * VW), epoch);
      } else if (t == W * VW + VW) {
        mb_put(myslot, (uint32_t)grow, epoch);
      } else if (p == 0 && t >= 128 && t < 128 + W * VW) {
        mb_put(mbp + (size_t)kMaxParts * STRIDE + (t - 128), ((const uint32_t*)diag_v[par])[t - 128], epoch);
      }
      // 5. stage the parts' words, each validated by its epoch.  With many
      // parts every workgroup reading every part's whole row crowds the few
      // L2 lines they sit in, so it reads the keys and rows first, then only
      // the winner's values and the diagonal row (two round trips, a tenth of
      // the data; 10% faster over a 32-part leaf, 2% slower over 8 parts).
      uint32_t* stg = stage[par];
      const bool two_phase = parts > kTwoPhaseParts;  // uniform
      auto stage_word = [&](const uint64_t* src, int dst) {
        uint64_t wv;
        long spins = 0;
        for (;;) {
          wv = __hip_atomic_load(src, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if ((uint32_t)wv == epoch) break;
          if (aborted || ++spins > kSpinLimit) { aborted = 1; break; }
          __builtin_amdgcn_s_sleep(1);
        }
        stg[dst] = (uint32_t)(wv >> 32);
      };
      if (two_phase) {
        constexpr int HW = VW + 1;  // header words of a slot: row, key
        for (int idx = t; idx < parts * HW; idx += kLeafT) {
          const int q = idx / HW, f = idx % HW;
          stage_word(mbp + (size_t)q * STRIDE + f, q * PSTRIDE + (f == 0 ? W * VW + VW : W * VW + f - 1));
        }
      } else {
        const int words = parts * STRIDE + W * VW;
        for (int idx = t; idx < words; idx += kLeafT) {
          const uint64_t* src =
              idx < parts * STRIDE ? mbp + idx : mbp + (size_t)kMaxParts * STRIDE + (idx - parts * STRIDE);
          int dst;
          if (idx < parts * STRIDE) {
            const int q = idx / STRIDE, f = idx % STRIDE;  // mailbox slot: [row | key | values]
            dst = q * PSTRIDE + (f == 0 ? W * VW + VW : (f <= VW ? W * VW + f - 1 : f - 1 - VW));
          } else {
            dst = kMaxParts * PSTRIDE + (idx - parts * STRIDE);
          }
          stage_word(src, dst);
        }
      }
      __syncthreads();
      // 6. the global winner: largest key, then smallest row (lane q reads
      // part q's; parts <= 32)
      Key pk = 0;
      int pr = INT_MAX;
      if (lane < parts) {
        const uint32_t* sq = stg + lane * PSTRIDE;
        pk = PK::from(sq + W * VW);
        pr = (int)sq[W * VW + VW];
      }
      const Key bk = PK::wave_max(pk);
      int brow = INT_MAX, bp = 0;
      if (bk != 0) {
        brow = (int)wave_umin(pk == bk ? (uint32_t)pr : 0xFFFFFFFFu);
        bp = __ffsll((long long)__ballot(pk == bk && pr == brow)) - 1;
      }
      if (two_phase) {
        // the winner's row values and the diagonal row (part bp published its
        // values whether or not it had a candidate)
        for (int idx = t; idx < 2 * W * VW; idx += kLeafT)
          stage_word(idx < W * VW ? mbp + (size_t)bp * STRIDE + 1 + VW + idx
                                  : mbp + (size_t)kMaxParts * STRIDE + (idx - W * VW),
                     idx < W * VW ? bp * PSTRIDE + idx : kMaxParts * PSTRIDE + (idx - W * VW));
        __syncthreads();
      }
      gk = bk;
      grow = brow;
      urow = (const T*)(stg + bp * PSTRIDE);
      drow = (const T*)(stg + kMaxParts * PSTRIDE);
    }

    const int wr = (gk == 0 || aborted) ? gr : grow;
    if (wr == gr) urow = drow;
    // FP32: the pivot row into registers once (both rows of the thread use
    // it); FP64 reads it from LDS per row (registers are at their limit)
    constexpr bool kRowRegs = sizeof(T) == 4;
    T u[kRowRegs ? W : 1];
#pragma unroll
    for (int c = 0; c < (kRowRegs ? W : 1); c++) u[c] = urow[c];
    auto urow_at = [&](int c) -> T { return kRowRegs ? u[kRowRegs ? c : 0] : urow[c]; };
    // 7. interchange inside the leaf's columns, record the pivot
#pragma unroll
    for (int i = 0; i < ROWS; i++) {
      const int row = row0 + kLeafT * i;
      if (wr != gr && row < n) {
        if (row == gr) {
#pragma unroll
          for (int c = 0; c < W; c++) v[i][c] = urow_at(c);
        } else if (row == wr) {
#pragma unroll
          for (int c = 0; c < W; c++) v[i][c] = drow[c];
        }
      }
    }
    T pivot = urow_at(0);
    if (t == 0 && p == 0) {
      ipiv[gc] = wr;
      if (fabs(pivot) == T(0) && !info_set) {
        *info = gc + 1;
        info_set = 1;
      }
    }
    if (pivot == T(0)) pivot = T(kPivotFloor);
    // 8. multipliers and the rank-1 update of the leaf's remaining columns.
    // The bound is opaque (a vector register), so the guard is a per-element
    // select rather than 31 uniform branches; columns past it are finished
    // ones rotated to the end and keep their values (a select, not a
    // multiply by 0, which an infinite multiplier would turn into NaN).
    int rem = w - j;
    asm volatile("" : "+v"(rem));
#pragma unroll
    for (int i = 0; i < ROWS; i++) {
      const int row = row0 + kLeafT * i;
      if (row > gr && row < n) {
        const T l = v[i][0] / pivot;
        v[i][0] = l;
#pragma unroll
        for (int c = 1; c < W; c++) {
          const T nv = v[i][c] - l * urow_at(c);
          v[i][c] = c < rem ? nv : v[i][c];
        }
      }
      rotate_left(v[i]);
    }
  }
  for (int j = w; j < W; j++) {
#pragma unroll
    for (int i = 0; i < ROWS; i++) rotate_left(v[i]);
  }

  if (aborted && t == 0) atomicExch(err, NMX_ERR_SYNC_TIMEOUT);
#pragma unroll
  for (int i = 0; i < ROWS; i++) {
    const int row = row0 + kLeafT * i;
    if (row < n) {
      T* dst = A + (int64_t)row * lda + c0;
#pragma unroll
      for (int c = 0; c < W; c++)
        if (c < w) dst[c] = v[i][c];
    }
  }
}

// Replay the interchanges ipiv[k1..k2) (k2 - k1 <= 32) over the columns
// [col_lo, col_hi) minus [skip_lo, skip_lo + skip_w): laswp / laswp_matrix
// (numax/linalg/panel.mojo:1824-1852, :2279-2313).  The net effect of the
// sequence on the <= 2(k2-k1) rows it touches is worked out once per
// workgroup in LDS; each thread then moves one column with all its loads in
// flight together.
template <typename T>
__global__ void __launch_bounds__(256) laswp_gather_kernel(T* __restrict__ A, int lda, int col_lo, int ncols,
                                                           int skip_lo, int skip_w, const int32_t* __restrict__ ipiv,
                                                           int k1, int k2) {
  __shared__ int mv_dst[2 * kLeafW], mv_src[2 * kLeafW], mv_count;
  const int t = threadIdx.x;
  const int w = k2 - k1;
  if (t < 64) {
    // wavefront 0: lane e tracks involved row e and the original row now in it
    int myrow = (t < w) ? k1 + t : -1;
    int mine = myrow;
    int nr = w;
    // the pivots once, one per lane (a load inside the walk waited a scalar
    // round trip per step)
    const int qv = t < w ? ipiv[k1 + t] : 0;
    for (int j = 0; j < w; j++) {
      const int q = __builtin_amdgcn_readlane(qv, j);
      const unsigned long long hit = __ballot(t < nr && myrow == q);
      int pos;
      if (hit) {
        pos = __ffsll((long long)hit) - 1;
      } else {
        pos = nr;
        if (t == nr) { myrow = q; mine = q; }
        nr++;
      }
      const int cj = __shfl(mine, j);
      const int cp = __shfl(mine, pos);
      if (t == j) mine = cp;
      if (t == pos) mine = cj;
    }
    const bool moved = (t < nr) && (mine != myrow);
    const unsigned long long mm = __ballot(moved);
    if (moved) {
      const int at = __popcll(mm & ((1ull << t) - 1ull));
      mv_dst[at] = myrow;
      mv_src[at] = mine;
    }
    if (t == 0) mv_count = __popcll(mm);
  }
  __syncthreads();
  const int moves = mv_count;
  if (moves == 0) return;
  // lane e of each wavefront takes move e out of LDS (one read), and every
  // move's rows then come to the whole wavefront by v_readlane: read per
  // move, the list cost one LDS round trip ahead of each load.  Every lane
  // stays active to the end (v_readlane reads lanes the branch would have
  // turned off); lanes past the last column load from the first one and
  // store nothing.
  static_assert(2 * kLeafW == 64, "one move per lane");
  const int lane = t & 63;
  const int my_src = mv_src[lane < moves ? lane : moves - 1];
  const int my_dst = mv_dst[lane < moves ? lane : moves - 1];
  const int i = blockIdx.x * blockDim.x + t;
  const bool act = i < ncols;
  int c = col_lo + (act ? i : 0);
  if (c >= skip_lo) c += skip_w;
  T vals[2 * kLeafW];
#pragma unroll
  for (int e = 0; e < 2 * kLeafW; e++) {
    // unconditional loads (a clamped, valid row) so all are in flight together
    vals[e] = A[(int64_t)__builtin_amdgcn_readlane(my_src, e) * lda + c];
  }
#pragma unroll
  for (int e = 0; e < 2 * kLeafW; e++)
    if (act && e < moves) A[(int64_t)__builtin_amdgcn_readlane(my_dst, e) * lda + c] = vals[e];
}

// Replay interchanges k1..k2-1 (in order) over the columns [col_lo, col_hi)
// minus [skip_lo, skip_hi): laswp / laswp_matrix (panel.mojo:1824-1852,
// :2279-2313), one thread per column.
template <typename T>
__global__ void laswp_kernel(T* __restrict__ A, int lda, int col_lo, int ncols, int skip_lo, int skip_w,
                             const int32_t* __restrict__ ipiv, int k1, int k2) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncols) return;
  int c = col_lo + i;
  if (c >= skip_lo) c += skip_w;
  for (int j = k1; j < k2; j++) {
    int o = ipiv[j];
    if (o != j) {
      T a = A[(int64_t)j * lda + c];
      T b = A[(int64_t)o * lda + c];
      A[(int64_t)j * lda + c] = b;
      A[(int64_t)o * lda + c] = a;
    }
  }
}

// Net permutation of the interchange sequence ipiv[0..n): row i of the
// permuted matrix is row perm[i] of the original one, which is what laswp
// over all n pivots (panel.mojo:1824-1852) leaves behind.  The sequence is a
// dependent chain, so one thread replays it on an identity vector; in LDS a
// step costs a few dozen cycles instead of the two dependent global
// round trips of a swap in place.  Pivot values are clamped to [0, n), so a
// corrupt ipiv yields a wrong permutation but no out-of-range access.
// The same permutation with the swaps taken 64 at a time.  A chunk's 64
// swaps, applied in order, permute at most 128 positions: its own 64 and up
// to 64 targets below it (ipiv[i] >= i).  A wavefront replays the chunk on
// labels -- slot s holds position pos_s and, after each swap, which
// position's value sits there (cont_s) -- so the chunk's net effect is
// perm'[pos_s] = perm[cont_s] for every touched slot: all reads, then all
// writes.  Sixteen wavefronts work out sixteen chunks at a time into LDS;
// wavefront 0 then applies them in order to the permutation held in LDS.
constexpr int kPermBatch = 16;  // chunks per round (one per wavefront)
__global__ void __launch_bounds__(1024) perm_chunks_kernel(int n, const int32_t* __restrict__ ipiv,
                                                           int32_t* __restrict__ perm) {
  extern __shared__ int s[];  // the permutation, n entries
  __shared__ int mdst[kPermBatch * 128], msrc[kPermBatch * 128], mcnt[kPermBatch];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  for (int i = t; i < n; i += blockDim.x) s[i] = i;
  const int nchunks = (n + 63) / 64;
  for (int c0 = 0; c0 < nchunks; c0 += kPermBatch) {
    const int c = c0 + wv;
    if (wv < kPermBatch && c < nchunks) {
      const int p0 = 64 * c, m = min(64, n - p0);
      const int iv = lane < m ? min(max(ipiv[p0 + lane], 0), n - 1) : p0 + lane;
      int pos0 = p0 + lane, cont0 = pos0;  // slot lane: position p0 + lane
      int pos1 = -1, cont1 = -1;           // slot 64 + lane: a target outside the chunk
      int nhi = 0;
      for (int j = 0; j < m; j++) {
        const int q = __builtin_amdgcn_readlane(iv, j);
        if (q == p0 + j) continue;
        int tq;
        if (q >= p0 && q < p0 + 64) {
          tq = q - p0;
        } else {
          const unsigned long long hit = __ballot(lane < nhi && pos1 == q);
          if (hit) {
            tq = 64 + __ffsll((long long)hit) - 1;
          } else {
            tq = 64 + nhi;
            if (lane == nhi) {
              pos1 = q;
              cont1 = q;
            }
            nhi++;
          }
        }
        const int a = __builtin_amdgcn_readlane(cont0, j);
        const int b = tq < 64 ? __builtin_amdgcn_readlane(cont0, tq) : __builtin_amdgcn_readlane(cont1, tq - 64);
        if (lane == j) cont0 = b;
        if (tq < 64) {
          if (lane == tq) cont0 = a;
        } else if (lane == tq - 64) {
          cont1 = a;
        }
      }
      // compact the moved slots
      const bool mv0 = cont0 != pos0, mv1 = lane < nhi && cont1 != pos1;
      const unsigned long long b0 = __ballot(mv0), b1 = __ballot(mv1);
      const int base1 = __popcll(b0);
      int* d = mdst + wv * 128;
      int* sr = msrc + wv * 128;
      if (mv0) {
        const int at = __popcll(b0 & ((1ull << lane) - 1ull));
        d[at] = pos0;
        sr[at] = cont0;
      }
      if (mv1) {
        const int at = base1 + __popcll(b1 & ((1ull << lane) - 1ull));
        d[at] = pos1;
        sr[at] = cont1;
      }
      if (lane == 0) mcnt[wv] = base1 + __popcll(b1);
    }
    __syncthreads();  // this round's moves, and (first round) the identity, are in LDS
    if (wv == 0) {
      for (int cc = 0; cc < kPermBatch && c0 + cc < nchunks; cc++) {
        const int cnt = mcnt[cc];
        const int* d = mdst + cc * 128;
        const int* sr = msrc + cc * 128;
        const int v0 = lane < cnt ? s[sr[lane]] : 0;
        const int v1 = lane + 64 < cnt ? s[sr[lane + 64]] : 0;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");  // every read before any write
        if (lane < cnt) s[d[lane]] = v0;
        if (lane + 64 < cnt) s[d[lane + 64]] = v1;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
      }
    }
    __syncthreads();  // the moves' LDS is free for the next round
  }
  for (int i = t; i < n; i += blockDim.x) perm[i] = s[i];
}

template <int IN_LDS>
__global__ void __launch_bounds__(1024) perm_kernel(int n, const int32_t* __restrict__ ipiv,
                                                     int32_t* __restrict__ perm) {
  extern __shared__ int s_perm[];
  int* s = IN_LDS ? s_perm : perm;
  const int t = threadIdx.x;
  for (int i = t; i < n; i += blockDim.x) s[i] = i;
  __syncthreads();
  if (t == 0) {
    const auto rs = buf_rsrc(ipiv);
    for (int j0 = 0; j0 < n; j0 += 16) {
      int q[16];
#pragma unroll
      for (int u = 0; u < 16; u++)
        q[u] = (int)__builtin_amdgcn_raw_buffer_load_b32(rs, j0 + u < n ? 4u * (uint32_t)(j0 + u) : kBufOob, 0, 0);
#pragma unroll
      for (int u = 0; u < 16; u++) {
        const int j = j0 + u;
        if (j < n) {
          const int p = min(max(q[u], 0), n - 1);
          const int a = s[j], b = s[p];
          s[j] = b;
          s[p] = a;
        }
      }
    }
  }
  __syncthreads();
  if (IN_LDS)
    for (int i = t; i < n; i += blockDim.x) perm[i] = s[i];
}

// out[i][c] = in[perm[i]][c] for c < ncols, converting Ti to To.
template <typename Ti, typename To>
__global__ void __launch_bounds__(256) gather_rows_kernel(int n, int ncols, const int32_t* __restrict__ perm,
                                                          const Ti* __restrict__ in, int ldi, To* __restrict__ out,
                                                          int ldo) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * ncols) return;
  const int i = (int)(idx / ncols), c = (int)(idx % ncols);
  out[(int64_t)i * ldo + c] = (To)in[(int64_t)perm[i] * ldi + c];
}

// Base triangular solve for n <= NB (32 or 64): one thread per vector, the
// vector in registers, op(T) staged in LDS by columns.  Right-looking with
// every index static: x_j is finished, then x_i -= op(T)[i][j] x_j for the
// rows after it -- independent updates, so the substitution is throughput-
// rather than latency-bound.  A lower solve takes each row's products in
// ascending column order, as trsm_diag (panel.mojo:2221-2276): running
// subtraction from the right-hand side, then the division unless `unit`; an
// upper solve takes them in descending order.  VEC_ROWS picks where the
// vectors live: the columns of B (left solves) or its rows (the Cholesky
// panel X L^-T).
template <typename T, int NB, int UPPER, int UNIT, int TRANS, int VEC_ROWS>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1))) trsm_base(int n, int nvec, const T* __restrict__ Tm, int ldt, T* __restrict__ B,
                                                int ldb) {
  __shared__ T Tc[NB][NB + 16 / sizeof(T)];  // Tc[j][i] = op(T)[i][j]
  const int t = threadIdx.x;
  const int v = blockIdx.x * 64 + t;
  // padded to NB with the identity, so the sweeps below need no bounds; all
  // NB^2/64 loads of a thread are issued before the first LDS store
  {
    constexpr int PER = NB * NB / 64;
    const __amdgpu_buffer_rsrc_t rt = buf_rsrc(Tm);
    T stg[PER];
#pragma unroll
    for (int i = 0; i < PER; i++) {
      const int idx = t + 64 * i, r = idx / NB, c = idx % NB;  // T[r][c], read along rows
      const uint32_t off = (uint32_t)r * (uint32_t)ldt + (uint32_t)c;
      stg[i] = buf_load<T>(rt, (r < n && c < n) ? off * (uint32_t)sizeof(T) : kBufOob);
    }
#pragma unroll
    for (int i = 0; i < PER; i++) {
      const int idx = t + 64 * i, r = idx / NB, c = idx % NB;
      const T val = (r < n && c < n) ? stg[i] : T(r == c ? 1 : 0);
      if (TRANS) Tc[r][c] = val;  // op(T)[c][r]
      else Tc[c][r] = val;        // op(T)[r][c]
    }
  }
  __syncthreads();
  // the block's 64 vectors through one descriptor (32-bit offsets, past-the-
  // end lanes and rows read 0 and store nothing)
  const int v0 = blockIdx.x * 64;
  const __amdgpu_buffer_rsrc_t rb = buf_rsrc(VEC_ROWS ? B + (int64_t)v0 * ldb : B + v0);
  const bool act = v < nvec;
  T x[NB];
#pragma unroll
  for (int i = 0; i < NB; i++) {
    const uint32_t off = VEC_ROWS ? (uint32_t)t * (uint32_t)ldb + (uint32_t)i : (uint32_t)i * (uint32_t)ldb + (uint32_t)t;
    x[i] = buf_load<T>(rb, (act & (i < n)) ? (uint32_t)sizeof(T) * off : kBufOob);
  }
  // rows and columns past n are the identity and x is 0 there: they change
  // nothing.  Each step's column is addressed through an index made opaque
  // on a value of the step before, so the compiler cannot hoist the whole
  // triangle into registers (NB^2/2 of them), and a scheduling barrier makes
  // the step's LDS reads go out together: left to itself the compiler issues
  // them one pair at a time between the products, a full LDS round trip per
  // pair (2.3x slower at NB = 64).  FP32 reads step j's column during step
  // j-1, right after the product that finishes x_j and before the rest, so
  // the latency hides under those products; FP64 (twice the registers) reads
  // it at the start of step j, which measured faster there.  Every x_i takes
  // the same products in the same order as plain substitution.
  constexpr bool kAhead = sizeof(T) == 4;
  if (!UPPER && !kAhead) {
#pragma unroll
    for (int j = 0; j < NB; j++) {
      const T xj = UNIT ? x[j] : x[j] / Tc[j][j];
      x[j] = xj;
      int jj = j;
      asm volatile("" : "+s"(jj) : "v"(xj));
      const T* col = Tc[jj];
      T cv[NB];
#pragma unroll
      for (int i = j + 1; i < NB; i++) cv[i] = col[i];
      __builtin_amdgcn_sched_barrier(0);  // the step's reads all go out before its products
#pragma unroll
      for (int i = j + 1; i < NB; i++) x[i] = x[i] - cv[i] * xj;
    }
  } else if (UPPER && !kAhead) {
#pragma unroll
    for (int j = NB - 1; j >= 0; j--) {
      const T xj = UNIT ? x[j] : x[j] / Tc[j][j];
      x[j] = xj;
      int jj = j;
      asm volatile("" : "+s"(jj) : "v"(xj));
      const T* col = Tc[jj];
      T cv[NB];
#pragma unroll
      for (int i = 0; i < j; i++) cv[i] = col[i];
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int i = 0; i < j; i++) x[i] = x[i] - cv[i] * xj;
    }
  } else if (!UPPER) {
    T cv[NB];  // the current step's column, read one step ahead
    {
      int jj = 0;
      asm volatile("" : "+s"(jj));
      const T* col = Tc[jj];
#pragma unroll
      for (int i = 1; i < NB; i++) cv[i] = col[i];
    }
#pragma unroll
    for (int j = 0; j < NB; j++) {
      const T xj = UNIT ? x[j] : x[j] / Tc[j][j];
      x[j] = xj;
      if (j + 1 < NB) {
        __builtin_amdgcn_sched_barrier(0);
        x[j + 1] = x[j + 1] - cv[j + 1] * xj;  // the next step's unknown first
        T cn[NB];
        int jj = j + 1;
        asm volatile("" : "+s"(jj) : "v"(x[j + 1]));
        const T* col = Tc[jj];
#pragma unroll
        for (int i = j + 2; i < NB; i++) cn[i] = col[i];
        __builtin_amdgcn_sched_barrier(0);  // the reads go out before this step's products
#pragma unroll
        for (int i = j + 2; i < NB; i++) x[i] = x[i] - cv[i] * xj;
#pragma unroll
        for (int i = j + 2; i < NB; i++) cv[i] = cn[i];
      }
    }
  } else {
    T cv[NB];
    {
      int jj = NB - 1;
      asm volatile("" : "+s"(jj));
      const T* col = Tc[jj];
#pragma unroll
      for (int i = 0; i < NB - 1; i++) cv[i] = col[i];
    }
#pragma unroll
    for (int j = NB - 1; j >= 0; j--) {
      const T xj = UNIT ? x[j] : x[j] / Tc[j][j];
      x[j] = xj;
      if (j > 0) {
        __builtin_amdgcn_sched_barrier(0);
        x[j - 1] = x[j - 1] - cv[j - 1] * xj;
        T cn[NB];
        int jj = j - 1;
        asm volatile("" : "+s"(jj) : "v"(x[j - 1]));
        const T* col = Tc[jj];
#pragma unroll
        for (int i = 0; i < j - 1; i++) cn[i] = col[i];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < j - 1; i++) x[i] = x[i] - cv[i] * xj;
#pragma unroll
        for (int i = 0; i < j - 1; i++) cv[i] = cn[i];
      }
    }
  }
  // the store offsets are recomputed from an opaque lane index: shared with
  // the loads' they would stay live (NB registers) through the sweep
  int ts = t;
  asm volatile("" : "+v"(ts));
  const bool acts = ts + v0 < nvec;
#pragma unroll
  for (int i = 0; i < NB; i++) {
    const uint32_t off =
        VEC_ROWS ? (uint32_t)ts * (uint32_t)ldb + (uint32_t)i : (uint32_t)i * (uint32_t)ldb + (uint32_t)ts;
    buf_store(rb, (acts & (i < n)) ? (uint32_t)sizeof(T) * off : kBufOob, x[i]);
  }
}

// potrf_diag (panel.mojo:216-286) for nb <= 64, the block staged in LDS:
// d <= 0 floored to 1e-30 with the first failure recorded 1-based, column
// scaled by the root, then the lower triangle's rank-1 update.
template <typename T>
__global__ void __launch_bounds__(256) potrf_diag_kernel(T* __restrict__ A, int lda, int k0, int nb,
                                                         int32_t* __restrict__ info) {
  // thread t: row i = t % 64 of the block, columns 16 q .. 16 q + 15 in
  // registers (q = t / 64, one wavefront per column quarter).  Column j's
  // owner wavefront scales it and publishes it through LDS (two buffers by
  // parity); one barrier per column, then every row updates its own columns.
  __shared__ T col[2][64];
  __shared__ int info_set;
  const int t = threadIdx.x, i = t & 63, q = t >> 6;
  T R[16];
#pragma unroll
  for (int cc = 0; cc < 16; cc++) {
    const int c = 16 * q + cc;
    R[cc] = (i < nb && c <= i) ? A[(int64_t)(k0 + i) * lda + k0 + c] : T(0);
  }
  if (t == 0) info_set = (*info != 0);
  __syncthreads();
  for (int qj = 0; qj < 4; qj++) {
#pragma unroll
    for (int cj = 0; cj < 16; cj++) {
      const int j = 16 * qj + cj;
      if (j >= nb) break;
      T* cv = col[j & 1];
      if (q == qj) {
        const T d = __shfl(R[cj], j);  // the diagonal entry, from lane j
        T dd = d;
        if (dd <= T(0)) dd = T(kPivotFloor);
        const T root = sqrt(dd);
        if (i > j && i < nb) {
          R[cj] = R[cj] / root;
          cv[i] = R[cj];
        } else if (i == j) {
          R[cj] = root;
          if (d <= T(0) && !info_set) {
            *info = k0 + j + 1;
            info_set = 1;
          }
        }
      }
      __syncthreads();
      if (i > j && i < nb) {
        const T lij = cv[i];
#pragma unroll
        for (int cc = 0; cc < 16; cc++) {
          const int c = 16 * q + cc;
          if (c > j && c <= i) R[cc] = R[cc] - lij * cv[c];
        }
      }
    }
  }
#pragma unroll
  for (int cc = 0; cc < 16; cc++) {
    const int c = 16 * q + cc;
    if (i < nb && c <= i) A[(int64_t)(k0 + i) * lda + k0 + c] = R[cc];
  }
}

template <typename T>
__global__ void diag_kernel(int n, const T* __restrict__ A, int lda, T* __restrict__ d) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = A[(int64_t)i * lda + i];
}

template <typename T>
__global__ void tril_kernel(int n, T* __restrict__ A, int lda) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * n) return;
  int i = (int)(idx / n), c = (int)(idx % n);
  if (c > i) A[(int64_t)i * lda + c] = T(0);
}

__global__ void zero_i32(int32_t* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// inv(op(T)) for n <= 64 into a dense 64 x 64 (ld 64) block: one thread per
// column of the inverse, substitution against the triangle staged in LDS.
template <typename T, int UPPER, int UNIT, int TRANS>
__global__ void __launch_bounds__(1024) tri_inv_small(int n, const T* __restrict__ Tm, int ldt, T* __restrict__ inv) {
  // Gauss-Jordan sweep on [op(T) | I]: one pivot row per step, every other
  // row updated in parallel (n <= 64 steps, each a barrier apart).
  __shared__ T S[64][65];
  __shared__ T X[64][65];
  const int t = threadIdx.x;
  for (int idx = t; idx < 64 * 64; idx += 1024) {
    const int i = idx / 64, q = idx % 64;
    S[i][q] = (i < n && q < n) ? (TRANS ? Tm[(int64_t)q * ldt + i] : Tm[(int64_t)i * ldt + q]) : T(0);
    X[i][q] = (i == q) ? T(1) : T(0);
  }
  __syncthreads();
  for (int s = 0; s < n; s++) {
    const int j = UPPER ? n - 1 - s : s;
    if (!UNIT) {
      if (t < 64) X[j][t] = X[j][t] / S[j][j];
      __syncthreads();
    }
    for (int idx = t; idx < 64 * 64; idx += 1024) {
      const int r = idx / 64, c = idx % 64;
      const bool below = UPPER ? (r < j) : (r > j && r < n);
      if (below) X[r][c] = X[r][c] - S[r][j] * X[j][c];
    }
    __syncthreads();
  }
  for (int idx = t; idx < 64 * 64; idx += 1024) {
    const int i = idx / 64, q = idx % 64;
    inv[idx] = (i < n && q < n) ? X[i][q] : T(0);
  }
}

// Apply a 64 x 64 (ld 64) inverse to 64 vectors per workgroup, in place:
// each vector x of B becomes inv x.  Thread t owns vector t % 64 and 16 of
// its n entries; the inverse and the vectors are staged in LDS.
template <typename T, int VEC_ROWS>
__global__ void __launch_bounds__(256) tri_apply_small(int n, int nvec, const T* __restrict__ inv,
                                                       T* __restrict__ B, int ldb) {
  __shared__ T Is[64][65];
  __shared__ T Xs[64][65];
  const int t = threadIdx.x;
  const int v0 = blockIdx.x * 64;
  for (int idx = t; idx < 64 * 64; idx += 256) Is[idx / 64][idx % 64] = inv[idx];
  if (VEC_ROWS) {
    for (int idx = t; idx < n * 64; idx += 256) {
      int v = idx / n, i = idx % n;
      if (v0 + v < nvec) Xs[i][v] = B[(int64_t)(v0 + v) * ldb + i];
    }
  } else {
    for (int idx = t; idx < n * 64; idx += 256) {
      int i = idx / 64, v = idx % 64;
      if (v0 + v < nvec) Xs[i][v] = B[(int64_t)i * ldb + v0 + v];
    }
  }
  __syncthreads();
  const int v = t & 63, i0 = (t >> 6) * 16;
  T y[16];
#pragma unroll
  for (int r = 0; r < 16; r++) y[r] = T(0);
  for (int q = 0; q < n; q++) {
    const T xq = Xs[q][v];
#pragma unroll
    for (int r = 0; r < 16; r++) y[r] += Is[i0 + r][q] * xq;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 16; r++) Xs[i0 + r][v] = y[r];
  __syncthreads();
  if (VEC_ROWS) {
    for (int idx = t; idx < n * 64; idx += 256) {
      int vv = idx / n, i = idx % n;
      if (v0 + vv < nvec) B[(int64_t)(v0 + vv) * ldb + i] = Xs[i][vv];
    }
  } else {
    for (int idx = t; idx < n * 64; idx += 256) {
      int i = idx / 64, vv = idx % 64;
      if (v0 + vv < nvec) B[(int64_t)i * ldb + v0 + vv] = Xs[i][vv];
    }
  }
}

constexpr int kBaseInverseMinVec = 1 << 30;  // substitution for every width

template <typename T, int UPPER, int UNIT, int TRANS, int VEC_ROWS>
void launch_base(nmx_ctx* ctx, int n, int nvec, const T* Tm, int ldt, T* B, int ldb) {
  // the explicit inverse only pays off for very wide right-hand sides: the
  // register-resident substitution above runs 64 vectors per wavefront
  if (nvec > kBaseInverseMinVec && ctx->tri_inv) {
    T* inv = (T*)ctx->tri_inv;
    ProfScope ps(ctx, K_TRSM, 2.0 * n * (double)n * nvec, sizeof(T) * (2.0 * n * nvec + 64.0 * 64));
    tri_inv_small<T, UPPER, UNIT, TRANS><<<1, 1024, 0, ctx->stream>>>(n, Tm, ldt, inv);
    tri_apply_small<T, VEC_ROWS><<<ceil_div(nvec, 64), 256, 0, ctx->stream>>>(n, nvec, inv, B, ldb);
    return;
  }
  ProfScope ps(ctx, K_TRSM, (double)n * n * nvec, sizeof(T) * ((double)n * n + 2.0 * n * nvec));
  if (n <= 32)
    trsm_base<T, 32, UPPER, UNIT, TRANS, VEC_ROWS><<<ceil_div(nvec, 64), 64, 0, ctx->stream>>>(n, nvec, Tm, ldt, B,
                                                                                             ldb);
  else
    trsm_base<T, 64, UPPER, UNIT, TRANS, VEC_ROWS><<<ceil_div(nvec, 64), 64, 0, ctx->stream>>>(n, nvec, Tm, ldt, B,
                                                                                             ldb);
}

template <typename T, int VEC_ROWS>
int base_dispatch(nmx_ctx* ctx, int upper, int unit, int trans, int n, int nvec, const T* Tm, int ldt, T* B,
                  int ldb) {
  int key = upper * 4 + unit * 2 + trans;
  switch (key) {
    case 0: launch_base<T, 0, 0, 0, VEC_ROWS>(ctx, n, nvec, Tm, ldt, B, ldb); break;
    case 1: launch_base<T, 0, 0, 1, VEC_ROWS>(ctx, n, nvec, Tm, ldt, B, ldb); break;
    case 2: launch_base<T, 0, 1, 0, VEC_ROWS>(ctx, n, nvec, Tm, ldt, B, ldb); break;
    case 3: launch_base<T, 0, 1, 1, VEC_ROWS>(ctx, n, nvec, Tm, ldt, B, ldb); break;
    case 4: launch_base<T, 1, 0, 0, VEC_ROWS>(ctx, n, nvec, Tm, ldt, B, ldb); break;
    case 5: launch_base<T, 1, 0, 1, VEC_ROWS>(ctx, n, nvec, Tm, ldt, B, ldb); break;
    case 6: launch_base<T, 1, 1, 0, VEC_ROWS>(ctx, n, nvec, Tm, ldt, B, ldb); break;
    default: launch_base<T, 1, 1, 1, VEC_ROWS>(ctx, n, nvec, Tm, ldt, B, ldb); break;
  }
  NMX_HIP(ctx, hipGetLastError());
  return NMX_OK;
}

// Recursive triangular solve: split op(T) in two, solve one diagonal block,
// update the other half of the right-hand sides with the MFMA GEMM, solve the
// other block.  VEC_ROWS = 0: the vectors are B's columns (B is n x nvec);
// VEC_ROWS = 1: they are B's rows (B is nvec x n).
template <typename T, int VEC_ROWS>
int trsm_rec(nmx_ctx* ctx, int upper, int unit, int trans, int n, int nvec, const T* Tm, int ldt, T* B, int ldb) {
  if (n <= 0 || nvec <= 0) return NMX_OK;
  if (n <= 64) return base_dispatch<T, VEC_ROWS>(ctx, upper, unit, trans, n, nvec, Tm, ldt, B, ldb);
  int n1 = ((n / 2 + 63) / 64) * 64;
  if (n1 >= n) n1 = n - 64;
  const int n2 = n - n1;
  const T* T11 = Tm;
  const T* T22 = Tm + (int64_t)n1 * ldt + n1;
  const T* Tlo = Tm + (int64_t)n1 * ldt;  // stored rows n1.., cols 0..n1
  const T* Tup = Tm + n1;                 // stored rows 0..n1, cols n1..
  // vector halves
  T* B1 = B;
  T* B2 = VEC_ROWS ? B + n1 : B + (int64_t)n1 * ldb;
  // op(T)21 (forward) or op(T)12 (backward) as stored block + transpose flag
  if (!upper) {
    NMX_TRY((trsm_rec<T, VEC_ROWS>(ctx, upper, unit, trans, n1, nvec, T11, ldt, B1, ldb)));
    // op(T)21 = trans ? (T12)^T : T21
    const T* blk = trans ? Tup : Tlo;
    if (!VEC_ROWS) {
      NMX_TRY(gemm_launch<T>(ctx, trans, 0, n2, nvec, n1, T(-1), blk, ldt, B1, ldb, T(1), B2, ldb, 0));
    } else {
      // rows: X2 -= X1 op(T)21^T
      NMX_TRY(gemm_launch<T>(ctx, 0, trans ? 0 : 1, nvec, n2, n1, T(-1), B1, ldb, blk, ldt, T(1), B2, ldb, 0));
    }
    NMX_TRY((trsm_rec<T, VEC_ROWS>(ctx, upper, unit, trans, n2, nvec, T22, ldt, B2, ldb)));
  } else {
    NMX_TRY((trsm_rec<T, VEC_ROWS>(ctx, upper, unit, trans, n2, nvec, T22, ldt, B2, ldb)));
    // op(T)12 = trans ? (T21)^T : T12
    const T* blk = trans ? Tlo : Tup;
    if (!VEC_ROWS) {
      NMX_TRY(gemm_launch<T>(ctx, trans, 0, n1, nvec, n2, T(-1), blk, ldt, B2, ldb, T(1), B1, ldb, 0));
    } else {
      NMX_TRY(gemm_launch<T>(ctx, 0, trans ? 0 : 1, nvec, n1, n2, T(-1), B2, ldb, blk, ldt, T(1), B1, ldb, 0));
    }
    NMX_TRY((trsm_rec<T, VEC_ROWS>(ctx, upper, unit, trans, n1, nvec, T11, ldt, B1, ldb)));
  }
  return NMX_OK;
}

}  // namespace

template <typename T>
int trsm_left(nmx_ctx* ctx, int upper, int unit, int trans, int n, int nrhs, const T* Tm, int ldt, T* B, int ldb) {
  return trsm_rec<T, 0>(ctx, upper, unit, trans, n, nrhs, Tm, ldt, B, ldb);
}

template <typename T>
int trsm_right_lower_t(nmx_ctx* ctx, int m, int n, const T* L, int ldl, T* X, int ldx) {
  // X := X L^-T  <=>  each row x of X solves L x^T = x^T (forward, non-unit)
  return trsm_rec<T, 1>(ctx, 0, 0, 0, n, m, L, ldl, X, ldx);
}

// X := X U^-1 with U upper n x n (rows of X solve U^T x^T = a^T).
int trsm_rec_rows_upper(nmx_ctx* ctx, int m, int n, const float* U, int ldu, float* X, int ldx) {
  return trsm_rec<float, 1>(ctx, 0, 0, 1, n, m, U, ldu, X, ldx);
}

template <typename T>
int laswp_launch(nmx_ctx* ctx, T* A, int lda, int col_lo, int col_hi, int skip_lo, int skip_hi, const int32_t* ipiv,
                 int k1, int k2) {
  if (k2 <= k1 || col_hi <= col_lo) return NMX_OK;
  int s_lo = skip_lo < col_lo ? col_lo : skip_lo;
  int s_hi = skip_hi > col_hi ? col_hi : skip_hi;
  int skip_w = s_hi > s_lo ? s_hi - s_lo : 0;
  if (skip_w == 0) s_lo = col_hi;
  int ncols = col_hi - col_lo - skip_w;
  if (ncols <= 0) return NMX_OK;
  ProfScope ps(ctx, K_LASWP, 0.0, 4.0 * sizeof(T) * (double)ncols * (k2 - k1));
  if (k2 - k1 <= kLeafW)
    laswp_gather_kernel<T><<<ceil_div(ncols, 256), 256, 0, ctx->stream>>>(A, lda, col_lo, ncols, s_lo, skip_w, ipiv, k1,
                                                                           k2);
  else
    laswp_kernel<T><<<ceil_div(ncols, 256), 256, 0, ctx->stream>>>(A, lda, col_lo, ncols, s_lo, skip_w, ipiv, k1, k2);
  NMX_HIP(ctx, hipGetLastError());
  return NMX_OK;
}

int perm_from_ipiv(nmx_ctx* ctx, int n, const int32_t* ipiv, int32_t* perm) {
  if (n <= 0) return NMX_OK;
  ProfScope ps(ctx, K_LASWP, 0.0, 8.0 * n);
  const size_t lds = sizeof(int32_t) * (size_t)n;
  if (lds <= kPermLdsBytes)
    perm_chunks_kernel<<<1, 1024, lds, ctx->stream>>>(n, ipiv, perm);
  else
    perm_kernel<0><<<1, 1024, 0, ctx->stream>>>(n, ipiv, perm);
  NMX_HIP(ctx, hipGetLastError());
  return NMX_OK;
}

template <typename Ti, typename To>
int gather_rows(nmx_ctx* ctx, int n, int ncols, const int32_t* perm, const Ti* in, int ldi, To* out, int ldo) {
  const int64_t total = (int64_t)n * ncols;
  if (total <= 0) return NMX_OK;
  ProfScope ps(ctx, K_LASWP, 0.0, (double)total * (sizeof(Ti) + sizeof(To)) + 4.0 * n);
  gather_rows_kernel<Ti, To><<<(unsigned)((total + 255) / 256), 256, 0, ctx->stream>>>(n, ncols, perm, in, ldi, out,
                                                                                       ldo);
  NMX_HIP(ctx, hipGetLastError());
  return NMX_OK;
}
template int gather_rows<double, double>(nmx_ctx*, int, int, const int32_t*, const double*, int, double*, int);
template int gather_rows<float, float>(nmx_ctx*, int, int, const int32_t*, const float*, int, float*, int);
template int gather_rows<double, float>(nmx_ctx*, int, int, const int32_t*, const double*, int, float*, int);

template <typename T>
int getf2_leaf_launch(nmx_ctx* ctx, int n, T* A, int lda, int r0, int c0, int w, int32_t* ipiv, int32_t* info,
                      int swap_others) {
  const int m = n - r0;
  if (m <= 0 || w <= 0) return NMX_OK;
  if (w > kLeafW) return NMX_ERR_UNSUPPORTED;
  // one row per thread once the panel spans several workgroups (and fits 32
  // of 512 rows), two otherwise
  const bool one_row = m > kLeafT * 2 && m <= kLeafT * kMaxParts;
  int parts = ceil_div(m, one_row ? kLeafT : kLeafT * 2);
  if (parts > kMaxParts) return NMX_ERR_UNSUPPORTED;  // n - r0 <= kLeafMaxRows
  uint32_t epoch0 = 0;
  NMX_TRY(nmx_epoch_reserve(ctx, (uint32_t)w, &epoch0));
  uint64_t* mb = (uint64_t*)ctx->mailbox;
  int* err = ctx->err;
  {
    ProfScope ps(ctx, K_LEAF, (double)m * w * w, 2.0 * sizeof(T) * (double)m * w);
    // At most 8 x 32 workgroups of 512 threads: one per CU of an otherwise
    // idle device (the stream orders this launch after the previous kernel).
    // A participant that is not resident makes the others' bounded waits
    // expire, which the context reports as NMX_ERR_SYNC_TIMEOUT.
    const int grid = parts == 1 ? 1 : parts * kXcds;
    if (one_row)
      getf2_leaf<T, 1><<<grid, kLeafT, 0, ctx->stream>>>(A, lda, n, r0, c0, w, ipiv, info, mb, epoch0, parts, err);
    else
      getf2_leaf<T, 2><<<grid, kLeafT, 0, ctx->stream>>>(A, lda, n, r0, c0, w, ipiv, info, mb, epoch0, parts, err);
    NMX_HIP(ctx, hipGetLastError());
  }
  if (swap_others) return laswp_launch<T>(ctx, A, lda, 0, n, c0, c0 + w, ipiv, r0, r0 + w);
  return NMX_OK;
}

template <typename T>
int diag_launch(nmx_ctx* ctx, int n, const T* A, int lda, T* d) {
  if (n <= 0) return NMX_OK;
  diag_kernel<T><<<ceil_div(n, 256), 256, 0, ctx->stream>>>(n, A, lda, d);
  NMX_HIP(ctx, hipGetLastError());
  return NMX_OK;
}

int zero_info(nmx_ctx* ctx, int32_t* info) {
  zero_i32<<<1, 64, 0, ctx->stream>>>(info, 1);
  NMX_HIP(ctx, hipGetLastError());
  return NMX_OK;
}

template <typename T>
int potrf_diag_launch(nmx_ctx* ctx, T* A, int lda, int k0, int nb, int32_t* info) {
  if (nb > 64) return NMX_ERR_UNSUPPORTED;
  ProfScope ps(ctx, K_POTRF_DIAG, (double)nb * nb * nb / 3.0, 2.0 * sizeof(T) * nb * nb);
  potrf_diag_kernel<T><<<1, 256, 0, ctx->stream>>>(A, lda, k0, nb, info);
  NMX_HIP(ctx, hipGetLastError());
  return NMX_OK;
}

template <typename T>
int tril_launch(nmx_ctx* ctx, int n, T* A, int lda) {
  int64_t total = (int64_t)n * n;
  if (total <= 0) return NMX_OK;
  tril_kernel<T><<<(unsigned)((total + 255) / 256), 256, 0, ctx->stream>>>(n, A, lda);
  NMX_HIP(ctx, hipGetLastError());
  return NMX_OK;
}

#define INST(T)                                                                                                     \
  template int trsm_left<T>(nmx_ctx*, int, int, int, int, int, const T*, int, T*, int);                            \
  template int trsm_right_lower_t<T>(nmx_ctx*, int, int, const T*, int, T*, int);                                  \
  template int laswp_launch<T>(nmx_ctx*, T*, int, int, int, int, int, const int32_t*, int, int);                   \
  template int getf2_leaf_launch<T>(nmx_ctx*, int, T*, int, int, int, int, int32_t*, int32_t*, int);               \
  template int diag_launch<T>(nmx_ctx*, int, const T*, int, T*);                                                   \
  template int potrf_diag_launch<T>(nmx_ctx*, T*, int, int, int, int32_t*);                                        \
  template int tril_launch<T>(nmx_ctx*, int, T*, int);
INST(double)
INST(float)
