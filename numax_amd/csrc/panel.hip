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

// ---- mailbox layout --------------------------------------------------------
struct Slot {
  double mag;
  int row;
  int pad[13];
  double vals[kLeafW];
};
static_assert(sizeof(Slot) == kMailboxSlotBytes, "mailbox slot size");

__device__ __forceinline__ uint32_t* mb_flag(unsigned char* mb, int p) { return (uint32_t*)(mb + (size_t)p * 64); }
__device__ __forceinline__ Slot* mb_slot(unsigned char* mb, int parity, int s) {
  return (Slot*)(mb + kMailboxFlagsBytes) + (size_t)parity * (kMaxParts + 1) + s;
}

// ---------------------------------------------------------------------------
// Cooperative leaf of the recursive LU panel.
//
// Factors columns c0..c0+w (w <= 32) over rows r0..n with partial pivoting,
// exactly getrf_panel's arithmetic (panel.mojo:289-431): first-maximum pivot
// (ties to the lower row; the diagonal row counts as magnitude >= 0, so an
// all-zero or all-NaN column keeps its diagonal), interchange, an exact-zero
// pivot divided as 1e-30 with `info` recording the first, true division for
// the multipliers and the rank-1 update of the panel's own columns.  Only the
// leaf's columns are interchanged here; the driver replays the leaf's pivots
// over every other column afterwards (laswp), which leaves the same matrix as
// numax's full-width swaps.
//
// Workgroup p of `parts` owns rows r0 + p*R ... (R = RPT*1024), thread t
// holds row r0 + p*R + q*1024 + t in registers.  With parts > 1 the grid is
// 8*parts workgroups of which every eighth participates, so on MI355X (where
// workgroup b lands on XCD b % 8) all participants share one XCD and its L2;
// correctness does not depend on that placement, since every exchange is an
// agent-scope release/acquire.
// ---------------------------------------------------------------------------
template <typename T, int RPT>
__global__ void __launch_bounds__(1024) getf2_leaf(T* __restrict__ A, int lda, int n, int r0, int c0, int w,
                                                   int32_t* __restrict__ ipiv, int32_t* __restrict__ info,
                                                   unsigned char* mb, uint32_t epoch0, int parts, int* err) {
  constexpr int W = kLeafW;
  constexpr int R = RPT * kLeafThreads;
  const bool multi = parts > 1;
  if (multi && (blockIdx.x % kXcds) != 0) return;
  const int p = multi ? (int)(blockIdx.x / kXcds) : 0;
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;

  __shared__ T red_mag[16];
  __shared__ int red_row[16];
  __shared__ T cand_mag[kMaxParts];
  __shared__ int cand_row[kMaxParts];
  __shared__ T u_s[W];  // pivot row
  __shared__ T d_s[W];  // old diagonal row
  __shared__ int win_row, win_slot, aborted;
  __shared__ int info_set;

  if (t == 0) {
    aborted = 0;
    info_set = (*info != 0);
  }

  T v[RPT][W];
  int rowq[RPT];
#pragma unroll
  for (int q = 0; q < RPT; q++) {
    rowq[q] = r0 + p * R + q * kLeafThreads + t;
    const bool ok = rowq[q] < n;
    const T* src = A + (int64_t)rowq[q] * lda + c0;
#pragma unroll
    for (int c = 0; c < W; c++) v[q][c] = (ok && c < w) ? src[c] : T(0);
  }
  __syncthreads();

#pragma unroll W
  for (int j = 0; j < W; j++) {
    if (j < w) {
    const int gr = r0 + j;
    const int gc = c0 + j;
    const uint32_t epoch = epoch0 + (uint32_t)j + 1u;
    const int parity = (int)(epoch & 1u);

    // 1. this thread's candidate, then the wavefront's, then the workgroup's
    T bm = T(-1);
    int br = INT_MAX;
#pragma unroll
    for (int q = 0; q < RPT; q++) {
      if (rowq[q] >= gr && rowq[q] < n) {
        T mag = fabs(v[q][j]);
        if (!(mag == mag)) mag = T(-1);        // NaN never wins ...
        if (rowq[q] == gr && mag < T(0)) mag = T(0);  // ... the diagonal row starts at 0
        if (better(mag, rowq[q], bm, br)) { bm = mag; br = rowq[q]; }
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      T om = __shfl_xor(bm, off);
      int orr = __shfl_xor(br, off);
      if (better(om, orr, bm, br)) { bm = om; br = orr; }
    }
    if (lane == 0) { red_mag[wave] = bm; red_row[wave] = br; }
    __syncthreads();
    if (t < 64) {
      bm = t < 16 ? red_mag[t] : T(-1);
      br = t < 16 ? red_row[t] : INT_MAX;
#pragma unroll
      for (int off = 8; off > 0; off >>= 1) {
        T om = __shfl_xor(bm, off);
        int orr = __shfl_xor(br, off);
        if (better(om, orr, bm, br)) { bm = om; br = orr; }
      }
      if (t == 0) { red_mag[0] = bm; red_row[0] = br; }
    }
    __syncthreads();
    const T wg_mag = red_mag[0];
    const int wg_row = red_row[0];

    if (!multi) {
      // the pivot row and the old diagonal row straight through LDS
#pragma unroll
      for (int q = 0; q < RPT; q++) {
        if (rowq[q] == wg_row) {
#pragma unroll
          for (int c = 0; c < W; c++) u_s[c] = v[q][c];
        }
        if (rowq[q] == gr) {
#pragma unroll
          for (int c = 0; c < W; c++) d_s[c] = v[q][c];
        }
      }
      if (t == 0) win_row = wg_row;
      __syncthreads();
    } else {
      // 2. publish this workgroup's candidate (and the diagonal row if owned)
#pragma unroll
      for (int q = 0; q < RPT; q++) {
        if (rowq[q] == wg_row && wg_row != INT_MAX) {
          Slot* s = mb_slot(mb, parity, p);
          s->mag = (double)wg_mag;
          s->row = wg_row;
#pragma unroll
          for (int c = 0; c < W; c++) s->vals[c] = (double)v[q][c];
        }
        if (rowq[q] == gr) {
          Slot* s = mb_slot(mb, parity, kMaxParts);
#pragma unroll
          for (int c = 0; c < W; c++) s->vals[c] = (double)v[q][c];
        }
      }
      if (wg_row == INT_MAX && t == 0) {
        Slot* s = mb_slot(mb, parity, p);
        s->mag = -1.0;
        s->row = INT_MAX;
      }
      __syncthreads();  // every store of the workgroup has completed
      if (t == 0) __hip_atomic_store(mb_flag(mb, p), epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      // 3. wait for every participant's flag, then read its candidate
      if (t < parts) {
        long spins = 0;
        while ((int)(__hip_atomic_load(mb_flag(mb, t), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - epoch) < 0) {
          if (aborted || ++spins > kSpinLimit) { aborted = 1; break; }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        const Slot* s = mb_slot(mb, parity, t);
        cand_mag[t] = (T)s->mag;
        cand_row[t] = s->row;
      }
      __syncthreads();
      if (t < 64) {
        T cm = t < parts ? cand_mag[t] : T(-1);
        int cr = t < parts ? cand_row[t] : INT_MAX;
        int cs = t;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
          T om = __shfl_xor(cm, off);
          int orr = __shfl_xor(cr, off);
          int os = __shfl_xor(cs, off);
          if (better(om, orr, cm, cr)) { cm = om; cr = orr; cs = os; }
        }
        if (t == 0) { win_row = cr; win_slot = cs; }
      }
      __syncthreads();
      if (t < W) {
        const Slot* ws = mb_slot(mb, parity, win_slot);
        const Slot* ds = mb_slot(mb, parity, kMaxParts);
        u_s[t] = (T)ws->vals[t];
        d_s[t] = (T)ds->vals[t];
      }
      __syncthreads();
    }

    const int wrow = (win_row == INT_MAX || aborted) ? gr : win_row;
    // 4. interchange inside the leaf's columns, record the pivot
    if (wrow != gr) {
#pragma unroll
      for (int q = 0; q < RPT; q++) {
        if (rowq[q] == gr) {
#pragma unroll
          for (int c = 0; c < W; c++) v[q][c] = u_s[c];
        } else if (rowq[q] == wrow) {
#pragma unroll
          for (int c = 0; c < W; c++) v[q][c] = d_s[c];
        }
      }
    }
    T pivot = (wrow != gr) ? u_s[j] : d_s[j];
    if (p == 0 && t == 0) {
      ipiv[gc] = wrow;
      if (fabs(pivot) == T(0) && !info_set) {
        *info = gc + 1;
        info_set = 1;
      }
    }
    if (pivot == T(0)) pivot = T(kPivotFloor);
    // 5. multipliers and the rank-1 update of the leaf's remaining columns
    const T* urow = (wrow != gr) ? u_s : d_s;  // LDS broadcast reads
#pragma unroll
    for (int q = 0; q < RPT; q++) {
      if (rowq[q] > gr && rowq[q] < n) {
        const T l = v[q][j] / pivot;
        v[q][j] = l;
#pragma unroll
        for (int c = j + 1; c < W; c++) v[q][c] = v[q][c] - l * urow[c];
      }
    }
    __syncthreads();
    }
  }

  if (aborted && t == 0) atomicExch(err, NMX_ERR_SYNC_TIMEOUT);
#pragma unroll
  for (int q = 0; q < RPT; q++) {
    if (rowq[q] < n) {
      T* dst = A + (int64_t)rowq[q] * lda + c0;
#pragma unroll
      for (int c = 0; c < W; c++)
        if (c < w) dst[c] = v[q][c];
    }
  }
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

// Base triangular solve for n <= 64: one thread per vector, the triangle and
// 64 vectors staged in LDS.  Each step is trsm_diag's substitution
// (panel.mojo:2221-2276): total - sum of products taken in ascending p, then
// the division unless `unit`.  VEC_ROWS picks where the vectors live: the
// columns of B (left solves) or its rows (the Cholesky panel X L^-T).
template <typename T, int UPPER, int UNIT, int TRANS, int VEC_ROWS>
__global__ void __launch_bounds__(64) trsm_base(int n, int nvec, const T* __restrict__ Tm, int ldt, T* __restrict__ B,
                                                int ldb) {
  __shared__ T Ts[64][65];
  __shared__ T Xs[64][65];
  const int t = threadIdx.x;
  const int v0 = blockIdx.x * 64;
  for (int idx = t; idx < n * n; idx += 64) {
    int i = idx / n, q = idx % n;
    Ts[i][q] = TRANS ? Tm[(int64_t)q * ldt + i] : Tm[(int64_t)i * ldt + q];
  }
  if (VEC_ROWS) {
    for (int idx = t; idx < n * 64; idx += 64) {
      int v = idx / n, i = idx % n;
      if (v0 + v < nvec) Xs[i][v] = B[(int64_t)(v0 + v) * ldb + i];
    }
  } else {
    for (int idx = t; idx < n * 64; idx += 64) {
      int i = idx / 64, v = idx % 64;
      if (v0 + v < nvec) Xs[i][v] = B[(int64_t)i * ldb + v0 + v];
    }
  }
  __syncthreads();
  if (v0 + t < nvec) {
    if (!UPPER) {
      for (int i = 0; i < n; i++) {
        T s = Xs[i][t];
        for (int q = 0; q < i; q++) s = s - Ts[i][q] * Xs[q][t];
        Xs[i][t] = UNIT ? s : s / Ts[i][i];
      }
    } else {
      for (int i = n - 1; i >= 0; i--) {
        T s = Xs[i][t];
        for (int q = i + 1; q < n; q++) s = s - Ts[i][q] * Xs[q][t];
        Xs[i][t] = UNIT ? s : s / Ts[i][i];
      }
    }
  }
  __syncthreads();
  if (VEC_ROWS) {
    for (int idx = t; idx < n * 64; idx += 64) {
      int v = idx / n, i = idx % n;
      if (v0 + v < nvec) B[(int64_t)(v0 + v) * ldb + i] = Xs[i][v];
    }
  } else {
    for (int idx = t; idx < n * 64; idx += 64) {
      int i = idx / 64, v = idx % 64;
      if (v0 + v < nvec) B[(int64_t)i * ldb + v0 + v] = Xs[i][v];
    }
  }
}

// potrf_diag (panel.mojo:216-286) for nb <= 64, the block staged in LDS:
// d <= 0 floored to 1e-30 with the first failure recorded 1-based, column
// scaled by the root, then the lower triangle's rank-1 update.
template <typename T>
__global__ void __launch_bounds__(256) potrf_diag_kernel(T* __restrict__ A, int lda, int k0, int nb,
                                                         int32_t* __restrict__ info) {
  __shared__ T S[64][65];
  __shared__ int info_set;
  const int t = threadIdx.x;
  for (int idx = t; idx < nb * nb; idx += 256) {
    int i = idx / nb, c = idx % nb;
    if (c <= i) S[i][c] = A[(int64_t)(k0 + i) * lda + k0 + c];
  }
  if (t == 0) info_set = (*info != 0);
  __syncthreads();
  for (int j = 0; j < nb; j++) {
    if (t == 0) {
      T d = S[j][j];
      if (d <= T(0)) {
        if (!info_set) { *info = k0 + j + 1; info_set = 1; }
        d = T(kPivotFloor);
      }
      S[j][j] = sqrt(d);
    }
    __syncthreads();
    const T root = S[j][j];
    for (int i = j + 1 + t; i < nb; i += 256) S[i][j] = S[i][j] / root;
    __syncthreads();
    for (int idx = t; idx < nb * nb; idx += 256) {
      int i = idx / nb, c = idx % nb;
      if (i > j && c > j && c <= i) S[i][c] = S[i][c] - S[i][j] * S[c][j];
    }
    __syncthreads();
  }
  for (int idx = t; idx < nb * nb; idx += 256) {
    int i = idx / nb, c = idx % nb;
    if (c <= i) A[(int64_t)(k0 + i) * lda + k0 + c] = S[i][c];
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

template <typename T, int UPPER, int UNIT, int TRANS, int VEC_ROWS>
void launch_base(nmx_ctx* ctx, int n, int nvec, const T* Tm, int ldt, T* B, int ldb) {
  ProfScope ps(ctx, K_TRSM, (double)n * n * nvec, sizeof(T) * ((double)n * n + 2.0 * n * nvec));
  trsm_base<T, UPPER, UNIT, TRANS, VEC_ROWS><<<ceil_div(nvec, 64), 64, 0, ctx->stream>>>(n, nvec, Tm, ldt, B, ldb);
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
  laswp_kernel<T><<<ceil_div(ncols, 256), 256, 0, ctx->stream>>>(A, lda, col_lo, ncols, s_lo, skip_w, ipiv, k1, k2);
  NMX_HIP(ctx, hipGetLastError());
  return NMX_OK;
}

template <typename T>
int getf2_leaf_launch(nmx_ctx* ctx, int n, T* A, int lda, int r0, int c0, int w, int32_t* ipiv, int32_t* info) {
  const int m = n - r0;
  if (m <= 0 || w <= 0) return NMX_OK;
  if (w > kLeafW) return NMX_ERR_UNSUPPORTED;
  int rpt = 1;
  int parts = ceil_div(m, kLeafThreads);
  if (parts > kMaxParts) {
    if (sizeof(T) == 8) return NMX_ERR_UNSUPPORTED;  // FP64 leaf holds one row per thread
    rpt = 2;
    parts = ceil_div(m, 2 * kLeafThreads);
    if (parts > kMaxParts) return NMX_ERR_UNSUPPORTED;
  }
  uint32_t epoch0 = 0;
  NMX_TRY(nmx_epoch_reserve(ctx, (uint32_t)w, &epoch0));
  unsigned char* mb = (unsigned char*)ctx->mailbox;
  int* err = ctx->err;
  ProfScope ps(ctx, K_LEAF, (double)m * w * w, 2.0 * sizeof(T) * (double)m * w);
  if (parts == 1) {
    if (rpt == 1)
      getf2_leaf<T, 1><<<1, kLeafThreads, 0, ctx->stream>>>(A, lda, n, r0, c0, w, ipiv, info, mb, epoch0, parts, err);
    else
      getf2_leaf<T, 2><<<1, kLeafThreads, 0, ctx->stream>>>(A, lda, n, r0, c0, w, ipiv, info, mb, epoch0, parts, err);
    NMX_HIP(ctx, hipGetLastError());
    return NMX_OK;
  }
  void* args[] = {&A, &lda, (void*)&n, &r0, &c0, &w, &ipiv, &info, &mb, &epoch0, &parts, &err};
  const void* fn = rpt == 1 ? (const void*)getf2_leaf<T, 1> : (const void*)getf2_leaf<T, 2>;
  NMX_HIP(ctx, hipLaunchCooperativeKernel(fn, dim3(parts * kXcds), dim3(kLeafThreads), args, 0, ctx->stream));
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
  template int getf2_leaf_launch<T>(nmx_ctx*, int, T*, int, int, int, int, int32_t*, int32_t*);                    \
  template int diag_launch<T>(nmx_ctx*, int, const T*, int, T*);                                                   \
  template int potrf_diag_launch<T>(nmx_ctx*, T*, int, int, int, int32_t*);                                        \
  template int tril_launch<T>(nmx_ctx*, int, T*, int);
INST(double)
INST(float)
