// Internal declarations shared by the HIP translation units of libnumax_hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/numax_hip.h"

#define NMX_API extern "C" __attribute__((visibility("default")))

// Return on a failing HIP call with NMX_ERR_HIP (the failing call's error is
// kept in ctx->last_hip for diagnostics).
#define NMX_HIP(ctx, call)                          \
  do {                                              \
    hipError_t e_ = (call);                         \
    if (e_ != hipSuccess) {                         \
      if (ctx) (ctx)->last_hip = (int)e_;           \
      return NMX_ERR_HIP;                           \
    }                                               \
  } while (0)

#define NMX_TRY(expr)           \
  do {                          \
    int s_ = (expr);            \
    if (s_ != NMX_OK) return s_; \
  } while (0)

// Leaf width of the pivoting panel and the cross-workgroup mailbox shape.
constexpr int kLeafW = 32;           // columns one cooperative leaf factors
constexpr int kLeafThreads = 1024;   // threads per participating workgroup
constexpr int kMaxParts = 32;        // participating workgroups at most
constexpr int kXcds = 8;             // MI355X accelerator complex dies
// Rows the cooperative pivoting leaf holds: kMaxParts workgroups of 512
// threads with two rows each.  lu_factor / getrf take n up to this.
constexpr int kLeafMaxRows = 32768;

// Kernel classes timed by the context's optional event profiler.
enum NmxKernel {
  K_GEMM = 0,     // MFMA GEMM update (FP64 / FP32)
  K_LEAF,         // cooperative pivoting leaf
  K_LASWP,        // row interchanges
  K_TRSM,         // triangular-solve base blocks
  K_POTRF_DIAG,   // Cholesky diagonal block
  K_MP_GEMM,      // mixed-precision (FP16/BF16 MFMA) trailing update
  K_MP_PANEL,     // mixed-precision diagonal block / panel
  K_RESIDUAL,     // FP64 residual b - A x
  K_MP_TRSV,      // low-precision triangular solves inside refinement
  K_OTHER,
  K_BATCH,        // batched uneven-size solves (all size classes of one call)
  K_EIGH,         // batched Jacobi eigensolver
  K_TSQR,         // batched tall-skinny least squares
  K_QR_PANEL,     // Householder panel (geqr2)
  K_COUNT
};

struct ProfRec {
  hipEvent_t a, b;
  int kind;
  double flops, bytes;
};

struct nmx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int num_cus = 0;
  int last_hip = 0;
  // grow-only scratch
  void* ws = nullptr;
  size_t ws_bytes = 0;
  // captured graphs still alive that may replay kernels reading the
  // context's buffers: while any exists, a workspace that grows keeps its
  // old buffer (retired, freed with the last graph), and nmx_ctx_destroy
  // defers the release of every buffer to the last nmx_graph_destroy
  int live_graphs = 0;
  bool destroy_pending = false;
  std::vector<void*> retired;
  // orders a newly set stream after the work queued on the previous one
  hipEvent_t switch_ev = nullptr;
  // second stream (default priority) for work that overlaps the main
  // stream's within one call, and the two events that order them: created
  // on first use, joined back into `stream` before the call returns
  hipStream_t aux = nullptr;
  hipEvent_t ahead_ev = nullptr, behind_ev = nullptr;
  // this device's copy of the library's static scratch array (the device
  // address hipGetSymbolAddress returned at context creation) and its size:
  // the allocation-free entry points work in it
  void* static_scratch = nullptr;
  size_t static_scratch_bytes = 0;
  // cross-workgroup mailbox for the cooperative panel: flags then slots
  void* mailbox = nullptr;
  size_t mailbox_bytes = 0;
  uint32_t epoch = 0;  // last epoch handed to a kernel (monotonic)
  // device error word (sticky): NMX_ERR_SYNC_TIMEOUT from bounded waits
  int* err = nullptr;
  // 64 x 64 scratch for the inverse of a triangular base block (FP64-sized)
  void* tri_inv = nullptr;
  // buffers of the mixed-precision factorization, kept across calls
  void* mp_cache = nullptr;
  void (*mp_cache_free)(void*) = nullptr;
  // the stream is being captured into a graph (nmx_graph_begin .. _end):
  // nothing may synchronize, allocate or advance host-side epochs
  bool capturing = false;
  // event profiler (off unless nmx_ctx_set_profiling)
  bool profiling = false;
  std::vector<ProfRec> recs;
  std::vector<hipEvent_t> pool;
  int64_t prof_launches[K_COUNT] = {};
  double prof_ms[K_COUNT] = {}, prof_flops[K_COUNT] = {}, prof_bytes[K_COUNT] = {};
};

// Brackets one launch with events when the context is profiling.
struct ProfScope {
  nmx_ctx* ctx;
  ProfRec rec;
  bool on;
  ProfScope(nmx_ctx* c, int kind, double flops, double bytes);
  ~ProfScope();
};

// The context's second stream and its two events, created on first use.
int nmx_aux_stream(nmx_ctx* ctx);

// Points the context at another stream for a scope (launch helpers and
// ProfScope use ctx->stream), then back.
struct StreamSwap {
  nmx_ctx* ctx;
  hipStream_t saved;
  StreamSwap(nmx_ctx* c, hipStream_t s) : ctx(c), saved(c->stream) { c->stream = s; }
  ~StreamSwap() { ctx->stream = saved; }
};

// Workspace of at least `bytes`, 256-byte aligned; grows (with a stream sync)
// when needed.
int nmx_workspace(nmx_ctx* ctx, size_t bytes, void** out);

// Cross-workgroup mailbox of the cooperative panel (epoch-tagged 64-bit words).
int nmx_mailbox(nmx_ctx* ctx, void** out);
constexpr size_t kMailboxBytes = 64 * 1024;

// Advance and return the epoch base for `columns` upcoming columns; resets
// the flags (with a fill kernel) before the 32-bit counter could wrap.
int nmx_epoch_reserve(nmx_ctx* ctx, uint32_t columns, uint32_t* base);

// ---- templated building blocks (instantiated for float and double) --------
template <typename T>
int gemm_launch(nmx_ctx* ctx, int transa, int transb, int m, int n, int k, T alpha, const T* A, int lda, const T* B,
                int ldb, T beta, T* C, int ldc, int lower_only = 0);
// ksplit K slices of A B (no transposes) into ksplit outputs C + q csplit
template <typename T>
int gemm_launch_splitk(nmx_ctx* ctx, int m, int n, int k, T alpha, const T* A, int lda, const T* B, int ldb, T* C,
                       int ldc, int ksplit, int64_t csplit);

// Left-side triangular solve op(T) X = B (vectors are columns of B).
template <typename T>
int trsm_left(nmx_ctx* ctx, int upper, int unit, int trans, int n, int nrhs, const T* Tm, int ldt, T* B, int ldb);

// Right-side X := X L^-T with L lower n x n; X is m x n (the Cholesky panel).
template <typename T>
int trsm_right_lower_t(nmx_ctx* ctx, int m, int n, const T* L, int ldl, T* X, int ldx);

int trsm_rec_rows_upper(nmx_ctx* ctx, int m, int n, const float* U, int ldu, float* X, int ldx);

template <typename T>
int laswp_launch(nmx_ctx* ctx, T* A, int lda, int col_lo, int col_hi, int skip_lo, int skip_hi, const int32_t* ipiv,
                 int k1, int k2);

// perm[i] = original row that the interchanges ipiv[0..n) bring to row i;
// gather_rows applies it (out[i] = in[perm[i]], with a type conversion).
constexpr size_t kPermLdsBytes = 64 * 1024;
int perm_from_ipiv(nmx_ctx* ctx, int n, const int32_t* ipiv, int32_t* perm);
template <typename Ti, typename To>
int gather_rows(nmx_ctx* ctx, int n, int ncols, const int32_t* perm, const Ti* in, int ldi, To* out, int ldo);

template <typename T>
int getf2_leaf_launch(nmx_ctx* ctx, int n, T* A, int lda, int r0, int c0, int w, int32_t* ipiv, int32_t* info,
                      int swap_others);

template <typename T>
int getrf_impl(nmx_ctx* ctx, int n, T* A, int lda, int32_t* ipiv, int32_t* info, int block, int base);

// Block-level trailing update hook: A22 -= L21 U12 for the step at column k
// with panel width nb and trailing order m.
typedef int (*TrailingUpdate)(nmx_ctx* ctx, void* user, int k, int nb, int m);
// Panel hook (FP32 only): factors the panel columns k..k+nb over rows k..n,
// writes ipiv[k..k+nb) and applies the interchanges to every column of A,
// like panel_rec.
typedef int (*PanelHook)(nmx_ctx* ctx, void* user, int n, float* A, int lda, int k, int nb, int32_t* ipiv,
                         int32_t* info);
template <typename T>
int getrf_impl_ex(nmx_ctx* ctx, int n, T* A, int lda, int32_t* ipiv, int32_t* info, int block, int base,
                  TrailingUpdate update, void* user, PanelHook panel = nullptr);

// Tournament-pivoted panel of the mixed-precision factorization (tslu.hip):
// scratch of tslu_scratch_bytes(n), cleared once by tslu_scratch_init.
size_t tslu_scratch_bytes(int n);
int tslu_scratch_init(nmx_ctx* ctx, void* scratch);
// A panel's interchange records go to record slot `slot` (0 or 1); a slot
// keeps them until the panel after next.
int tslu_panel_launch(nmx_ctx* ctx, void* scratch, size_t scratch_bytes, int n, float* A, int lda, int k, int nb,
                      int32_t* ipiv, int32_t* info, int swap_outside, int slot = 0);
// The interchanges of the panel at k (records in `slot`) over columns
// c_lo .. c_hi outside it.
int tslu_swap_columns(nmx_ctx* ctx, void* scratch, float* A, int lda, int k, int nb, int c_lo, int c_hi,
                      int slot = 0);

template <typename T>
int potrf_impl(nmx_ctx* ctx, int n, T* A, int lda, int32_t* info, int block);

inline int ceil_div(int a, int b) { return (a + b - 1) / b; }

// ---- buffer loads: the hardware range check returns 0 past num_records, so
// bounds-checked tile loads need no branches ---------------------------------
constexpr uint32_t kBufOob = 0xFFFFFFF0u;  // an offset no descriptor covers

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, 0x7FFFFFFF, 0x00020000);
}
template <typename T>
__device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off);
template <>
__device__ __forceinline__ double buf_load<double>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
template <>
__device__ __forceinline__ float buf_load<float>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
// vector-memory stores through a descriptor: an out-of-range offset (kBufOob)
// is dropped by the range check, so bounds-checked tile stores need no branches
__device__ __forceinline__ void buf_store(__amdgpu_buffer_rsrc_t r, uint32_t off, double v) {
  typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, off, 0, 0);
}
__device__ __forceinline__ void buf_store(__amdgpu_buffer_rsrc_t r, uint32_t off, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, off, 0, 0);
}

// ---- wavefront reductions on DPP (no LDS traffic): every lane of a 16-lane
// row gets the row's extreme, then four lanes are read into scalar registers.
__device__ __forceinline__ uint32_t wave_umax(uint32_t v) {
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false));   // quad_perm 1,0,3,2
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false));   // quad_perm 2,3,0,1
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, false));  // row_half_mirror
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, false));  // row_mirror
  const uint32_t a = __builtin_amdgcn_readlane(v, 0), b = __builtin_amdgcn_readlane(v, 16);
  const uint32_t c = __builtin_amdgcn_readlane(v, 32), d = __builtin_amdgcn_readlane(v, 48);
  return max(max(a, b), max(c, d));
}
__device__ __forceinline__ uint32_t wave_umin(uint32_t v) {
  v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false));
  v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false));
  v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, false));
  v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, false));
  const uint32_t a = __builtin_amdgcn_readlane(v, 0), b = __builtin_amdgcn_readlane(v, 16);
  const uint32_t c = __builtin_amdgcn_readlane(v, 32), d = __builtin_amdgcn_readlane(v, 48);
  return min(min(a, b), min(c, d));
}

__device__ __forceinline__ float readlane_f(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
