// Blocked LU and Cholesky drivers and their C-ABI entry points.
#include <type_traits>

#include "nmx_internal.h"

template <typename T>
int diag_launch(nmx_ctx* ctx, int n, const T* A, int lda, T* d);
int zero_info(nmx_ctx* ctx, int32_t* info);
template <typename T>
int potrf_diag_launch(nmx_ctx* ctx, T* A, int lda, int k0, int nb, int32_t* info);
template <typename T>
int tril_launch(nmx_ctx* ctx, int n, T* A, int lda);

namespace {

#define AT(p, ld, i, j) ((p) + (int64_t)(i) * (ld) + (j))

// getrf2 (numax/linalg/panel.mojo:445-586): split the panel's columns until a
// half fits the cooperative leaf; the left half's effect on the right half is
// a triangular solve over the panel's own rows and an MFMA GEMM below them.
// After every leaf its interchanges are replayed over all other columns.
template <typename T>
int panel_rec(nmx_ctx* ctx, int n, T* A, int lda, int32_t* ipiv, int32_t* info, int k, int nb, int base) {
  if (nb <= base) return getf2_leaf_launch<T>(ctx, n, A, lda, k, k, nb, ipiv, info, 1);
  int half = nb / 2;
  if (half > base) half = (half / base) * base;
  NMX_TRY(panel_rec<T>(ctx, n, A, lda, ipiv, info, k, half, base));
  NMX_TRY(trsm_left<T>(ctx, 0, 1, 0, half, nb - half, AT(A, lda, k, k), lda, AT(A, lda, k, k + half), lda));
  const int rows = n - k - half, cols = nb - half;
  if (rows > 0 && cols > 0)
    NMX_TRY(gemm_launch<T>(ctx, 0, 0, rows, cols, half, T(-1), AT(A, lda, k + half, k), lda,
                           AT(A, lda, k, k + half), lda, T(1), AT(A, lda, k + half, k + half), lda, 0));
  return panel_rec<T>(ctx, n, A, lda, ipiv, info, k + half, nb - half, base);
}

// One panel: the recursive exact-pivoting panel, or (FP32 with a hook) the
// hook's panel.
template <typename T>
int panel(nmx_ctx* ctx, int n, T* A, int lda, int32_t* ipiv, int32_t* info, int k, int nb, int base, PanelHook hook,
          void* user) {
  if constexpr (std::is_same<T, float>::value) {
    if (hook) return hook(ctx, user, n, A, lda, k, nb, ipiv, info);
  }
  return panel_rec<T>(ctx, n, A, lda, ipiv, info, k, nb, base);
}

}  // namespace

// lu_factor (numax/linalg/lu.mojo:454-701): right-looking blocked GEPP.
// `update`, when given, replaces the FP64/FP32 GEMM of the block-level
// trailing update A22 -= L21 U12 (the mixed-precision factorization plugs
// its FP16 / BF16 MFMA update in here).
template <typename T>
int getrf_impl_ex(nmx_ctx* ctx, int n, T* A, int lda, int32_t* ipiv, int32_t* info, int block, int base,
                  TrailingUpdate update, void* user, PanelHook panel_hook) {
  if (block <= 0) block = 256;
  if (base <= 0 || base > kLeafW) base = kLeafW;
  NMX_TRY(zero_info(ctx, info));
  for (int k = 0; k < n; k += block) {
    const int nb = block < n - k ? block : n - k;
    NMX_TRY(panel<T>(ctx, n, A, lda, ipiv, info, k, nb, base, panel_hook, user));
    const int m = n - k - nb;
    if (m > 0) {
      // U12 := L11^-1 A12 over the whole block row, then A22 -= L21 U12.
      NMX_TRY(trsm_left<T>(ctx, 0, 1, 0, nb, m, AT(A, lda, k, k), lda, AT(A, lda, k, k + nb), lda));
      if (update) {
        NMX_TRY(update(ctx, user, k, nb, m));
      } else {
        NMX_TRY(gemm_launch<T>(ctx, 0, 0, m, m, nb, T(-1), AT(A, lda, k + nb, k), lda, AT(A, lda, k, k + nb), lda,
                               T(1), AT(A, lda, k + nb, k + nb), lda, 0));
      }
    }
  }
  return NMX_OK;
}

template <typename T>
int getrf_impl(nmx_ctx* ctx, int n, T* A, int lda, int32_t* ipiv, int32_t* info, int block, int base) {
  return getrf_impl_ex<T>(ctx, n, A, lda, ipiv, info, block, base, nullptr, nullptr);
}
template int getrf_impl_ex<float>(nmx_ctx*, int, float*, int, int32_t*, int32_t*, int, int, TrailingUpdate, void*,
                                   PanelHook);

// _potrf_blocked (numax/linalg/cholesky.mojo:59-230): diagonal block, panel
// solve X := X L11^-T, and the lower block triangle of A22 -= L21 L21^T.
// The diagonal block itself recurses the same way down to 64 columns.
namespace {
template <typename T>
int potrf_rec(nmx_ctx* ctx, T* A, int lda, int k, int nb, int32_t* info) {
  if (nb <= 64) return potrf_diag_launch<T>(ctx, A, lda, k, nb, info);
  int half = ((nb / 2 + 63) / 64) * 64;
  if (half >= nb) half = nb - 64;
  NMX_TRY(potrf_rec<T>(ctx, A, lda, k, half, info));
  const int rest = nb - half;
  NMX_TRY(trsm_right_lower_t<T>(ctx, rest, half, AT(A, lda, k, k), lda, AT(A, lda, k + half, k), lda));
  NMX_TRY(gemm_launch<T>(ctx, 0, 1, rest, rest, half, T(-1), AT(A, lda, k + half, k), lda, AT(A, lda, k + half, k),
                         lda, T(1), AT(A, lda, k + half, k + half), lda, 1));
  return potrf_rec<T>(ctx, A, lda, k + half, rest, info);
}
}  // namespace

template <typename T>
int potrf_impl(nmx_ctx* ctx, int n, T* A, int lda, int32_t* info, int block) {
  if (block <= 0) block = 256;
  NMX_TRY(zero_info(ctx, info));
  for (int k = 0; k < n; k += block) {
    const int nb = block < n - k ? block : n - k;
    NMX_TRY(potrf_rec<T>(ctx, A, lda, k, nb, info));
    const int m = n - k - nb;
    if (m > 0) {
      NMX_TRY(trsm_right_lower_t<T>(ctx, m, nb, AT(A, lda, k, k), lda, AT(A, lda, k + nb, k), lda));
      NMX_TRY(gemm_launch<T>(ctx, 0, 1, m, m, nb, T(-1), AT(A, lda, k + nb, k), lda, AT(A, lda, k + nb, k), lda,
                             T(1), AT(A, lda, k + nb, k + nb), lda, 1));
    }
  }
  return tril_launch<T>(ctx, n, A, lda);
}

template int getrf_impl<double>(nmx_ctx*, int, double*, int, int32_t*, int32_t*, int, int);
template int getrf_impl<float>(nmx_ctx*, int, float*, int, int32_t*, int32_t*, int, int);
template int potrf_impl<double>(nmx_ctx*, int, double*, int, int32_t*, int);
template int potrf_impl<float>(nmx_ctx*, int, float*, int, int32_t*, int);

// ---- C ABI -------------------------------------------------------------------
template <typename T>
static int getrf_api(nmx_ctx* ctx, int n, T* A, int lda, int32_t* ipiv, int32_t* info, int block, int base) {
  if (!ctx) return -1;
  if (n < 0) return -2;
  if (n > 0 && !A) return -3;
  if (lda < (n > 1 ? n : 1)) return -4;
  if (n > 0 && !ipiv) return -5;
  if (!info) return -6;
  if (block < 0) return -7;
  if (base < 0) return -8;
  if (n > kLeafMaxRows) return NMX_ERR_UNSUPPORTED;  // the pivoting leaf's row capacity
  if (n == 0) return zero_info(ctx, info);
  return getrf_impl<T>(ctx, n, A, lda, ipiv, info, block, base);
}

NMX_API int nmx_getrf_f64(nmx_ctx* ctx, int n, double* A, int lda, int32_t* ipiv, int32_t* info, int block,
                          int base) {
  return getrf_api<double>(ctx, n, A, lda, ipiv, info, block, base);
}
NMX_API int nmx_getrf_f32(nmx_ctx* ctx, int n, float* A, int lda, int32_t* ipiv, int32_t* info, int block, int base) {
  return getrf_api<float>(ctx, n, A, lda, ipiv, info, block, base);
}

// laswp over all n pivots (panel.mojo:1824-1852): the net permutation is
// formed once and applied as one gather through the workspace.
template <typename T>
static int laswp_full(nmx_ctx* ctx, int n, int nrhs, T* B, int ldb, const int32_t* ipiv) {
  void* ws = nullptr;
  const size_t vec_bytes = ((sizeof(T) * (size_t)n * nrhs + 255) / 256) * 256;
  NMX_TRY(nmx_workspace(ctx, vec_bytes + sizeof(int32_t) * (size_t)n, &ws));
  T* W = (T*)ws;
  int32_t* perm = (int32_t*)((char*)ws + vec_bytes);
  NMX_TRY(perm_from_ipiv(ctx, n, ipiv, perm));
  NMX_TRY((gather_rows<T, T>(ctx, n, nrhs, perm, B, ldb, W, nrhs)));
  NMX_HIP(ctx, hipMemcpy2DAsync(B, sizeof(T) * (size_t)ldb, W, sizeof(T) * (size_t)nrhs, sizeof(T) * (size_t)nrhs, n,
                                hipMemcpyDeviceToDevice, ctx->stream));
  return NMX_OK;
}

// LU.solve (lu.mojo:181-354): laswp, forward (unit) then back substitution.
template <typename T>
static int getrs_api(nmx_ctx* ctx, int n, int nrhs, const T* LU, int lda, const int32_t* ipiv, T* B, int ldb) {
  if (!ctx) return -1;
  if (n < 0) return -2;
  if (nrhs < 0) return -3;
  if (n > 0 && !LU) return -4;
  if (lda < (n > 1 ? n : 1)) return -5;
  if (n > 0 && !ipiv) return -6;
  if (n > 0 && nrhs > 0 && !B) return -7;
  if (ldb < (nrhs > 1 ? nrhs : 1)) return -8;
  if (n == 0 || nrhs == 0) return NMX_OK;
  NMX_TRY(laswp_full<T>(ctx, n, nrhs, B, ldb, ipiv));
  NMX_TRY(trsm_left<T>(ctx, 0, 1, 0, n, nrhs, LU, lda, B, ldb));
  return trsm_left<T>(ctx, 1, 0, 0, n, nrhs, LU, lda, B, ldb);
}

NMX_API int nmx_getrs_f64(nmx_ctx* ctx, int n, int nrhs, const double* LU, int lda, const int32_t* ipiv, double* B,
                          int ldb) {
  return getrs_api<double>(ctx, n, nrhs, LU, lda, ipiv, B, ldb);
}
NMX_API int nmx_getrs_f32(nmx_ctx* ctx, int n, int nrhs, const float* LU, int lda, const int32_t* ipiv, float* B,
                          int ldb) {
  return getrs_api<float>(ctx, n, nrhs, LU, lda, ipiv, B, ldb);
}

template <typename T>
__global__ void eye_kernel(int n, T* A, int lda) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * n) return;
  const int i = (int)(e / n), j = (int)(e % n);
  A[(int64_t)i * lda + j] = i == j ? T(1) : T(0);
}

// inverse (basic.mojo:149-195) after lu_factor: LU.solve against the whole
// identity, n right-hand sides, so the substitutions are blocked TRSMs with
// MFMA updates between the diagonal blocks.
template <typename T>
static int getri_api(nmx_ctx* ctx, int n, const T* LU, int lda, const int32_t* ipiv, T* Ainv, int ldi) {
  if (!ctx) return -1;
  if (n < 0) return -2;
  if (lda < (n > 1 ? n : 1)) return -4;
  if (ldi < (n > 1 ? n : 1)) return -7;
  if (n == 0) return NMX_OK;
  if (!LU) return -3;
  if (!ipiv) return -5;
  if (!Ainv) return -6;
  eye_kernel<T><<<(unsigned)(((int64_t)n * n + 255) / 256), 256, 0, ctx->stream>>>(n, Ainv, ldi);
  NMX_HIP(ctx, hipGetLastError());
  return getrs_api<T>(ctx, n, n, LU, lda, ipiv, Ainv, ldi);
}
NMX_API int nmx_getri_f64(nmx_ctx* ctx, int n, const double* LU, int lda, const int32_t* ipiv, double* Ainv,
                          int ldi) {
  return getri_api<double>(ctx, n, LU, lda, ipiv, Ainv, ldi);
}
NMX_API int nmx_getri_f32(nmx_ctx* ctx, int n, const float* LU, int lda, const int32_t* ipiv, float* Ainv,
                          int ldi) {
  return getri_api<float>(ctx, n, LU, lda, ipiv, Ainv, ldi);
}

// A = P L U from lu_factor's packed factor (lu, lu.mojo:849-975): L (unit
// diagonal) and U split out, perm[i] = the row of A that row i of L U came
// from (the interchanges replayed in order on the identity), P[perm[c]][c]
// = 1.
template <typename T>
__global__ void lu_unpack_kernel(int n, const T* __restrict__ LU, int lda, const int32_t* __restrict__ perm,
                                 T* __restrict__ P, int ldp, T* __restrict__ L, int ldl, T* __restrict__ U, int ldu) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * n) return;
  const int i = (int)(e / n), j = (int)(e % n);
  const T v = LU[(int64_t)i * lda + j];
  L[(int64_t)i * ldl + j] = i > j ? v : (i == j ? T(1) : T(0));
  U[(int64_t)i * ldu + j] = i <= j ? v : T(0);
  P[(int64_t)i * ldp + j] = perm[j] == i ? T(1) : T(0);
}
template <typename T>
static int lu_unpack_api(nmx_ctx* ctx, int n, const T* LU, int lda, const int32_t* ipiv, int32_t* perm, T* P,
                         int ldp, T* L, int ldl, T* U, int ldu) {
  if (!ctx) return -1;
  if (n < 0) return -2;
  const int ld1 = n > 1 ? n : 1;
  if (lda < ld1) return -4;
  if (ldp < ld1) return -8;
  if (ldl < ld1) return -10;
  if (ldu < ld1) return -12;
  if (n == 0) return NMX_OK;
  if (!LU) return -3;
  if (!ipiv) return -5;
  if (!perm) return -6;
  if (!P) return -7;
  if (!L) return -9;
  if (!U) return -11;
  NMX_TRY(perm_from_ipiv(ctx, n, ipiv, perm));
  lu_unpack_kernel<T><<<(unsigned)(((int64_t)n * n + 255) / 256), 256, 0, ctx->stream>>>(n, LU, lda, perm, P, ldp,
                                                                                          L, ldl, U, ldu);
  NMX_HIP(ctx, hipGetLastError());
  return NMX_OK;
}
NMX_API int nmx_lu_unpack_f64(nmx_ctx* ctx, int n, const double* LU, int lda, const int32_t* ipiv, int32_t* perm,
                              double* P, int ldp, double* L, int ldl, double* U, int ldu) {
  return lu_unpack_api<double>(ctx, n, LU, lda, ipiv, perm, P, ldp, L, ldl, U, ldu);
}
NMX_API int nmx_lu_unpack_f32(nmx_ctx* ctx, int n, const float* LU, int lda, const int32_t* ipiv, int32_t* perm,
                              float* P, int ldp, float* L, int ldl, float* U, int ldu) {
  return lu_unpack_api<float>(ctx, n, LU, lda, ipiv, perm, P, ldp, L, ldl, U, ldu);
}

template <typename T>
static int laswp_api(nmx_ctx* ctx, int n, int nrhs, T* B, int ldb, const int32_t* ipiv) {
  if (!ctx) return -1;
  if (n < 0) return -2;
  if (nrhs < 0) return -3;
  if (ldb < (nrhs > 1 ? nrhs : 1)) return -5;
  if (n == 0 || nrhs == 0) return NMX_OK;
  if (!B) return -4;
  if (!ipiv) return -6;
  return laswp_full<T>(ctx, n, nrhs, B, ldb, ipiv);
}
NMX_API int nmx_laswp_f64(nmx_ctx* ctx, int n, int nrhs, double* B, int ldb, const int32_t* ipiv) {
  return laswp_api<double>(ctx, n, nrhs, B, ldb, ipiv);
}
NMX_API int nmx_laswp_f32(nmx_ctx* ctx, int n, int nrhs, float* B, int ldb, const int32_t* ipiv) {
  return laswp_api<float>(ctx, n, nrhs, B, ldb, ipiv);
}

template <typename T>
static int diag_api(nmx_ctx* ctx, int n, const T* A, int lda, T* d) {
  if (!ctx) return -1;
  if (n < 0) return -2;
  if (lda < (n > 1 ? n : 1)) return -4;
  if (n == 0) return NMX_OK;
  if (!A) return -3;
  if (!d) return -5;
  return diag_launch<T>(ctx, n, A, lda, d);
}
NMX_API int nmx_diag_f64(nmx_ctx* ctx, int n, const double* A, int lda, double* diag) {
  return diag_api<double>(ctx, n, A, lda, diag);
}
NMX_API int nmx_diag_f32(nmx_ctx* ctx, int n, const float* A, int lda, float* diag) {
  return diag_api<float>(ctx, n, A, lda, diag);
}

template <typename T>
static int trsm_api(nmx_ctx* ctx, int upper, int unit, int trans, int n, int nrhs, const T* Tm, int ldt, T* B,
                    int ldb) {
  if (!ctx) return -1;
  if (n < 0) return -5;
  if (nrhs < 0) return -6;
  if (ldt < (n > 1 ? n : 1)) return -8;
  if (ldb < (nrhs > 1 ? nrhs : 1)) return -10;
  if (n == 0 || nrhs == 0) return NMX_OK;
  if (!Tm) return -7;
  if (!B) return -9;
  return trsm_left<T>(ctx, upper != 0, unit != 0, trans != 0, n, nrhs, Tm, ldt, B, ldb);
}
NMX_API int nmx_trsm_f64(nmx_ctx* ctx, int upper, int unit, int trans, int n, int nrhs, const double* T, int ldt,
                         double* B, int ldb) {
  return trsm_api<double>(ctx, upper, unit, trans, n, nrhs, T, ldt, B, ldb);
}
NMX_API int nmx_trsm_f32(nmx_ctx* ctx, int upper, int unit, int trans, int n, int nrhs, const float* T, int ldt,
                         float* B, int ldb) {
  return trsm_api<float>(ctx, upper, unit, trans, n, nrhs, T, ldt, B, ldb);
}

template <typename T>
static int potrf_api(nmx_ctx* ctx, int n, T* A, int lda, int32_t* info, int block) {
  if (!ctx) return -1;
  if (n < 0) return -2;
  if (n > 0 && !A) return -3;
  if (lda < (n > 1 ? n : 1)) return -4;
  if (!info) return -5;
  if (block < 0) return -6;
  if (n == 0) return zero_info(ctx, info);
  return potrf_impl<T>(ctx, n, A, lda, info, block);
}
NMX_API int nmx_potrf_f64(nmx_ctx* ctx, int n, double* A, int lda, int32_t* info, int block) {
  return potrf_api<double>(ctx, n, A, lda, info, block);
}
NMX_API int nmx_potrf_f32(nmx_ctx* ctx, int n, float* A, int lda, int32_t* info, int block) {
  return potrf_api<float>(ctx, n, A, lda, info, block);
}

// cholesky_solve (cholesky.mojo:471-580): L y = b, then L^T x = y reading L
// transposed in place.
template <typename T>
static int potrs_api(nmx_ctx* ctx, int n, int nrhs, const T* L, int lda, T* B, int ldb) {
  if (!ctx) return -1;
  if (n < 0) return -2;
  if (nrhs < 0) return -3;
  if (lda < (n > 1 ? n : 1)) return -5;
  if (ldb < (nrhs > 1 ? nrhs : 1)) return -7;
  if (n == 0 || nrhs == 0) return NMX_OK;
  if (!L) return -4;
  if (!B) return -6;
  NMX_TRY(trsm_left<T>(ctx, 0, 0, 0, n, nrhs, L, lda, B, ldb));
  return trsm_left<T>(ctx, 1, 0, 1, n, nrhs, L, lda, B, ldb);
}
NMX_API int nmx_potrs_f64(nmx_ctx* ctx, int n, int nrhs, const double* L, int lda, double* B, int ldb) {
  return potrs_api<double>(ctx, n, nrhs, L, lda, B, ldb);
}
NMX_API int nmx_potrs_f32(nmx_ctx* ctx, int n, int nrhs, const float* L, int lda, float* B, int ldb) {
  return potrs_api<float>(ctx, n, nrhs, L, lda, B, ldb);
}
