// Symmetric indefinite factorization A = L D L^T with Bunch-Kaufman
// pivoting (numax's `ldl`, numax/linalg/ldl.mojo:59-148): LAPACK's sytf2 on
// the lower triangle (numax/linalg/panel.mojo:2732-2894), then SciPy's
// unpacking into (lu, d, perm) (panel.mojo:2897-3003).
//
// One workgroup of 1024 threads does the whole factorization, as numax's
// single-block kernel does: every step is a pivot decision the next step
// depends on.  The pivot scans here are workgroup reductions (each thread
// scans its rows in order, keeping the first maximum, then the threads'
// pairs are combined by larger value, then lower row) instead of numax's
// serial scan on thread 0, and the trailing update runs over the whole
// lower triangle at once, one element per thread-step in row-major order
// (neighbouring threads, neighbouring columns).  Every product and
// difference is rounded as numax's loop rounds it (no fused multiply-add),
// so the factor equals the CPU restatement of sytf2 bit for bit.
#include <math.h>

#include "nmx_internal.h"

#pragma clang fp contract(off)

namespace {

constexpr int kLdlThreads = 1024;
constexpr double kBunchKaufmanAlpha = 0.6403882032022076;  // (1 + sqrt(17)) / 8

#define A_(i, j) A[(int64_t)(i) * lda + (j)]

// (value, index) of the first largest |a| over this thread's strided rows,
// combined over the workgroup: larger value wins, ties to the lower index;
// values that are not > 0 (zeros, NaN) never win (sequential `v > best`
// from best = 0).
template <typename T>
__device__ void block_argmax(T v, int idx, T* sv, int* si, T& best, int& bidx) {
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  if (!(v > T(0))) {
    v = T(0);
    idx = INT32_MAX;
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    const T ov = __shfl_xor(v, m);
    const int oi = __shfl_xor(idx, m);
    if (ov > v || (ov == v && oi < idx)) {
      v = ov;
      idx = oi;
    }
  }
  if (lane == 0) {
    sv[wv] = v;
    si[wv] = idx;
  }
  __syncthreads();
  if (t == 0) {
    T b = T(0);
    int bi = INT32_MAX;
    for (int w = 0; w < kLdlThreads / 64; w++) {
      if (sv[w] > b || (sv[w] == b && si[w] < bi)) {
        b = sv[w];
        bi = si[w];
      }
    }
    sv[0] = b;
    si[0] = bi;
  }
  __syncthreads();
  best = sv[0];
  bidx = si[0];
  __syncthreads();
}

// sytf2, uplo = 'L', in place; ipiv 1-based: kp + 1 for a 1 x 1 block
// interchanged with row kp, -(kp + 1) on both rows of a 2 x 2 block.
// info (device) = first k + 1 whose pivot block is exactly zero.
template <typename T>
__global__ void __launch_bounds__(kLdlThreads) sytf2_lower_kernel(int n, T* __restrict__ A, int lda,
                                                                  int32_t* __restrict__ ipiv,
                                                                  int32_t* __restrict__ info) {
  __shared__ T sv[kLdlThreads / 64];
  __shared__ int si[kLdlThreads / 64];
  const int t = threadIdx.x, nt = kLdlThreads;
  const T alpha = T(kBunchKaufmanAlpha);
  if (t == 0) *info = 0;
  int k = 0;
  while (k < n) {
    // ---- pivot choice (identical on every thread after the reductions)
    const T absakk = fabs(A_(k, k));
    T cv = T(0);
    int ci = k;
    for (int i = k + 1 + t; i < n; i += nt) {
      const T v = fabs(A_(i, k));
      if (v > cv) {
        cv = v;
        ci = i;
      }
    }
    T colmax;
    int imax;
    block_argmax(cv, ci, sv, si, colmax, imax);
    if (imax == INT32_MAX) imax = k;
    int kstep = 1, kp = k;
    if (absakk == T(0) && colmax == T(0)) {
      kp = k;
      if (t == 0 && *info == 0) *info = k + 1;
    } else if (absakk >= alpha * colmax) {
      kp = k;
    } else {
      T rv = T(0);
      for (int j = k + t; j < imax; j += nt) rv = fmax(rv, fabs(A_(imax, j)));
      for (int i = imax + 1 + t; i < n; i += nt) rv = fmax(rv, fabs(A_(i, imax)));
      T rowmax;
      int unused;
      block_argmax(rv, 0, sv, si, rowmax, unused);
      if (absakk >= alpha * colmax * (colmax / rowmax)) {
        kp = k;
      } else if (fabs(A_(imax, imax)) >= alpha * rowmax) {
        kp = imax;
      } else if (k + 1 < n) {
        kp = imax;
        kstep = 2;
      } else {
        // the last column (imax == k): only NaN comparisons after a zero
        // pivot get here, and a 2 x 2 block would start past the matrix
        kp = k;
      }
    }
    __syncthreads();  // every thread has read what the decision needs before anything moves
    if (t == 0) {
      if (kstep == 1) {
        ipiv[k] = kp + 1;
      } else {
        ipiv[k] = -(kp + 1);
        ipiv[k + 1] = -(kp + 1);
      }
    }
    // ---- interchange rows and columns kk and kp of the trailing triangle
    const int kk = k + kstep - 1;
    if (kp != kk) {
      for (int i = kp + 1 + t; i < n; i += nt) {
        const T h = A_(i, kk);
        A_(i, kk) = A_(i, kp);
        A_(i, kp) = h;
      }
      for (int j = kk + 1 + t; j < kp; j += nt) {
        const T h = A_(j, kk);
        A_(j, kk) = A_(kp, j);
        A_(kp, j) = h;
      }
      if (t == 0) {
        const T h = A_(kk, kk);
        A_(kk, kk) = A_(kp, kp);
        A_(kp, kp) = h;
        if (kstep == 2) {
          const T h2 = A_(k + 1, k);
          A_(k + 1, k) = A_(kp, k);
          A_(kp, k) = h2;
        }
      }
    }
    __syncthreads();
    // ---- update of the trailing lower triangle, then the pivot columns
    if (kstep == 1) {
      if (k < n - 1) {
        const T d11 = T(1) / A_(k, k);
        const int m = n - k - 1;
        for (int64_t e = t; e < (int64_t)m * m; e += nt) {
          const int i = k + 1 + (int)(e / m), j = k + 1 + (int)(e % m);
          if (i >= j) A_(i, j) = A_(i, j) - d11 * A_(i, k) * A_(j, k);
        }
        __syncthreads();
        for (int i = k + 1 + t; i < n; i += nt) A_(i, k) = A_(i, k) * d11;
      }
    } else if (k < n - 2) {
      T d21 = A_(k + 1, k);
      const T d11 = A_(k + 1, k + 1) / d21;
      const T d22 = A_(k, k) / d21;
      const T tt = T(1) / (d11 * d22 - T(1));
      d21 = tt / d21;
      const int m = n - k - 2;
      for (int64_t e = t; e < (int64_t)m * m; e += nt) {
        const int i = k + 2 + (int)(e / m), j = k + 2 + (int)(e % m);
        if (i >= j) {
          const T wk = d21 * (d11 * A_(j, k) - A_(j, k + 1));
          const T wkp1 = d21 * (d22 * A_(j, k + 1) - A_(j, k));
          A_(i, j) = A_(i, j) - A_(i, k) * wk - A_(i, k + 1) * wkp1;
        }
      }
      __syncthreads();
      for (int j = k + 2 + t; j < n; j += nt) {
        const T wk = d21 * (d11 * A_(j, k) - A_(j, k + 1));
        const T wkp1 = d21 * (d22 * A_(j, k + 1) - A_(j, k));
        A_(j, k) = wk;
        A_(j, k + 1) = wkp1;
      }
    }
    __syncthreads();
    k += kstep;
  }
}

// SciPy's post-processing (_ldl_sanitize_ipiv, _ldl_get_d_and_l,
// _ldl_construct_tri_factor) of sytf2's packed factor: lu, d (n x n), perm.
// scratch: 3 n int32 (swap, block sizes, the permutation before the final
// inversion).
template <typename T>
__global__ void __launch_bounds__(kLdlThreads) ldl_unpack_kernel(int n, const T* __restrict__ A, int lda,
                                                                 const int32_t* __restrict__ ipiv, T* __restrict__ L,
                                                                 int ldl, T* __restrict__ D, int ldd,
                                                                 int64_t* __restrict__ perm,
                                                                 int32_t* __restrict__ scratch) {
  const int t = threadIdx.x, nt = kLdlThreads;
  int32_t* swap = scratch;
  int32_t* piv = scratch + n;
  int32_t* pr = scratch + 2 * n;
  if (t == 0) {
    for (int i = 0; i < n; i++) {
      swap[i] = i;
      piv[i] = 0;
      pr[i] = i;
    }
    int ind = 0;
    while (ind < n) {
      const int cur = ipiv[ind];
      if (cur > 0) {
        if (cur != ind + 1) swap[ind] = swap[cur - 1];
        piv[ind] = 1;
        ind += 1;
      } else {
        if (-cur != ind + 2) swap[ind + 1] = swap[-cur - 1];
        piv[ind] = 2;
        ind += 2;
      }
    }
  }
  __syncthreads();
  for (int64_t e = t; e < (int64_t)n * n; e += nt) {
    const int i = (int)(e / n), j = (int)(e % n);
    D[(int64_t)i * ldd + j] = i == j ? A_(i, i) : T(0);
    L[(int64_t)i * ldl + j] = i == j ? T(1) : (i > j ? A_(i, j) : T(0));
  }
  __syncthreads();
  if (t == 0) {
    int blk = 0;
    while (blk < n) {
      if (piv[blk] == 2) {
        const T v = A_(blk + 1, blk);
        D[(int64_t)(blk + 1) * ldd + blk] = v;
        D[(int64_t)blk * ldd + blk + 1] = v;
        L[(int64_t)(blk + 1) * ldl + blk] = T(0);
        blk += 2;
      } else {
        blk += 1;
      }
    }
  }
  __syncthreads();
  for (int ind = n - 1; ind >= 0; ind--) {
    const int s = swap[ind];
    if (s != ind) {
      const int col_s = piv[ind] == 0 ? ind - 1 : ind;
      for (int c = col_s + t; c < n; c += nt) {
        const T h = L[(int64_t)s * ldl + c];
        L[(int64_t)s * ldl + c] = L[(int64_t)ind * ldl + c];
        L[(int64_t)ind * ldl + c] = h;
      }
      if (t == 0) {
        const int32_t ps = pr[s];
        pr[s] = pr[ind];
        pr[ind] = ps;
      }
    }
    __syncthreads();
  }
  for (int i = t; i < n; i += nt) perm[pr[i]] = i;
}

template <typename T>
int sytrf_api(nmx_ctx* ctx, int n, T* A, int lda, int32_t* ipiv, int32_t* info) {
  if (!ctx) return -1;
  if (n < 0) return -2;
  if (lda < (n > 1 ? n : 1)) return -4;
  if (!info) return -6;
  if (n > 0 && !A) return -3;
  if (n > 0 && !ipiv) return -5;
  ProfScope ps(ctx, K_OTHER, (double)n * n * n / 3.0, 2.0 * sizeof(T) * (double)n * n);
  sytf2_lower_kernel<T><<<1, kLdlThreads, 0, ctx->stream>>>(n, A, lda, ipiv, info);
  NMX_HIP(ctx, hipGetLastError());
  return NMX_OK;
}

template <typename T>
int ldl_unpack_api(nmx_ctx* ctx, int n, const T* A, int lda, const int32_t* ipiv, T* L, int ldl, T* D, int ldd,
                   int64_t* perm) {
  if (!ctx) return -1;
  if (n < 0) return -2;
  const int ld1 = n > 1 ? n : 1;
  if (lda < ld1) return -4;
  if (ldl < ld1) return -7;
  if (ldd < ld1) return -9;
  if (n == 0) return NMX_OK;
  if (!A) return -3;
  if (!ipiv) return -5;
  if (!L) return -6;
  if (!D) return -8;
  if (!perm) return -10;
  void* ws = nullptr;
  NMX_TRY(nmx_workspace(ctx, sizeof(int32_t) * 3 * (size_t)n, &ws));
  ldl_unpack_kernel<T><<<1, kLdlThreads, 0, ctx->stream>>>(n, A, lda, ipiv, L, ldl, D, ldd, perm, (int32_t*)ws);
  NMX_HIP(ctx, hipGetLastError());
  return NMX_OK;
}

#undef A_

}  // namespace

NMX_API int nmx_sytrf_f64(nmx_ctx* ctx, int n, double* A, int lda, int32_t* ipiv, int32_t* info) {
  return sytrf_api<double>(ctx, n, A, lda, ipiv, info);
}
NMX_API int nmx_sytrf_f32(nmx_ctx* ctx, int n, float* A, int lda, int32_t* ipiv, int32_t* info) {
  return sytrf_api<float>(ctx, n, A, lda, ipiv, info);
}
NMX_API int nmx_ldl_unpack_f64(nmx_ctx* ctx, int n, const double* LDU, int lda, const int32_t* ipiv, double* L,
                               int ldl, double* D, int ldd, int64_t* perm) {
  return ldl_unpack_api<double>(ctx, n, LDU, lda, ipiv, L, ldl, D, ldd, perm);
}
NMX_API int nmx_ldl_unpack_f32(nmx_ctx* ctx, int n, const float* LDU, int lda, const int32_t* ipiv, float* L, int ldl,
                               float* D, int ldd, int64_t* perm) {
  return ldl_unpack_api<float>(ctx, n, LDU, lda, ipiv, L, ldl, D, ldd, perm);
}
