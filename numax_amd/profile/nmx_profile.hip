// Standalone driver for hardware-counter and kernel-trace runs of the
// library's hot paths, with no framework in the process: every kernel a
// profiler sees is either the library's or the calibration copy below.
//
//   nmx_profile copy  BYTES      plain float4 copy of BYTES (HBM calibration)
//   nmx_profile lu    N [NRHS]   FP64 getrf + getrs, A ~ N(0,1), NRHS (1) right-hand sides
//   nmx_profile chol  N [NRHS]   FP64 potrf + potrs, A = G G^T / n + I, NRHS (1) right-hand sides
//   nmx_profile mpir  N [PREC]   mixed-precision LU-IR, A ~ N(0,1) + 0 I
//   nmx_profile mprhs N [NRHS]   the same for NRHS (64) right-hand sides
//   nmx_profile batch COUNT      uneven batch, n ~ U{8..256}, half LU / half
//                                Cholesky, FP32 factor + FP64 refinement
//   nmx_profile eigh  COUNT      batched Jacobi eigh, n in {24, 32, 48, 64} in equal
//                                shares, (G + G^T) / 2, FP32, 12 sweeps, vectors
//   nmx_profile lstsq COUNT      batched least squares, A 1024 x 32 ~ N(0,1), FP32
//
// Each mode runs its path twice (a warm-up, then the measured call) and
// prints one line per step.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <random>
#include <vector>

#include "numax_hip.h"

#define CHECK(x)                                                                    \
  do {                                                                              \
    int rc_ = (x);                                                                  \
    if (rc_ != 0) {                                                                 \
      fprintf(stderr, "%s failed: %d (%s)\n", #x, rc_, nmx_status_string(rc_));    \
      exit(1);                                                                      \
    }                                                                               \
  } while (0)
#define HCHECK(x)                                                                   \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) {                                                         \
      fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));                \
      exit(1);                                                                      \
    }                                                                               \
  } while (0)

__global__ void __launch_bounds__(256) copy_kernel(const float4* __restrict__ src, float4* __restrict__ dst,
                                                   long long n) {
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (long long)gridDim.x * 256) dst[i] = src[i];
}

static int run_copy(long long bytes) {
  const long long n = bytes / 16;
  float4 *a, *b;
  HCHECK(hipMalloc(&a, n * 16));
  HCHECK(hipMalloc(&b, n * 16));
  HCHECK(hipMemset(a, 0, n * 16));
  for (int it = 0; it < 2; it++) {
    copy_kernel<<<4096, 256>>>(a, b, n);
    HCHECK(hipGetLastError());
    HCHECK(hipDeviceSynchronize());
    printf("copy step %d: %lld bytes read, %lld written\n", it, n * 16, n * 16);
  }
  HCHECK(hipFree(a));
  HCHECK(hipFree(b));
  return 0;
}

static int run_lu(nmx_ctx* ctx, int n, int nrhs) {
  double *A, *LU, *b;
  int32_t *ipiv, *info;
  HCHECK(hipMalloc(&A, sizeof(double) * n * n));
  HCHECK(hipMalloc(&LU, sizeof(double) * n * n));
  HCHECK(hipMalloc(&b, sizeof(double) * n * nrhs));
  HCHECK(hipMalloc(&ipiv, sizeof(int32_t) * n));
  HCHECK(hipMalloc(&info, sizeof(int32_t)));
  CHECK(nmx_fill_random_f64(ctx, 1, 7, 0, (int64_t)n * n, 0, 0, A));
  for (int it = 0; it < 2; it++) {
    CHECK(nmx_fill_random_f64(ctx, 1, 8, 0, (int64_t)n * nrhs, 0, 0, b));
    HCHECK(hipMemcpy(LU, A, sizeof(double) * n * n, hipMemcpyDeviceToDevice));
    CHECK(nmx_getrf_f64(ctx, n, LU, n, ipiv, info, 256, 0));
    CHECK(nmx_getrs_f64(ctx, n, nrhs, LU, n, ipiv, b, nrhs));
    CHECK(nmx_sync(ctx));
    printf("lu step %d: n=%d nrhs=%d done\n", it, n, nrhs);
    fflush(stdout);
  }
  return 0;
}

__global__ void add_identity_kernel(int n, double* A, double scale) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) A[(int64_t)i * n + i] += scale;
}

static int run_chol(nmx_ctx* ctx, int n, int nrhs) {
  double *G, *A, *L, *b;
  int32_t* info;
  HCHECK(hipMalloc(&G, sizeof(double) * n * n));
  HCHECK(hipMalloc(&A, sizeof(double) * n * n));
  HCHECK(hipMalloc(&L, sizeof(double) * n * n));
  HCHECK(hipMalloc(&b, sizeof(double) * n * nrhs));
  HCHECK(hipMalloc(&info, sizeof(int32_t)));
  CHECK(nmx_fill_random_f64(ctx, 1, 9, 0, (int64_t)n * n, 0, 0, G));
  CHECK(nmx_gemm_f64(ctx, 0, 1, n, n, n, 1.0 / n, G, n, G, n, 0.0, A, n));
  add_identity_kernel<<<(n + 255) / 256, 256>>>(n, A, 1.0);
  HCHECK(hipGetLastError());
  HCHECK(hipDeviceSynchronize());
  for (int it = 0; it < 2; it++) {
    CHECK(nmx_fill_random_f64(ctx, 1, 8, 0, (int64_t)n * nrhs, 0, 0, b));
    HCHECK(hipMemcpy(L, A, sizeof(double) * n * n, hipMemcpyDeviceToDevice));
    CHECK(nmx_potrf_f64(ctx, n, L, n, info, 256));
    CHECK(nmx_potrs_f64(ctx, n, nrhs, L, n, b, nrhs));
    CHECK(nmx_sync(ctx));
    int h = 0;
    HCHECK(hipMemcpy(&h, info, 4, hipMemcpyDeviceToHost));
    printf("chol step %d: n=%d nrhs=%d info %d\n", it, n, nrhs, h);
    fflush(stdout);
  }
  return 0;
}

static int run_eigh(nmx_ctx* ctx, int count) {
  std::mt19937_64 rng(13);
  std::normal_distribution<float> nd(0.0f, 1.0f);
  const int sizes[4] = {24, 32, 48, 64};
  std::vector<int32_t> n(count);
  std::vector<int64_t> oa(count), ow(count);
  int64_t ae = 0, we = 0;
  for (int p = 0; p < count; p++) {
    n[p] = sizes[p % 4];
    oa[p] = ae;
    ow[p] = we;
    ae += (int64_t)n[p] * n[p];
    we += n[p];
  }
  std::vector<float> a(ae);
  for (int p = 0; p < count; p++) {
    const int m = n[p];
    float* ap = a.data() + oa[p];
    for (int i = 0; i < m; i++)
      for (int j = 0; j <= i; j++) ap[i * m + j] = ap[j * m + i] = nd(rng);
  }
  int32_t* dn;
  int64_t *doa, *dow;
  float *dA, *dW, *dV;
  HCHECK(hipMalloc(&dn, 4 * count));
  HCHECK(hipMalloc(&doa, 8 * count));
  HCHECK(hipMalloc(&dow, 8 * count));
  HCHECK(hipMalloc(&dA, 4 * ae));
  HCHECK(hipMalloc(&dW, 4 * we));
  HCHECK(hipMalloc(&dV, 4 * ae));
  HCHECK(hipMemcpy(dn, n.data(), 4 * count, hipMemcpyHostToDevice));
  HCHECK(hipMemcpy(doa, oa.data(), 8 * count, hipMemcpyHostToDevice));
  HCHECK(hipMemcpy(dow, ow.data(), 8 * count, hipMemcpyHostToDevice));
  HCHECK(hipMemcpy(dA, a.data(), 4 * ae, hipMemcpyHostToDevice));
  for (int it = 0; it < 2; it++) {
    CHECK(nmx_batched_syevj_f32(ctx, count, dn, doa, dow, dA, ae, dW, we, dV, 12));
    CHECK(nmx_sync(ctx));
    printf("eigh step %d: %d matrices\n", it, count);
    fflush(stdout);
  }
  return 0;
}

static int run_lstsq(nmx_ctx* ctx, int count) {
  const int m = 1024, n = 32;
  float *A, *B, *X;
  HCHECK(hipMalloc(&A, sizeof(float) * (size_t)count * m * n));
  HCHECK(hipMalloc(&B, sizeof(float) * (size_t)count * m));
  HCHECK(hipMalloc(&X, sizeof(float) * (size_t)count * n));
  std::mt19937_64 rng(17);
  std::normal_distribution<float> nd(0.0f, 1.0f);
  std::vector<float> a((size_t)count * m * n), b((size_t)count * m);
  for (auto& v : a) v = nd(rng);
  for (auto& v : b) v = nd(rng);
  HCHECK(hipMemcpy(A, a.data(), 4 * a.size(), hipMemcpyHostToDevice));
  HCHECK(hipMemcpy(B, b.data(), 4 * b.size(), hipMemcpyHostToDevice));
  for (int it = 0; it < 2; it++) {
    CHECK(nmx_batched_gels_f32(ctx, count, m, n, A, B, X));
    CHECK(nmx_sync(ctx));
    printf("lstsq step %d: %d problems %d x %d\n", it, count, m, n);
    fflush(stdout);
  }
  return 0;
}

static int run_mpir(nmx_ctx* ctx, int n, int prec) {
  double *A, *b, *x;
  HCHECK(hipMalloc(&A, sizeof(double) * n * n));
  HCHECK(hipMalloc(&b, sizeof(double) * n));
  HCHECK(hipMalloc(&x, sizeof(double) * n));
  CHECK(nmx_fill_random_f64(ctx, 1, 7, 0, (int64_t)n * n, 0, 0, A));
  CHECK(nmx_fill_random_f64(ctx, 1, 8, 0, n, 0, 0, b));
  nmx_ir_cfg cfg;
  nmx_ir_cfg_default(&cfg);
  cfg.precision = prec;
  for (int it = 0; it < 2; it++) {
    nmx_ir_report rep;
    CHECK(nmx_gesv_mpir(ctx, n, A, n, b, x, &cfg, &rep));
    CHECK(nmx_sync(ctx));
    printf("mpir step %d: n=%d iterations %d backward error %.3e factor %.2f ms refine %.2f ms\n", it, n,
           rep.iterations, rep.backward_error, rep.factor_ms, rep.refine_ms);
    fflush(stdout);
  }
  return 0;
}

static int run_mprhs(nmx_ctx* ctx, int n, int nrhs) {
  double *A, *B, *X;
  HCHECK(hipMalloc(&A, sizeof(double) * n * n));
  HCHECK(hipMalloc(&B, sizeof(double) * n * nrhs));
  HCHECK(hipMalloc(&X, sizeof(double) * n * nrhs));
  CHECK(nmx_fill_random_f64(ctx, 1, 7, 0, (int64_t)n * n, 0, 0, A));
  CHECK(nmx_fill_random_f64(ctx, 1, 8, 0, (int64_t)n * nrhs, 0, 0, B));
  nmx_ir_cfg cfg;
  nmx_ir_cfg_default(&cfg);
  std::vector<nmx_ir_report> reps(nrhs);
  for (int it = 0; it < 2; it++) {
    CHECK(nmx_gesv_mpir_nrhs(ctx, n, nrhs, A, n, B, nrhs, X, nrhs, &cfg, reps.data()));
    CHECK(nmx_sync(ctx));
    int steps = 0;
    for (auto& r : reps) steps = std::max(steps, r.iterations);
    printf("mprhs step %d: n=%d nrhs=%d iterations %d factor %.2f ms refine %.2f ms\n", it, n, nrhs, steps,
           reps[0].factor_ms, reps[0].refine_ms);
    fflush(stdout);
  }
  return 0;
}

static int run_batch(nmx_ctx* ctx, int count) {
  std::mt19937_64 rng(11);
  std::uniform_int_distribution<int> un(8, 256);
  std::uniform_real_distribution<double> uu(-1.0, 1.0);
  std::normal_distribution<double> nd(0.0, 1.0);
  std::vector<int32_t> n(count);
  std::vector<uint8_t> kind(count);
  std::vector<int64_t> oa(count), ob(count);
  int64_t ae = 0, be = 0;
  for (int p = 0; p < count; p++) {
    n[p] = un(rng);
    kind[p] = (uint8_t)(p & 1);
    oa[p] = ae;
    ob[p] = be;
    ae += (int64_t)n[p] * n[p];
    be += n[p];
  }
  std::vector<double> a(ae), b(be), g;
  for (int p = 0; p < count; p++) {
    const int m = n[p];
    double* ap = a.data() + oa[p];
    if (kind[p] == 0) {
      for (int i = 0; i < m * m; i++) ap[i] = uu(rng);
      for (int i = 0; i < m; i++) ap[i * m + i] += 0.1 * m;
    } else {
      g.resize((size_t)m * m);
      for (auto& v : g) v = nd(rng);
      for (int i = 0; i < m; i++)
        for (int j = 0; j <= i; j++) {
          double s = 0;
          for (int t = 0; t < m; t++) s += g[(size_t)i * m + t] * g[(size_t)j * m + t];
          ap[i * m + j] = ap[j * m + i] = s / m + (i == j ? 1.0 : 0.0);
        }
    }
    for (int i = 0; i < m; i++) b[ob[p] + i] = nd(rng);
  }
  int32_t *dn, *dinfo, *diters;
  int64_t *doa, *dob;
  uint8_t* dk;
  double *dA, *dB, *dX, *dberr;
  HCHECK(hipMalloc(&dn, 4 * count));
  HCHECK(hipMalloc(&dinfo, 4 * count));
  HCHECK(hipMalloc(&diters, 4 * count));
  HCHECK(hipMalloc(&doa, 8 * count));
  HCHECK(hipMalloc(&dob, 8 * count));
  HCHECK(hipMalloc(&dk, count));
  HCHECK(hipMalloc(&dA, 8 * ae));
  HCHECK(hipMalloc(&dB, 8 * be));
  HCHECK(hipMalloc(&dX, 8 * be));
  HCHECK(hipMalloc(&dberr, 8 * count));
  HCHECK(hipMemcpy(dn, n.data(), 4 * count, hipMemcpyHostToDevice));
  HCHECK(hipMemcpy(doa, oa.data(), 8 * count, hipMemcpyHostToDevice));
  HCHECK(hipMemcpy(dob, ob.data(), 8 * count, hipMemcpyHostToDevice));
  HCHECK(hipMemcpy(dk, kind.data(), count, hipMemcpyHostToDevice));
  HCHECK(hipMemcpy(dA, a.data(), 8 * ae, hipMemcpyHostToDevice));
  HCHECK(hipMemcpy(dB, b.data(), 8 * be, hipMemcpyHostToDevice));
  std::vector<int32_t> info(count);
  for (int it = 0; it < 2; it++) {
    CHECK(nmx_batched_gesv(ctx, count, dn, doa, dob, dk, dA, ae, dB, be, dX, dinfo, diters, dberr, nullptr));
    CHECK(nmx_sync(ctx));
    HCHECK(hipMemcpy(info.data(), dinfo, 4 * count, hipMemcpyDeviceToHost));
    int bad = 0;
    for (int p = 0; p < count; p++) bad += info[p] != 0;
    printf("batch step %d: %d systems, %lld matrix elements, %d with info != 0\n", it, count, (long long)ae, bad);
    fflush(stdout);
  }
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: nmx_profile copy BYTES | lu N [NRHS] | chol N [NRHS] | mpir N [PREC] | mprhs N [NRHS] | batch COUNT | "
                    "eigh COUNT | lstsq COUNT\n");
    return 2;
  }
  const char* mode = argv[1];
  if (!strcmp(mode, "copy")) return run_copy(atoll(argv[2]));
  nmx_ctx* ctx;
  CHECK(nmx_ctx_create(0, nullptr, &ctx));
  int rc = 2;
  if (!strcmp(mode, "lu")) rc = run_lu(ctx, atoi(argv[2]), argc > 3 ? atoi(argv[3]) : 1);
  else if (!strcmp(mode, "chol")) rc = run_chol(ctx, atoi(argv[2]), argc > 3 ? atoi(argv[3]) : 1);
  else if (!strcmp(mode, "eigh")) rc = run_eigh(ctx, atoi(argv[2]));
  else if (!strcmp(mode, "lstsq")) rc = run_lstsq(ctx, atoi(argv[2]));
  else if (!strcmp(mode, "mpir")) rc = run_mpir(ctx, atoi(argv[2]), argc > 3 ? atoi(argv[3]) : 2);
  else if (!strcmp(mode, "mprhs")) rc = run_mprhs(ctx, atoi(argv[2]), argc > 3 ? atoi(argv[3]) : 64);
  else if (!strcmp(mode, "batch")) rc = run_batch(ctx, atoi(argv[2]));
  else fprintf(stderr, "unknown mode %s\n", mode);
  CHECK(nmx_ctx_destroy(ctx));
  return rc;
}
