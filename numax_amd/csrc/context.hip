// Context, workspace, mailbox, status strings and the Philox input generator.
#include <stdio.h>
#include <math.h>
#include "nmx_internal.h"

NMX_API int nmx_version(void) { return 1; }

NMX_API const char* nmx_status_string(int status) {
  if (status < 0) return "invalid argument";
  switch (status) {
    case NMX_OK: return "ok";
    case NMX_ERR_HIP: return "HIP runtime error";
    case NMX_ERR_ALLOC: return "device allocation failed";
    case NMX_ERR_UNSUPPORTED: return "unsupported size or option";
    case NMX_ERR_SYNC_TIMEOUT: return "cross-workgroup wait timed out";
    case NMX_ERR_NOT_CONVERGED: return "iterative refinement did not converge";
    case NMX_ERR_IO: return "file could not be opened, read or written";
    case NMX_ERR_FORMAT: return "not a .npy / NMX1 file, or a truncated or unsupported header";
    case NMX_ERR_DTYPE: return "the file holds another dtype";
    case NMX_ERR_SHAPE: return "the file's rank, shape or payload size disagrees with the request";
    case NMX_ERR_LAYOUT: return "Fortran-ordered or big-endian .npy";
    default: return "unknown status";
  }
}

// Scratch of the allocation-free entry points (nmx_batched_gesv_noalloc):
// every device the library runs on holds its own copy from module load to
// process end.  Host code never uses this name as an address -- in host code
// it is a stand-in, not device memory: nmx_ctx_create takes this device's
// address of it with hipGetSymbolAddress and keeps it in the context.
constexpr size_t kStaticScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

__global__ void fill_u32(uint32_t* p, size_t n, uint32_t v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

NMX_API int nmx_ctx_create(int device, void* stream, nmx_ctx** out) {
  if (!out) return -3;
  *out = nullptr;
  int count = 0;
  const hipError_t ec = hipGetDeviceCount(&count);
  if (ec != hipSuccess || count <= 0) {
    fprintf(stderr, "nmx_ctx_create: no HIP device (%s)\n", hipGetErrorString(ec));
    return NMX_ERR_HIP;
  }
  if (device < 0 || device >= count) return -1;
  nmx_ctx* c = new nmx_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess) {
    fprintf(stderr, "nmx_ctx_create: hipSetDevice(%d) failed\n", device);
    delete c;
    return NMX_ERR_HIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) { delete c; return NMX_ERR_HIP; }
  c->num_cus = prop.multiProcessorCount;
  if (stream) {
    c->stream = (hipStream_t)stream;
  } else {
    if (hipStreamCreate(&c->stream) != hipSuccess) {
      fprintf(stderr, "nmx_ctx_create: hipStreamCreate failed\n");
      delete c;
      return NMX_ERR_HIP;
    }
    c->own_stream = true;
  }
  if (hipMalloc(&c->err, 256) != hipSuccess) { nmx_ctx_destroy(c); return NMX_ERR_ALLOC; }
  if (hipMalloc(&c->mailbox, kMailboxBytes) != hipSuccess) { nmx_ctx_destroy(c); return NMX_ERR_ALLOC; }
  if (hipMalloc(&c->tri_inv, 64 * 64 * sizeof(double)) != hipSuccess) { nmx_ctx_destroy(c); return NMX_ERR_ALLOC; }
  c->mailbox_bytes = kMailboxBytes;
  // an error another runtime user of this process left behind is not ours
  (void)hipGetLastError();
  // this device's copy of the static scratch array (device `device` is current)
  void* sp = nullptr;
  const hipError_t es = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
  if (es != hipSuccess || !sp) {
    fprintf(stderr, "nmx_ctx_create: hipGetSymbolAddress(scratch) on device %d: %s\n", device, hipGetErrorString(es));
    nmx_ctx_destroy(c);
    return NMX_ERR_HIP;
  }
  c->static_scratch = sp;
  c->static_scratch_bytes = kStaticScratchBytes;
  size_t words = kMailboxBytes / 4;
  fill_u32<<<(unsigned)((words + 255) / 256), 256, 0, c->stream>>>((uint32_t*)c->mailbox, words, 0u);
  fill_u32<<<1, 64, 0, c->stream>>>((uint32_t*)c->err, 64, 0u);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    fprintf(stderr, "nmx_ctx_create: initializing the context on device %d: %s\n", device, hipGetErrorString(e));
    nmx_ctx_destroy(c);
    return NMX_ERR_HIP;
  }
  *out = c;
  return NMX_OK;
}

int nmx_aux_stream(nmx_ctx* c) {
  if (!c->aux) {
    if (c->capturing) return NMX_ERR_UNSUPPORTED;  // creating a stream under capture
    NMX_HIP(c, hipSetDevice(c->device));
    NMX_HIP(c, hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking));
  }
  if (!c->ahead_ev) NMX_HIP(c, hipEventCreateWithFlags(&c->ahead_ev, hipEventDisableTiming));
  if (!c->behind_ev) NMX_HIP(c, hipEventCreateWithFlags(&c->behind_ev, hipEventDisableTiming));
  return NMX_OK;
}

// Releases every buffer of the context (the caller made sure no captured
// graph can still replay kernels that read them).
static void ctx_release(nmx_ctx* c) {
  if (c->stream) hipStreamSynchronize(c->stream);
  for (auto& r : c->recs) { hipEventDestroy(r.a); hipEventDestroy(r.b); }
  for (auto e : c->pool) hipEventDestroy(e);
  for (void* p : c->retired) hipFree(p);
  if (c->switch_ev) hipEventDestroy(c->switch_ev);
  if (c->aux) {
    hipStreamSynchronize(c->aux);
    hipStreamDestroy(c->aux);
  }
  if (c->ahead_ev) hipEventDestroy(c->ahead_ev);
  if (c->behind_ev) hipEventDestroy(c->behind_ev);
  if (c->ws) hipFree(c->ws);
  if (c->mailbox) hipFree(c->mailbox);
  if (c->err) hipFree(c->err);
  if (c->tri_inv) hipFree(c->tri_inv);
  if (c->mp_cache && c->mp_cache_free) c->mp_cache_free(c->mp_cache);
  if (c->own_stream && c->stream) hipStreamDestroy(c->stream);
  delete c;
}

NMX_API int nmx_ctx_destroy(nmx_ctx* c) {
  if (!c) return -1;
  if (c->destroy_pending) return -1;
  if (c->live_graphs > 0) {
    // graphs captured on this context still hold its buffers and stream:
    // the last nmx_graph_destroy releases them
    c->destroy_pending = true;
    return NMX_OK;
  }
  ctx_release(c);
  return NMX_OK;
}

NMX_API int nmx_ctx_set_stream(nmx_ctx* c, void* stream) {
  if (!c) return -1;
  if (c->capturing) return NMX_ERR_UNSUPPORTED;
  hipStream_t next = (hipStream_t)stream;
  if (next && next == c->stream) return NMX_OK;
  if (c->own_stream && c->live_graphs > 0) return NMX_ERR_UNSUPPORTED;  // the graphs replay on it
  if (!next) {
    NMX_HIP(c, hipStreamCreate(&next));
  }
  // work already queued on the old stream (the shared workspace, mailbox and
  // caches it uses) completes before anything the new stream runs next
  if (c->stream) {
    if (!c->switch_ev) NMX_HIP(c, hipEventCreateWithFlags(&c->switch_ev, hipEventDisableTiming));
    NMX_HIP(c, hipEventRecord(c->switch_ev, c->stream));
    NMX_HIP(c, hipStreamWaitEvent(next, c->switch_ev, 0));
  }
  if (c->own_stream && c->stream) {
    hipStreamSynchronize(c->stream);
    hipStreamDestroy(c->stream);
  }
  c->stream = next;
  c->own_stream = stream == nullptr;
  return NMX_OK;
}

NMX_API int nmx_sync(nmx_ctx* c) {
  if (!c) return -1;
  if (c->capturing) return NMX_ERR_UNSUPPORTED;
  NMX_HIP(c, hipStreamSynchronize(c->stream));
  return NMX_OK;
}

NMX_API int nmx_ctx_check(nmx_ctx* c) {
  if (!c) return -1;
  if (c->capturing) return NMX_ERR_UNSUPPORTED;
  int h = 0;
  NMX_HIP(c, hipMemcpyAsync(&h, c->err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  NMX_HIP(c, hipStreamSynchronize(c->stream));
  if (h != 0) {
    fill_u32<<<1, 64, 0, c->stream>>>((uint32_t*)c->err, 1, 0u);
    NMX_HIP(c, hipGetLastError());
    NMX_HIP(c, hipStreamSynchronize(c->stream));
    return h;
  }
  return NMX_OK;
}

int nmx_workspace(nmx_ctx* c, size_t bytes, void** out) {
  bytes = (bytes + 255) & ~(size_t)255;
  if (bytes > c->ws_bytes) {
    if (c->capturing) return NMX_ERR_UNSUPPORTED;  // run the call once before capturing it
    NMX_HIP(c, hipStreamSynchronize(c->stream));
    if (c->ws) {
      // a live graph may replay kernels that read the old buffer
      if (c->live_graphs > 0) c->retired.push_back(c->ws);
      else hipFree(c->ws);
    }
    c->ws = nullptr;
    c->ws_bytes = 0;
    if (hipMalloc(&c->ws, bytes) != hipSuccess) return NMX_ERR_ALLOC;
    c->ws_bytes = bytes;
  }
  *out = c->ws;
  return NMX_OK;
}

int nmx_mailbox(nmx_ctx* c, void** out) {
  *out = c->mailbox;
  return NMX_OK;
}

int nmx_epoch_reserve(nmx_ctx* c, uint32_t columns, uint32_t* base) {
  // epochs are kernel arguments: a replayed graph would reuse them
  if (c->capturing) return NMX_ERR_UNSUPPORTED;
  if ((uint64_t)c->epoch + columns + 4 >= 0x7fffffffu) {
    size_t words = kMailboxBytes / 4;
    fill_u32<<<(unsigned)((words + 255) / 256), 256, 0, c->stream>>>((uint32_t*)c->mailbox, words, 0u);
    NMX_HIP(c, hipGetLastError());
    c->epoch = 0;
  }
  *base = c->epoch;
  c->epoch += columns;
  return NMX_OK;
}

// ---- Philox4x32-10 (Salmon et al., SC'11) --------------------------------
__device__ __forceinline__ void philox10(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; r++) {
    if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
  }
}

__device__ __forceinline__ double philox_u01(uint64_t seed, uint64_t word) {
  uint64_t off = word >> 1;
  uint32_t c[4] = {(uint32_t)off, (uint32_t)(off >> 32), 0u, 0u};
  philox10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  int pair = (int)(word & 1) * 2;
  uint64_t bits = ((uint64_t)c[pair] << 32) | c[pair + 1];
  return (double)(bits >> 11) * 1.1102230246251565e-16;
}

__global__ void fill_random_kernel(int kind, uint64_t seed, int64_t offset, int64_t count, double lo, double hi,
                                   double* out) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= count) return;
  uint64_t idx = (uint64_t)(offset + e);
  if (kind == 0) {
    // unfused, so the value matches the host stream bit for bit
    out[e] = __dadd_rn(lo, __dmul_rn(__dadd_rn(hi, -lo), philox_u01(seed, idx)));
  } else {
    // Both uniforms of normal i sit in Philox block i.
    uint32_t c[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), 0u, 0u};
    philox10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    uint64_t b1 = ((uint64_t)c[0] << 32) | c[1];
    uint64_t b2 = ((uint64_t)c[2] << 32) | c[3];
    double u1 = (double)(b1 >> 11) * 1.1102230246251565e-16;
    double u2 = (double)(b2 >> 11) * 1.1102230246251565e-16;
    out[e] = sqrt(-2.0 * log(1.0 - u1)) * cos(6.283185307179586 * u2);
  }
}

NMX_API int nmx_fill_random_f64(nmx_ctx* c, int kind, uint64_t seed, int64_t offset, int64_t count, double lo,
                                double hi, double* out) {
  if (!c) return -1;
  if (kind != 0 && kind != 1) return -2;
  if (count < 0) return -5;
  if (count == 0) return NMX_OK;
  if (!out) return -8;
  int64_t blocks = (count + 255) / 256;
  fill_random_kernel<<<(unsigned)blocks, 256, 0, c->stream>>>(kind, seed, offset, count, lo, hi, out);
  NMX_HIP(c, hipGetLastError());
  return NMX_OK;
}

// ---- event profiler ----------------------------------------------------------
static hipEvent_t pool_get(nmx_ctx* c) {
  if (!c->pool.empty()) {
    hipEvent_t e = c->pool.back();
    c->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  hipEventCreate(&e);
  return e;
}

ProfScope::ProfScope(nmx_ctx* c, int kind, double flops, double bytes) : ctx(c), on(c && c->profiling) {
  if (!on) return;
  rec.kind = kind;
  rec.flops = flops;
  rec.bytes = bytes;
  rec.a = pool_get(c);
  rec.b = pool_get(c);
  hipEventRecord(rec.a, c->stream);
}

ProfScope::~ProfScope() {
  if (!on) return;
  hipEventRecord(rec.b, ctx->stream);
  ctx->recs.push_back(rec);
}

static int prof_collect(nmx_ctx* c) {
  if (c->recs.empty()) return NMX_OK;
  NMX_HIP(c, hipStreamSynchronize(c->stream));
  for (auto& r : c->recs) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      c->prof_launches[r.kind] += 1;
      c->prof_ms[r.kind] += ms;
      c->prof_flops[r.kind] += r.flops;
      c->prof_bytes[r.kind] += r.bytes;
    }
    c->pool.push_back(r.a);
    c->pool.push_back(r.b);
  }
  c->recs.clear();
  return NMX_OK;
}

NMX_API int nmx_ctx_set_profiling(nmx_ctx* c, int on) {
  if (!c) return -1;
  NMX_TRY(prof_collect(c));
  c->profiling = on != 0;
  for (int k = 0; k < K_COUNT; k++) {
    c->prof_launches[k] = 0;
    c->prof_ms[k] = c->prof_flops[k] = c->prof_bytes[k] = 0.0;
  }
  return NMX_OK;
}

NMX_API int nmx_ctx_profile(nmx_ctx* c, int kind, int64_t* launches, double* ms, double* flops, double* bytes) {
  if (!c) return -1;
  if (kind < 0 || kind >= K_COUNT) return -2;
  NMX_TRY(prof_collect(c));
  if (launches) *launches = c->prof_launches[kind];
  if (ms) *ms = c->prof_ms[kind];
  if (flops) *flops = c->prof_flops[kind];
  if (bytes) *bytes = c->prof_bytes[kind];
  return NMX_OK;
}

NMX_API const char* nmx_kernel_name(int kind) {
  static const char* names[K_COUNT] = {"gemm",       "getf2_leaf", "laswp",    "trsm_base", "potrf_diag",
                                       "mp_gemm",    "mp_panel",   "residual", "mp_trsv",   "other",
                                       "batched_solve", "batched_eigh", "batched_lstsq", "qr_panel"};
  return (kind >= 0 && kind < K_COUNT) ? names[kind] : "unknown";
}

// ---- graph capture of the context stream -------------------------------------
struct nmx_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t stream = nullptr;
  nmx_ctx* ctx = nullptr;  // whose buffers the replayed kernels read
};

NMX_API int nmx_graph_begin(nmx_ctx* c) {
  if (!c) return -1;
  if (c->capturing || c->profiling) return NMX_ERR_UNSUPPORTED;
  NMX_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  c->capturing = true;
  return NMX_OK;
}

NMX_API int nmx_graph_end(nmx_ctx* c, nmx_graph** out) {
  if (!c) return -1;
  if (!out) return -2;
  *out = nullptr;
  if (!c->capturing) return NMX_ERR_UNSUPPORTED;
  c->capturing = false;
  hipGraph_t g = nullptr;
  NMX_HIP(c, hipStreamEndCapture(c->stream, &g));
  hipGraphExec_t e = nullptr;
  if (hipGraphInstantiate(&e, g, nullptr, nullptr, 0) != hipSuccess) {
    hipGraphDestroy(g);
    return NMX_ERR_HIP;
  }
  nmx_graph* r = new nmx_graph;
  r->graph = g;
  r->exec = e;
  r->stream = c->stream;
  r->ctx = c;
  c->live_graphs++;
  *out = r;
  return NMX_OK;
}

NMX_API int nmx_graph_launch(nmx_graph* g) {
  if (!g) return -1;
  return hipGraphLaunch(g->exec, g->stream) == hipSuccess ? NMX_OK : NMX_ERR_HIP;
}

NMX_API int nmx_graph_destroy(nmx_graph* g) {
  if (!g) return NMX_OK;
  if (g->stream) hipStreamSynchronize(g->stream);  // no replay still running
  if (g->exec) hipGraphExecDestroy(g->exec);
  if (g->graph) hipGraphDestroy(g->graph);
  nmx_ctx* c = g->ctx;
  delete g;
  if (c && --c->live_graphs == 0) {
    if (c->destroy_pending) {
      ctx_release(c);
    } else {
      for (void* p : c->retired) hipFree(p);
      c->retired.clear();
    }
  }
  return NMX_OK;
}
