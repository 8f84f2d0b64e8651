// Tensor files between disk and device memory: NumPy's .npy (numax.io.numpy,
// numax/io/npy.mojo) and numax's own NMX1 (numax.io.nmx, numax/io/io.mojo).
//
// numax reads a whole file into host memory and copies it to the device
// through Tensor's constructor.  Here the payload streams: two pinned
// staging buffers alternate, the file is read into one while the other is
// copied to the device on the context's stream (saving runs the same pipe
// backwards), so a 2 GiB matrix moves at the slower of disk and PCIe rather
// than their sum, and host memory stays at two chunks.  Headers are parsed
// and written here in C++ with numax's rules: the three keys of the .npy
// dict, '<' or '=' (or '|' for one-byte types) byte order, no Fortran order,
// .npy format 1.x / 2.x on load and 1.0 on save padded as numpy.save pads
// (so the bytes equal numpy.save's); NMX1 with its full dtype name.
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "nmx_internal.h"

namespace {

struct DtInfo {
  const char* descr;  // numpy.save's spelling
  const char* name;   // numax's DType name (NMX1)
  int size;
};
// dtype codes of the C ABI, in this order
const DtInfo kDt[] = {{"<f2", "float16", 2}, {"<f4", "float32", 4}, {"<f8", "float64", 8}, {"|i1", "int8", 1},
                      {"<i2", "int16", 2},   {"<i4", "int32", 4},   {"<i8", "int64", 8},   {"|u1", "uint8", 1},
                      {"<u2", "uint16", 2},  {"<u4", "uint32", 4},  {"<u8", "uint64", 8},  {"|b1", "bool", 1}};
constexpr int kNumDt = (int)(sizeof(kDt) / sizeof(kDt[0]));
constexpr int kMaxRank = 8;
constexpr size_t kChunk = 32u << 20;  // bytes per staging buffer

struct Header {
  int dtype = -1;
  int rank = 0;
  int64_t dims[kMaxRank] = {};
  int64_t offset = 0;  // first payload byte
};

struct File {
  FILE* f = nullptr;
  explicit File(const char* path, const char* mode) : f(fopen(path, mode)) {}
  ~File() {
    if (f) fclose(f);
  }
};

// '<' and '=' both mean little-endian on every target; a one-byte type may
// also say '|' (numax/io/npy.mojo:455-482)
bool descr_matches(const std::string& file, const char* expected) {
  if (file == expected) return true;
  if (file.size() != 3 || strlen(expected) != 3) return false;
  if (file[1] != expected[1] || file[2] != expected[2]) return false;
  const bool native = file[0] == '<' || file[0] == '=';
  if (file[2] == '1') return native || file[0] == '|';
  return native;
}

// the value of `key` in the header dict: from the colon after the key to the
// next comma at bracket depth zero, quotes and spaces dropped
// (numax/io/npy.mojo:484-531)
bool dict_value(const std::string& h, const char* key, std::string& out) {
  const size_t start = h.find(std::string("'") + key + "'");
  if (start == std::string::npos) return false;
  const size_t colon = h.find(':', start);
  if (colon == std::string::npos) return false;
  out.clear();
  int depth = 0;
  for (size_t i = colon + 1; i < h.size(); i++) {
    const char c = h[i];
    if (c == '(' || c == '[' || c == '{') {
      depth++;
    } else if (c == ')' || c == ']' || c == '}') {
      if (depth == 0) break;
      depth--;
    } else if ((c == ',' || c == '\n') && depth == 0) {
      break;
    }
    if (c == ' ' || c == '\'' || c == '"') continue;
    out.push_back(c);
  }
  return true;
}

// the integers of the shape tuple text, "(2, 3)" or "(3,)" or "()"
bool parse_shape(const std::string& t, Header& h) {
  h.rank = 0;
  int64_t v = 0;
  bool in = false;
  for (char c : t) {
    if (c >= '0' && c <= '9') {
      v = v * 10 + (c - '0');
      in = true;
    } else if (in) {
      if (h.rank == kMaxRank) return false;
      h.dims[h.rank++] = v;
      v = 0;
      in = false;
    }
  }
  if (in) {
    if (h.rank == kMaxRank) return false;
    h.dims[h.rank++] = v;
  }
  return true;
}

int read_exact(FILE* f, void* p, size_t n) { return fread(p, 1, n, f) == n ? NMX_OK : NMX_ERR_FORMAT; }

int read_header(FILE* f, int format, Header& h) {
  if (format == 0) {
    unsigned char pre[12];
    if (read_exact(f, pre, 10)) return NMX_ERR_FORMAT;  // too short to be a .npy file
    if (pre[0] != 0x93 || memcmp(pre + 1, "NUMPY", 5) != 0) return NMX_ERR_FORMAT;
    const int major = pre[6];
    size_t hlen, start;
    if (major == 1) {
      hlen = pre[8] | (pre[9] << 8);
      start = 10;
    } else if (major == 2 || major == 3) {
      if (read_exact(f, pre + 10, 2)) return NMX_ERR_FORMAT;
      hlen = (size_t)pre[8] | ((size_t)pre[9] << 8) | ((size_t)pre[10] << 16) | ((size_t)pre[11] << 24);
      start = 12;
    } else {
      return NMX_ERR_FORMAT;
    }
    std::string text(hlen, '\0');
    if (read_exact(f, &text[0], hlen)) return NMX_ERR_FORMAT;
    std::string descr, fortran, shape;
    if (!dict_value(text, "descr", descr) || !dict_value(text, "fortran_order", fortran) ||
        !dict_value(text, "shape", shape))
      return NMX_ERR_FORMAT;
    h.dtype = -1;
    for (int d = 0; d < kNumDt; d++)
      if (descr_matches(descr, kDt[d].descr)) h.dtype = d;
    if (h.dtype < 0) return descr.size() == 3 && descr[0] == '>' ? NMX_ERR_LAYOUT : NMX_ERR_DTYPE;
    if (fortran != "False") return NMX_ERR_LAYOUT;
    if (!parse_shape(shape, h)) return NMX_ERR_FORMAT;
    h.offset = (int64_t)(start + hlen);
    return NMX_OK;
  }
  unsigned char pre[6];
  if (read_exact(f, pre, 5)) return NMX_ERR_FORMAT;
  if (memcmp(pre, "NMX1", 4) != 0) return NMX_ERR_FORMAT;
  const int name_len = pre[4];
  std::string name(name_len, '\0');
  if (read_exact(f, &name[0], (size_t)name_len)) return NMX_ERR_FORMAT;
  h.dtype = -1;
  for (int d = 0; d < kNumDt; d++)
    if (name == kDt[d].name) h.dtype = d;
  if (h.dtype < 0) return NMX_ERR_DTYPE;
  unsigned char r;
  if (read_exact(f, &r, 1)) return NMX_ERR_FORMAT;
  if (r > kMaxRank) return NMX_ERR_FORMAT;
  h.rank = r;
  for (int i = 0; i < h.rank; i++) {
    unsigned char b[8];
    if (read_exact(f, b, 8)) return NMX_ERR_FORMAT;
    uint64_t v = 0;
    for (int k = 0; k < 8; k++) v |= (uint64_t)b[k] << (8 * k);
    h.dims[i] = (int64_t)v;
  }
  h.offset = 5 + name_len + 1 + 8 * (int64_t)h.rank;
  return NMX_OK;
}

int64_t payload_bytes(const Header& h) {
  int64_t count = 1;
  for (int i = 0; i < h.rank; i++) count *= h.dims[i];
  return count * kDt[h.dtype].size;
}

std::string npy_header(int dtype, int rank, const int64_t* dims) {
  std::string shape = "(";
  for (int i = 0; i < rank; i++) {
    shape += std::to_string((long long)dims[i]);
    if (i != rank - 1) shape += ", ";
  }
  if (rank == 1) shape += ",";
  shape += ")";
  std::string dict = std::string("{'descr': '") + kDt[dtype].descr + "', 'fortran_order': False, 'shape': " + shape +
                     ", }";
  // numpy.lib.format: magic (6) + version (2) + length (2) + dict + padding
  // + newline is a multiple of 64
  const size_t pad = 64 - ((10 + dict.size() + 1) % 64);
  dict.append(pad, ' ');
  dict.push_back('\n');
  std::string out;
  out.push_back((char)0x93);
  out += "NUMPY";
  out.push_back((char)1);
  out.push_back((char)0);
  out.push_back((char)(dict.size() & 0xFF));
  out.push_back((char)((dict.size() >> 8) & 0xFF));
  return out + dict;
}

std::string nmx_header(int dtype, int rank, const int64_t* dims) {
  std::string out = "NMX1";
  out.push_back((char)strlen(kDt[dtype].name));
  out += kDt[dtype].name;
  out.push_back((char)rank);
  for (int i = 0; i < rank; i++) {
    const uint64_t v = (uint64_t)dims[i];
    for (int k = 0; k < 8; k++) out.push_back((char)((v >> (8 * k)) & 0xFF));
  }
  return out;
}

// Two pinned staging buffers (hipHostMalloc) and an event per buffer.
struct Staging {
  void* buf[2] = {nullptr, nullptr};
  hipEvent_t done[2] = {nullptr, nullptr};
  ~Staging() {
    for (int i = 0; i < 2; i++) {
      if (done[i]) {
        hipEventSynchronize(done[i]);
        hipEventDestroy(done[i]);
      }
      if (buf[i]) hipHostFree(buf[i]);
    }
  }
  int init(nmx_ctx* ctx, size_t bytes) {
    for (int i = 0; i < 2; i++) {
      NMX_HIP(ctx, hipHostMalloc(&buf[i], bytes, hipHostMallocDefault));
      NMX_HIP(ctx, hipEventCreateWithFlags(&done[i], hipEventDisableTiming));
    }
    return NMX_OK;
  }
};

int stream_to_device(nmx_ctx* ctx, FILE* f, char* dst, int64_t nbytes) {
  if (nbytes == 0) return NMX_OK;
  Staging st;
  NMX_TRY(st.init(ctx, (size_t)std::min<int64_t>(nbytes, (int64_t)kChunk)));
  int k = 0;
  for (int64_t off = 0; off < nbytes; off += kChunk, k ^= 1) {
    const size_t len = (size_t)std::min<int64_t>(kChunk, nbytes - off);
    NMX_HIP(ctx, hipEventSynchronize(st.done[k]));  // this buffer's previous copy has left it
    if (fread(st.buf[k], 1, len, f) != len) return NMX_ERR_IO;
    NMX_HIP(ctx, hipMemcpyAsync(dst + off, st.buf[k], len, hipMemcpyHostToDevice, ctx->stream));
    NMX_HIP(ctx, hipEventRecord(st.done[k], ctx->stream));
  }
  NMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return NMX_OK;
}

int stream_from_device(nmx_ctx* ctx, FILE* f, const char* src, int64_t nbytes) {
  if (nbytes == 0) return NMX_OK;
  Staging st;
  NMX_TRY(st.init(ctx, (size_t)std::min<int64_t>(nbytes, (int64_t)kChunk)));
  // chunk i's copy is queued before chunk i-1 is written out
  const int64_t chunks = (nbytes + (int64_t)kChunk - 1) / (int64_t)kChunk;
  auto len_of = [&](int64_t i) { return (size_t)std::min<int64_t>(kChunk, nbytes - i * (int64_t)kChunk); };
  NMX_HIP(ctx, hipMemcpyAsync(st.buf[0], src, len_of(0), hipMemcpyDeviceToHost, ctx->stream));
  NMX_HIP(ctx, hipEventRecord(st.done[0], ctx->stream));
  for (int64_t i = 0; i < chunks; i++) {
    const int k = (int)(i & 1);
    if (i + 1 < chunks) {
      NMX_HIP(ctx, hipMemcpyAsync(st.buf[k ^ 1], src + (i + 1) * (int64_t)kChunk, len_of(i + 1),
                                  hipMemcpyDeviceToHost, ctx->stream));
      NMX_HIP(ctx, hipEventRecord(st.done[k ^ 1], ctx->stream));
    }
    NMX_HIP(ctx, hipEventSynchronize(st.done[k]));
    if (fwrite(st.buf[k], 1, len_of(i), f) != len_of(i)) return NMX_ERR_IO;
  }
  return NMX_OK;
}

int check_request(int format, int dtype, int rank, const int64_t* dims) {
  if (format != 0 && format != 1) return -3;
  if (dtype < 0 || dtype >= kNumDt) return -4;
  if (rank < 0 || rank > kMaxRank) return -5;
  if (rank > 0 && !dims) return -6;
  for (int i = 0; i < rank; i++)
    if (dims[i] < 0) return -6;
  return NMX_OK;
}

}  // namespace

NMX_API int nmx_tensor_file_info(const char* path, int format, int* dtype, int* rank, int64_t* dims) {
  if (!path) return -1;
  if (format != 0 && format != 1) return -2;
  if (!dtype || !rank || !dims) return -3;
  File file(path, "rb");
  if (!file.f) return NMX_ERR_IO;
  Header h;
  NMX_TRY(read_header(file.f, format, h));
  *dtype = h.dtype;
  *rank = h.rank;
  for (int i = 0; i < h.rank; i++) dims[i] = h.dims[i];
  return NMX_OK;
}

NMX_API int nmx_tensor_file_load(nmx_ctx* ctx, const char* path, int format, int dtype, int rank,
                                 const int64_t* dims, void* dst) {
  if (!ctx) return -1;
  if (!path) return -2;
  NMX_TRY(check_request(format, dtype, rank, dims));
  File file(path, "rb");
  if (!file.f) return NMX_ERR_IO;
  Header h;
  NMX_TRY(read_header(file.f, format, h));
  if (h.dtype != dtype) return NMX_ERR_DTYPE;
  if (h.rank != rank) return NMX_ERR_SHAPE;
  for (int i = 0; i < rank; i++)
    if (h.dims[i] != dims[i]) return NMX_ERR_SHAPE;
  const int64_t nbytes = payload_bytes(h);
  if (nbytes > 0 && !dst) return -7;
  if (fseeko(file.f, 0, SEEK_END) != 0) return NMX_ERR_IO;
  const int64_t size = (int64_t)ftello(file.f);
  if (size - h.offset != nbytes) return NMX_ERR_SHAPE;  // payload size disagrees with the header
  if (fseeko(file.f, (off_t)h.offset, SEEK_SET) != 0) return NMX_ERR_IO;
  return stream_to_device(ctx, file.f, (char*)dst, nbytes);
}

NMX_API int nmx_tensor_file_save(nmx_ctx* ctx, const char* path, int format, int dtype, int rank,
                                 const int64_t* dims, const void* src) {
  if (!ctx) return -1;
  if (!path) return -2;
  NMX_TRY(check_request(format, dtype, rank, dims));
  Header h;
  h.dtype = dtype;
  h.rank = rank;
  for (int i = 0; i < rank; i++) h.dims[i] = dims[i];
  const int64_t nbytes = payload_bytes(h);
  if (nbytes > 0 && !src) return -7;
  const std::string head = format == 0 ? npy_header(dtype, rank, dims) : nmx_header(dtype, rank, dims);
  File file(path, "wb");
  if (!file.f) return NMX_ERR_IO;
  if (fwrite(head.data(), 1, head.size(), file.f) != head.size()) return NMX_ERR_IO;
  NMX_TRY(stream_from_device(ctx, file.f, (const char*)src, nbytes));
  return fflush(file.f) == 0 ? NMX_OK : NMX_ERR_IO;
}
