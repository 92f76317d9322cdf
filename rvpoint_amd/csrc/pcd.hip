// pcd.hip -- PCD ingest (src/io/simple_pcd_loader.h:143-293) into host or device SoA, and the writer
// (savePCD, :295-328).
//
// The header rules follow the reference's parser: '#' lines and empty lines are skipped; FIELDS, SIZE, TYPE,
// COUNT (default 1), POINTS (WIDTH only while POINTS is unset) and DATA are read; x, y and z must be 4-byte
// fields.  DATA ascii parses one point per line (lines that fail to parse are skipped), binary holds
// interleaved records, binary_compressed holds LZF-compressed field-major blocks.  For the device path a
// binary payload goes to the GPU as is and is de-interleaved there; a compressed payload is decompressed on
// the host (LZF is sequential) and its x, y and z blocks are copied straight into the device arrays.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "common.hpp"

namespace rvp {

struct PcdHeader {
  std::vector<std::string> fields;
  std::vector<int> sizes, counts;
  std::vector<char> types;
  size_t points = 0;
  std::string data;
  size_t data_offset = 0;  // bytes before the payload
};

static bool pcd_header(std::ifstream &f, PcdHeader &h) {
  std::string line;
  while (std::getline(f, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line.empty() || line[0] == '#') continue;
    std::stringstream ss(line);
    std::string tag;
    ss >> tag;
    if (tag == "FIELDS") {
      h.fields.clear();
      std::string v;
      while (ss >> v) h.fields.push_back(v);
    } else if (tag == "SIZE") {
      h.sizes.clear();
      int v;
      while (ss >> v) h.sizes.push_back(v);
    } else if (tag == "TYPE") {
      h.types.clear();
      char v;
      while (ss >> v) h.types.push_back(v);
    } else if (tag == "COUNT") {
      h.counts.clear();
      int v;
      while (ss >> v) h.counts.push_back(v);
    } else if (tag == "POINTS") {
      ss >> h.points;
    } else if (tag == "WIDTH" && h.points == 0) {
      ss >> h.points;
    } else if (tag == "DATA") {
      ss >> h.data;
      break;
    }
  }
  if (h.fields.empty() || h.sizes.size() != h.fields.size() || h.types.size() != h.fields.size()) return false;
  if (h.counts.empty()) h.counts.assign(h.fields.size(), 1);
  if (h.counts.size() != h.fields.size()) return false;
  h.data_offset = (size_t)f.tellg();
  return h.points > 0 && !h.data.empty();
}

static int field_index(const PcdHeader &h, const char *name) {
  for (size_t i = 0; i < h.fields.size(); ++i)
    if (h.fields[i] == name) return (int)i;
  return -1;
}

// LZF (the format of liblzf): a control byte below 32 starts a literal run of ctrl + 1 bytes; otherwise
// the top 3 bits are a length (7: add the next byte) and the low 5 bits with the next byte a distance,
// and len + 2 bytes are copied from distance + 1 bytes back
static bool lzf_decode(const unsigned char *in, size_t in_len, unsigned char *out, size_t out_len) {
  size_t ip = 0, op = 0;
  while (ip < in_len) {
    unsigned ctrl = in[ip++];
    if (ctrl < 32) {
      const size_t run = ctrl + 1;
      if (ip + run > in_len || op + run > out_len) return false;
      std::memcpy(out + op, in + ip, run);
      ip += run;
      op += run;
    } else {
      size_t len = ctrl >> 5;
      if (ip >= in_len) return false;
      if (len == 7) {
        len += in[ip++];
        if (ip >= in_len) return false;
      }
      const size_t dist = ((size_t)(ctrl & 31) << 8) + in[ip++] + 1;
      if (dist > op || op + len + 2 > out_len) return false;
      for (size_t k = 0; k < len + 2; ++k, ++op) out[op] = out[op - dist];  // may overlap: byte by byte
    }
  }
  return op == out_len;
}

// what one load needs: the header, where x/y/z sit, and the payload in host memory
struct PcdData {
  PcdHeader h;
  int xi = -1, yi = -1, zi = -1;
  std::vector<unsigned char> raw;   // binary: interleaved records; compressed: decompressed blocks
  std::vector<float> ax, ay, az;    // ascii: parsed values
  size_t n = 0;
  size_t stride = 0;
  size_t off[3] = {0, 0, 0};        // binary: byte offset of x/y/z in a record; compressed: block starts
};

static int pcd_load(const char *path, PcdData &d) {
  std::ifstream f(path, std::ios::binary);
  RVP_CHECK(f.is_open(), RVP_E_ARG, std::string("cannot open ") + path);
  RVP_CHECK(pcd_header(f, d.h), RVP_E_ARG, "failed to parse the PCD header");
  const PcdHeader &h = d.h;
  d.xi = field_index(h, "x");
  d.yi = field_index(h, "y");
  d.zi = field_index(h, "z");
  RVP_CHECK(d.xi >= 0 && d.yi >= 0 && d.zi >= 0, RVP_E_ARG, "PCD is missing x/y/z fields");
  const int xyz[3] = {d.xi, d.yi, d.zi};
  if (h.data == "ascii") {
    std::string line;
    std::vector<float> v(h.fields.size());
    while (std::getline(f, line)) {
      if (line.empty()) continue;
      std::stringstream ss(line);
      std::fill(v.begin(), v.end(), 0.0f);
      for (size_t i = 0; i < v.size(); ++i) ss >> v[i];
      if (ss.fail()) continue;
      d.ax.push_back(v[d.xi]);
      d.ay.push_back(v[d.yi]);
      d.az.push_back(v[d.zi]);
    }
    d.n = d.ax.size();
    return RVP_OK;
  }
  for (int a = 0; a < 3; ++a)
    RVP_CHECK(h.sizes[xyz[a]] == 4, RVP_E_UNSUPPORTED, "x/y/z fields must be 4 bytes");
  std::vector<size_t> foff(h.fields.size());
  size_t stride = 0;
  for (size_t i = 0; i < h.fields.size(); ++i) {
    foff[i] = stride;
    stride += (size_t)h.sizes[i] * (size_t)h.counts[i];
  }
  d.n = h.points;
  d.stride = stride;
  if (h.data == "binary") {
    d.raw.resize(stride * h.points);
    f.read(reinterpret_cast<char *>(d.raw.data()), (std::streamsize)d.raw.size());
    RVP_CHECK((size_t)f.gcount() == d.raw.size(), RVP_E_ARG, "binary PCD payload is short");
    for (int a = 0; a < 3; ++a) d.off[a] = foff[xyz[a]];
    return RVP_OK;
  }
  if (h.data == "binary_compressed") {
    uint32_t csize = 0, usize = 0;
    f.read(reinterpret_cast<char *>(&csize), 4);
    f.read(reinterpret_cast<char *>(&usize), 4);
    RVP_CHECK((bool)f, RVP_E_ARG, "could not read the compressed sizes");
    std::vector<unsigned char> comp(csize);
    f.read(reinterpret_cast<char *>(comp.data()), (std::streamsize)csize);
    RVP_CHECK((size_t)f.gcount() == csize, RVP_E_ARG, "compressed PCD payload is short");
    d.raw.resize(usize);
    RVP_CHECK(lzf_decode(comp.data(), csize, d.raw.data(), usize), RVP_E_ARG, "LZF payload is corrupt");
    size_t cursor = 0;
    std::vector<size_t> block(h.fields.size());
    for (size_t i = 0; i < h.fields.size(); ++i) {
      block[i] = cursor;
      cursor += (size_t)h.sizes[i] * (size_t)h.counts[i] * h.points;
    }
    RVP_CHECK(cursor <= usize, RVP_E_ARG, "decompressed PCD payload is smaller than its fields");
    for (int a = 0; a < 3; ++a) d.off[a] = block[xyz[a]];
    return RVP_OK;
  }
  return fail(RVP_E_UNSUPPORTED, "unsupported PCD DATA mode '" + h.data + "'");
}

// de-interleave x/y/z of binary records (any stride and field offsets; byte loads when unaligned)
__global__ void pcd_deinterleave_k(const unsigned char *__restrict__ raw, size_t n, size_t stride, size_t ox,
                                   size_t oy, size_t oz, float *__restrict__ x, float *__restrict__ y,
                                   float *__restrict__ z) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned char *r = raw + i * stride;
  const size_t o[3] = {ox, oy, oz};
  float *dst[3] = {x, y, z};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const unsigned char *p = r + o[a];
    uint32_t v;
    if ((((uintptr_t)p) & 3u) == 0u) {
      v = *reinterpret_cast<const uint32_t *>(p);
    } else {
      v = (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
    }
    dst[a][i] = __uint_as_float(v);
  }
}

}  // namespace rvp

using namespace rvp;

extern "C" {

int rvp_pcd_info(const char *path, size_t *n_points, int *data_mode) {
  RVP_CHECK(path && n_points, RVP_E_ARG, "null argument");
  std::ifstream f(path, std::ios::binary);
  RVP_CHECK(f.is_open(), RVP_E_ARG, std::string("cannot open ") + path);
  PcdHeader h;
  RVP_CHECK(pcd_header(f, h), RVP_E_ARG, "failed to parse the PCD header");
  *n_points = h.points;
  if (data_mode) *data_mode = h.data == "ascii" ? 0 : (h.data == "binary" ? 1 : (h.data == "binary_compressed" ? 2 : -1));
  return RVP_OK;
}

int rvp_pcd_read(const char *path, float *x, float *y, float *z, size_t cap, size_t *n) {
  RVP_CHECK(path && n, RVP_E_ARG, "null argument");
  PcdData d;
  RVP_TRY(pcd_load(path, d));
  *n = d.n;
  RVP_CHECK(cap >= d.n, RVP_E_ARG, "output capacity below the point count (*n holds it)");
  RVP_CHECK(x && y && z, RVP_E_ARG, "null output");
  if (d.h.data == "ascii") {
    std::memcpy(x, d.ax.data(), d.n * 4);
    std::memcpy(y, d.ay.data(), d.n * 4);
    std::memcpy(z, d.az.data(), d.n * 4);
  } else if (d.h.data == "binary") {
    float *dst[3] = {x, y, z};
    for (size_t i = 0; i < d.n; ++i)
      for (int a = 0; a < 3; ++a) std::memcpy(dst[a] + i, d.raw.data() + i * d.stride + d.off[a], 4);
  } else {
    std::memcpy(x, d.raw.data() + d.off[0], d.n * 4);
    std::memcpy(y, d.raw.data() + d.off[1], d.n * 4);
    std::memcpy(z, d.raw.data() + d.off[2], d.n * 4);
  }
  return RVP_OK;
}

int rvp_pcd_read_device(const char *path, float *x, float *y, float *z, size_t cap, size_t *n, rvp_stream_t st) {
  RVP_CHECK(path && n, RVP_E_ARG, "null argument");
  RVP_TRY(pool_init());
  PcdData d;
  RVP_TRY(pcd_load(path, d));
  *n = d.n;
  RVP_CHECK(cap >= d.n, RVP_E_ARG, "output capacity below the point count (*n holds it)");
  RVP_CHECK(x && y && z, RVP_E_ARG, "null output");
  if (d.n == 0) return RVP_OK;
  hipStream_t s = S(st);
  if (d.h.data == "ascii") {
    RVP_HIP(hipMemcpyAsync(x, d.ax.data(), d.n * 4, hipMemcpyHostToDevice, s));
    RVP_HIP(hipMemcpyAsync(y, d.ay.data(), d.n * 4, hipMemcpyHostToDevice, s));
    RVP_HIP(hipMemcpyAsync(z, d.az.data(), d.n * 4, hipMemcpyHostToDevice, s));
  } else if (d.h.data == "binary") {
    Tmp raw;
    RVP_TRY(raw.alloc(d.raw.size(), s));
    RVP_HIP(hipMemcpyAsync(raw.p, d.raw.data(), d.raw.size(), hipMemcpyHostToDevice, s));
    RVP_LAUNCH("pcd_deinterleave_k", s, pcd_deinterleave_k<<<ceil_div(d.n, 256), 256, 0, s>>>(
                                            raw.as<unsigned char>(), d.n, d.stride, d.off[0], d.off[1], d.off[2],
                                            x, y, z));
  } else {
    RVP_HIP(hipMemcpyAsync(x, d.raw.data() + d.off[0], d.n * 4, hipMemcpyHostToDevice, s));
    RVP_HIP(hipMemcpyAsync(y, d.raw.data() + d.off[1], d.n * 4, hipMemcpyHostToDevice, s));
    RVP_HIP(hipMemcpyAsync(z, d.raw.data() + d.off[2], d.n * 4, hipMemcpyHostToDevice, s));
  }
  RVP_HIP(hipStreamSynchronize(s));  // the host payload dies with this call
  return RVP_OK;
}

int rvp_pcd_write(const char *path, const float *x, const float *y, const float *z, size_t n, int binary) {
  RVP_CHECK(path && (n == 0 || (x && y && z)), RVP_E_ARG, "null argument");
  std::ofstream f(path, binary ? (std::ios::binary | std::ios::out) : std::ios::out);
  RVP_CHECK(f.is_open(), RVP_E_ARG, std::string("cannot open ") + path + " for writing");
  f << "# .PCD v.7 - Point Cloud Data file format\n"
    << "VERSION .7\nFIELDS x y z\nSIZE 4 4 4\nTYPE F F F\nCOUNT 1 1 1\n"
    << "WIDTH " << n << "\nHEIGHT 1\nVIEWPOINT 0 0 0 1 0 0 0\nPOINTS " << n << "\n"
    << "DATA " << (binary ? "binary" : "ascii") << "\n";
  for (size_t i = 0; i < n; ++i) {
    if (binary) {
      f.write(reinterpret_cast<const char *>(x + i), 4);
      f.write(reinterpret_cast<const char *>(y + i), 4);
      f.write(reinterpret_cast<const char *>(z + i), 4);
    } else {
      f << x[i] << " " << y[i] << " " << z[i] << "\n";  // default stream precision, as savePCD
    }
  }
  RVP_CHECK((bool)f, RVP_E_ARG, "write failed");
  return RVP_OK;
}

}  // extern "C"
