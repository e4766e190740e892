// vh_io.cpp - settings, depth-image and pose loading (include/vh_io.h). Host code, no GPU.
//
// The reference reads these through OpenCV (cv::FileStorage, cv::imread, an ifstream;
// src/PointCloudGenerator.cpp:20-45, src/SaveFrame.cpp:154-218). OpenCV is not a dependency
// here: the settings file is read as the OpenCV-YAML scalar subset it uses, and depth images
// as 16-bit greyscale PNG through zlib.
#include <zlib.h>

#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "vh_error.h"
#include "vh_io.h"

namespace {

using vh::set_last_error;

int fail(int code, const std::string& msg) { return set_last_error(code, msg); }

std::string trim(const std::string& s) {
  const size_t a = s.find_first_not_of(" \t\r\n");
  if (a == std::string::npos) return "";
  const size_t b = s.find_last_not_of(" \t\r\n");
  return s.substr(a, b - a + 1);
}

bool parse_number(const std::string& v, double& out) {
  if (v.empty()) return false;
  char* end = nullptr;
  errno = 0;
  const double d = std::strtod(v.c_str(), &end);
  if (end == v.c_str() || errno == ERANGE || !std::isfinite(d)) return false;
  if (!trim(std::string(end)).empty()) return false;
  out = d;
  return true;
}

bool read_file(const char* path, std::vector<unsigned char>& out) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  unsigned char buf[1 << 16];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) out.insert(out.end(), buf, buf + n);
  const bool ok = !std::ferror(f);
  std::fclose(f);
  return ok;
}

uint32_t be32(const unsigned char* p) {
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | (uint32_t)p[3];
}
void put_be32(std::vector<unsigned char>& v, uint32_t x) {
  v.push_back((unsigned char)(x >> 24));
  v.push_back((unsigned char)(x >> 16));
  v.push_back((unsigned char)(x >> 8));
  v.push_back((unsigned char)x);
}

const unsigned char kPngSig[8] = {0x89, 'P', 'N', 'G', 0x0d, 0x0a, 0x1a, 0x0a};

// Parsed 16-bit greyscale PNG: size and the concatenated IDAT stream.
struct Png {
  uint32_t w = 0, h = 0;
  std::vector<unsigned char> idat;
};

int parse_png(const char* path, Png& png, bool need_data) {
  std::vector<unsigned char> f;
  if (!read_file(path, f)) return fail(VH_ERR_IO, std::string("cannot read ") + path);
  if (f.size() < 8 + 25 || std::memcmp(f.data(), kPngSig, 8) != 0)
    return fail(VH_ERR_INVALID_ARG, std::string(path) + ": not a PNG file");
  size_t p = 8;
  bool have_ihdr = false, have_end = false;
  while (p + 12 <= f.size()) {
    const uint32_t len = be32(&f[p]);
    if (len > f.size() - p - 12) return fail(VH_ERR_INVALID_ARG, std::string(path) + ": truncated chunk");
    const unsigned char* type = &f[p + 4];
    const unsigned char* data = &f[p + 8];
    const uint32_t crc = be32(&f[p + 8 + len]);
    if ((uint32_t)crc32(0, type, 4 + len) != crc) return fail(VH_ERR_INVALID_ARG, std::string(path) + ": bad chunk CRC");
    if (!std::memcmp(type, "IHDR", 4)) {
      if (len != 13) return fail(VH_ERR_INVALID_ARG, std::string(path) + ": bad IHDR");
      png.w = be32(data);
      png.h = be32(data + 4);
      const int depth = data[8], ctype = data[9], comp = data[10], filt = data[11], inter = data[12];
      if (depth != 16 || ctype != 0 || comp != 0 || filt != 0 || inter != 0)
        return fail(VH_ERR_INVALID_ARG, std::string(path) + ": only 16-bit greyscale, non-interlaced PNG is read");
      if (png.w == 0 || png.h == 0 || png.w > 65536 || png.h > 65536)
        return fail(VH_ERR_INVALID_ARG, std::string(path) + ": bad image size");
      have_ihdr = true;
    } else if (!std::memcmp(type, "IDAT", 4)) {
      if (!have_ihdr) return fail(VH_ERR_INVALID_ARG, std::string(path) + ": IDAT before IHDR");
      if (need_data) png.idat.insert(png.idat.end(), data, data + len);
    } else if (!std::memcmp(type, "IEND", 4)) {
      have_end = true;
      break;
    } else if (!(type[0] & 0x20)) {
      return fail(VH_ERR_INVALID_ARG, std::string(path) + ": unknown critical chunk");
    }
    if (!have_ihdr) return fail(VH_ERR_INVALID_ARG, std::string(path) + ": IHDR is not the first chunk");
    p += 12 + len;
  }
  if (!have_ihdr || !have_end) return fail(VH_ERR_INVALID_ARG, std::string(path) + ": incomplete PNG");
  return VH_OK;
}

int paeth(int a, int b, int c) {
  const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
  return (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
}

}  // namespace

extern "C" {

int vh_settings_load(const char* path, vh_settings* out) {
  if (!path || !out) return fail(VH_ERR_INVALID_ARG, "null argument");
  std::memset(out, 0, sizeof *out);
  std::ifstream in(path);
  if (!in) return fail(VH_ERR_IO, std::string("cannot read ") + path);
  struct Key {
    const char* name;
    uint32_t bit;
    int kind;  // 0 float, 1 int
    void* dst;
  };
  uint32_t origin_bits = 0, dim_bits = 0;
  const Key keys[] = {
      {"Camera.fx", VH_SET_FX, 0, &out->fx},
      {"Camera.fy", VH_SET_FY, 0, &out->fy},
      {"Camera.cx", VH_SET_CX, 0, &out->cx},
      {"Camera.cy", VH_SET_CY, 0, &out->cy},
      {"Camera.width", VH_SET_WIDTH, 1, &out->width},
      {"Camera.height", VH_SET_HEIGHT, 1, &out->height},
      {"DepthMapFactor", VH_SET_DEPTH_FACTOR, 0, &out->depth_factor},
      {"MaxDepth", VH_SET_MAX_DEPTH, 0, &out->max_depth},
      {"Voxel.Origin.x", 1u << 16, 0, &out->origin[0]},
      {"Voxel.Origin.y", 1u << 17, 0, &out->origin[1]},
      {"Voxel.Origin.z", 1u << 18, 0, &out->origin[2]},
      {"Voxel.Size", VH_SET_VOX_SIZE, 0, &out->vox_size},
      {"Voxel.TruncMargin", VH_SET_TRUNC, 0, &out->trunc},
      {"Voxel.Dim.x", 1u << 20, 1, &out->vox_dim[0]},
      {"Voxel.Dim.y", 1u << 21, 1, &out->vox_dim[1]},
      {"Voxel.Dim.z", 1u << 22, 1, &out->vox_dim[2]},
  };
  std::string line;
  int lineno = 0;
  uint32_t found = 0;
  while (std::getline(in, line)) {
    ++lineno;
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line = line.substr(0, hash);
    line = trim(line);
    if (line.empty() || line[0] == '%' || line == "---") continue;
    const size_t colon = line.find(':');
    if (colon == std::string::npos) continue;
    const std::string key = trim(line.substr(0, colon)), val = trim(line.substr(colon + 1));
    for (const Key& k : keys) {
      if (key != k.name) continue;
      double d;
      if (!parse_number(val, d))
        return fail(VH_ERR_INVALID_ARG, std::string(path) + ":" + std::to_string(lineno) + ": " + key +
                                            " is not a number");
      if (k.kind == 1) {
        if (d != std::floor(d) || std::fabs(d) > 2147483647.0)
          return fail(VH_ERR_INVALID_ARG, std::string(path) + ":" + std::to_string(lineno) + ": " + key +
                                              " is not an integer");
        *static_cast<int32_t*>(k.dst) = (int32_t)d;
      } else {
        *static_cast<float*>(k.dst) = (float)d;
      }
      if (k.bit >= (1u << 16) && k.bit < (1u << 20)) origin_bits |= k.bit;
      else if (k.bit >= (1u << 20)) dim_bits |= k.bit;
      else found |= k.bit;
    }
  }
  if (origin_bits == (7u << 16)) found |= VH_SET_ORIGIN;
  if (dim_bits == (7u << 20)) found |= VH_SET_VOX_DIM;
  out->found = found;
  return VH_OK;
}

int vh_settings_to_config(const vh_settings* s, vh_config* c) {
  if (!s || !c) return fail(VH_ERR_INVALID_ARG, "null argument");
  const uint32_t need = VH_SET_FX | VH_SET_FY | VH_SET_CX | VH_SET_CY | VH_SET_WIDTH | VH_SET_HEIGHT |
                        VH_SET_MAX_DEPTH | VH_SET_VOX_SIZE | VH_SET_TRUNC;
  if ((s->found & need) != need) return fail(VH_ERR_INVALID_ARG, "settings lack a camera, MaxDepth or voxel key");
  c->width = s->width;
  c->height = s->height;
  c->fx = s->fx;
  c->fy = s->fy;
  c->cx = s->cx;
  c->cy = s->cy;
  c->max_depth = s->max_depth;
  c->vox_size = s->vox_size;
  c->trunc = s->trunc;
  for (int k = 0; k < 3; ++k) {
    c->origin[k] = s->origin[k];
    c->vox_dim[k] = s->vox_dim[k];
  }
  return VH_OK;
}

int vh_png16_info(const char* path, int32_t* width, int32_t* height) {
  if (!path || !width || !height) return fail(VH_ERR_INVALID_ARG, "null argument");
  Png png;
  const int rc = parse_png(path, png, false);
  if (rc) return rc;
  *width = (int32_t)png.w;
  *height = (int32_t)png.h;
  return VH_OK;
}

int vh_png16_load(const char* path, uint16_t* px, int32_t width, int32_t height) {
  if (!path || !px || width <= 0 || height <= 0) return fail(VH_ERR_INVALID_ARG, "bad argument");
  Png png;
  int rc = parse_png(path, png, true);
  if (rc) return rc;
  if ((int32_t)png.w != width || (int32_t)png.h != height)
    return fail(VH_ERR_INVALID_ARG, std::string(path) + ": image is " + std::to_string(png.w) + "x" +
                                        std::to_string(png.h) + ", expected " + std::to_string(width) + "x" +
                                        std::to_string(height));
  const size_t stride = (size_t)png.w * 2;
  std::vector<unsigned char> raw((stride + 1) * png.h);
  uLongf raw_len = (uLongf)raw.size();
  const int zr = uncompress(raw.data(), &raw_len, png.idat.data(), (uLong)png.idat.size());
  if (zr != Z_OK || raw_len != raw.size())
    return fail(VH_ERR_INVALID_ARG, std::string(path) + ": corrupt image data (zlib " + std::to_string(zr) + ")");
  // undo the per-row filters (2 bytes per pixel: the left neighbour is 2 bytes back)
  std::vector<unsigned char> prev(stride, 0), cur(stride);
  for (uint32_t y = 0; y < png.h; ++y) {
    const unsigned char* row = &raw[y * (stride + 1)];
    const int ft = row[0];
    for (size_t i = 0; i < stride; ++i) {
      const int a = i >= 2 ? cur[i - 2] : 0, b = prev[i], c = i >= 2 ? prev[i - 2] : 0, x = row[1 + i];
      int v;
      switch (ft) {
        case 0: v = x; break;
        case 1: v = x + a; break;
        case 2: v = x + b; break;
        case 3: v = x + ((a + b) >> 1); break;
        case 4: v = x + paeth(a, b, c); break;
        default: return fail(VH_ERR_INVALID_ARG, std::string(path) + ": bad row filter");
      }
      cur[i] = (unsigned char)v;
    }
    for (uint32_t x = 0; x < png.w; ++x) px[(size_t)y * png.w + x] = (uint16_t)((cur[2 * x] << 8) | cur[2 * x + 1]);
    prev.swap(cur);
  }
  return VH_OK;
}

int vh_png16_save(const char* path, const uint16_t* px, int32_t width, int32_t height) {
  if (!path || !px || width <= 0 || height <= 0 || width > 65536 || height > 65536)
    return fail(VH_ERR_INVALID_ARG, "bad argument");
  const size_t stride = (size_t)width * 2;
  std::vector<unsigned char> raw((stride + 1) * (size_t)height);
  for (int32_t y = 0; y < height; ++y) {
    unsigned char* row = &raw[(size_t)y * (stride + 1)];
    row[0] = 0;
    for (int32_t x = 0; x < width; ++x) {
      const uint16_t v = px[(size_t)y * width + x];
      row[1 + 2 * x] = (unsigned char)(v >> 8);
      row[2 + 2 * x] = (unsigned char)v;
    }
  }
  uLongf zlen = compressBound((uLong)raw.size());
  std::vector<unsigned char> z(zlen);
  if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 6) != Z_OK)
    return fail(VH_ERR_IO, "zlib compression failed");
  z.resize(zlen);
  std::vector<unsigned char> out(kPngSig, kPngSig + 8);
  auto chunk = [&](const char* type, const unsigned char* data, size_t len) {
    put_be32(out, (uint32_t)len);
    const size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data, data + len);
    put_be32(out, (uint32_t)crc32(0, &out[start], (uInt)(4 + len)));
  };
  std::vector<unsigned char> ihdr;
  put_be32(ihdr, (uint32_t)width);
  put_be32(ihdr, (uint32_t)height);
  const unsigned char rest[5] = {16, 0, 0, 0, 0};
  ihdr.insert(ihdr.end(), rest, rest + 5);
  chunk("IHDR", ihdr.data(), ihdr.size());
  chunk("IDAT", z.data(), z.size());
  chunk("IEND", nullptr, 0);
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(VH_ERR_IO, std::string("cannot open ") + path);
  const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
  if ((std::fclose(f) != 0) || !ok) return fail(VH_ERR_IO, std::string("write failed: ") + path);
  return VH_OK;
}

int vh_pose_load(const char* path, float c2w[16]) {
  if (!path || !c2w) return fail(VH_ERR_INVALID_ARG, "null argument");
  std::ifstream in(path);
  if (!in) return fail(VH_ERR_IO, std::string("cannot read ") + path);
  std::string tok;
  for (int k = 0; k < 16; ++k) {
    double d;
    if (!(in >> tok) || !parse_number(tok, d))
      return fail(VH_ERR_INVALID_ARG, std::string(path) + ": expected 16 numbers");
    c2w[k] = (float)d;
  }
  return VH_OK;
}

int vh_frame_load(const char* dir, int32_t id, uint16_t* depth, int32_t width, int32_t height, float c2w[16]) {
  if (!dir || !depth || !c2w) return fail(VH_ERR_INVALID_ARG, "null argument");
  std::string base(dir);
  if (!base.empty() && base.back() != '/') base += '/';
  const std::string dpath = base + "depth/" + std::to_string(id) + ".png";
  FILE* probe = std::fopen(dpath.c_str(), "rb");
  if (!probe) return fail(VH_ERR_IO, "no frame " + std::to_string(id) + " (" + dpath + ")");
  std::fclose(probe);
  int rc = vh_png16_load(dpath.c_str(), depth, width, height);
  if (rc) return rc;
  return vh_pose_load((base + "tcw/" + std::to_string(id) + ".txt").c_str(), c2w);
}

}  // extern "C"
