// vh_oracle.cpp - CPU restatement of the reference's voxel-hashing TSDF path.
//
// TEST INFRASTRUCTURE ONLY. This library is the oracle of libvh's results: it is loaded by
// the tests, by the smoke check and by the CPU-baseline leg of the benchmark, and by nothing
// that ships. It shares no source with libvh; it restates the algorithm from the
// reference (citations below) with the documented corrections of the reference's defects,
// evaluating every FP32 operation in the same order as the GPU kernels so that an FP32
// (or FP16-storage) twin reproduces the GPU bit-for-bit, and carrying an FP64 shadow of
// the voxel values for accuracy measurements.
//
// Reference behaviour restated here:
//   lattice / block / in-block order          src/tsdf.cu:24-64
//   world<->camera transforms, projection      src/tsdf.cu:66-116 (base2cam, cam2base, cam2frame)
//   integration skip rules, truncation         src/tsdf.cu:599-751 (IntegrateHashKernel)
//   running average with the old weight        src/tsdf.cu:853-858 (dense Integrate: the intent;
//                                              evaluated as s + (dist - s) / (w + 1))
//   colour running average, 8-bit truncation   src/tsdf.cu:743-745, include/tsdf.cuh:122
//   block selection along the pixel ray        src/tsdf.cu:2088-2238 (re-specified: truncation band)
//   marching cubes corner order, sign rule     include/tsdf.cuh:234-241, src/tsdf.cu:978-1061
//   vertex interpolation with 1e-5 snaps       src/tsdf.cu:1640-1660 (position and colour)
//   Bourke triangle table                      include/tsdf.cuh:278-534
// Functions without a reference counterpart (trilinear query, its gradient, binary16
// rounding) are restated from their textbook definitions. Measurement helpers: marching
// cubes over the FP64 shadow (the FP64 reference mesh) and one-sided mesh-to-mesh distances
// (point to triangle: the closest-point construction of Ericson, Real-Time Collision
// Detection, 2005, sec. 5.1.5).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <array>
#include <cstring>
#include <map>
#include <tuple>
#include <type_traits>
#include <unordered_map>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ------------------------------------------------------------------ Bourke triTable
// One string per cube case: the triangle list as hex edge indices (3 per triangle).
const char* const kTri[256] = {
#include "mc_tri_strings.inc"
};
// corner offsets in the reference's order (include/tsdf.cuh:234-241)
const int kCorner[8][3] = {{0, 0, 0}, {0, 1, 0}, {1, 1, 0}, {1, 0, 0},
                           {0, 0, 1}, {0, 1, 1}, {1, 1, 1}, {1, 0, 1}};
// edge -> corner pair, as in the VertexInterp calls of src/tsdf.cu:1001-1037
const int kEdge[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                          {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

int hexval(char c) { return c <= '9' ? c - '0' : c - 'a' + 10; }

// ------------------------------------------------------------------ binary16 (RNE)
uint16_t to_half(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  const uint32_t sign = (u >> 31) << 15;
  const int exp = (int)((u >> 23) & 0xff);
  uint32_t man = u & 0x7fffff;
  if (exp == 0xff) return (uint16_t)(sign | 0x7c00 | (man ? 0x200 : 0));
  // value = 1.man * 2^(exp-127); half keeps 10 fraction bits at exponent >= -14
  int e = exp - 127;
  if (exp == 0) return (uint16_t)sign;  // float subnormals are far below half's range
  uint64_t sig = (uint64_t)(man | 0x800000);  // 24-bit significand, value = sig * 2^(e-23)
  int shift;                                  // bits to drop to reach half's last place
  uint32_t he;
  if (e >= -14) {
    shift = 13;
    he = (uint32_t)(e + 15);
  } else {
    shift = 13 + (-14 - e);
    he = 0;
  }
  if (shift > 40) return (uint16_t)sign;
  uint64_t q = sig >> shift;
  const uint64_t rem = sig & ((1ull << shift) - 1);
  const uint64_t halfway = 1ull << (shift - 1);
  if (rem > halfway || (rem == halfway && (q & 1))) ++q;
  uint32_t bits;
  if (he == 0) {
    bits = (uint32_t)q;  // subnormal; a carry into 0x400 is the smallest normal, correctly
  } else {
    // q has 11 significant bits (1.xxx); a carry to 0x800 bumps the exponent
    if (q == 0x800) {
      q = 0x400;
      ++he;
    }
    if (he >= 31) return (uint16_t)(sign | 0x7c00);
    bits = (he << 10) | (uint32_t)(q & 0x3ff);
  }
  return (uint16_t)(sign | bits);
}

float from_half(uint16_t h) {
  const int sign = (h >> 15) & 1, e = (h >> 10) & 31, m = h & 0x3ff;
  double v;
  if (e == 0) v = std::ldexp((double)m, -24);
  else if (e == 31) v = m ? NAN : INFINITY;
  else v = std::ldexp((double)(m | 0x400), e - 25);
  return (float)(sign ? -v : v);
}

// Compensated binary16: value = h + r * ulp(h)/256, r an int8 residual of the rounding.
double comp_scale(uint16_t h) {
  const int e = (h >> 10) & 31;
  return std::ldexp(1.0, e == 0 ? -32 : e - 33);
}
float decode_comp(uint16_t h, int8_t r) { return (float)((double)from_half(h) + (double)r * comp_scale(h)); }
void encode_comp(float s, uint16_t& h, int8_t& r) {
  h = to_half(s);
  // the nearest integer, ties to even (the quotient is exact: s - h exactly, a power-of-two scale)
  const double q = std::nearbyint(((double)s - (double)from_half(h)) / comp_scale(h));
  r = (int8_t)std::max(-128.0, std::min(127.0, q));
}

// ------------------------------------------------------------------ keys
inline int64_t key_of(int x, int y, int z) {
  return ((int64_t)(x + (1 << 20)) << 42) | ((int64_t)(y + (1 << 20)) << 21) | (int64_t)(z + (1 << 20));
}
inline void key_xyz(int64_t k, int& x, int& y, int& z) {
  x = (int)((k >> 42) & 0x1fffff) - (1 << 20);
  y = (int)((k >> 21) & 0x1fffff) - (1 << 20);
  z = (int)(k & 0x1fffff) - (1 << 20);
}
inline int floor_div8(int v) { return v >= 0 ? v / 8 : -((-v + 7) / 8); }

struct Oracle {
  int W, H;
  float fx, fy, cx, cy, max_depth, min_depth, vox, mu;
  int storage;  // 0: FP32 twin, 1: FP16 twin, 2: compensated-FP16 twin
  int cap;      // 0 = uncapped
  std::unordered_map<int64_t, int64_t> index;  // key -> block
  std::vector<int64_t> keys;                   // block -> key
  std::vector<float> s32, w32;                 // FP32 twin
  std::vector<uint16_t> s16, w16;              // FP16 twins (binary16 bits, weight)
  std::vector<int8_t> r8;                      // compensated FP16: rounding residual
  std::vector<double> s64, w64;                // FP64 shadow
  bool color = false;
  int carve = 0;  // 1: every allocated block may_observe keeps is a candidate each frame; 2: every one
  int shadow = 1;  // 0: the FP64 shadow is not updated (times the FP32 twin alone)
  std::vector<uint8_t> rgb8;                   // colour volumes: 3 channels per voxel
  std::vector<int64_t> touched;                // last frame, sorted keys
  int64_t frames = 0;

  int64_t add_block(int64_t k) {
    const int64_t b = (int64_t)keys.size();
    keys.push_back(k);
    index.emplace(k, b);
    s32.resize(s32.size() + 512, 0.0f);
    w32.resize(w32.size() + 512, 0.0f);
    s16.resize(s16.size() + 512, 0);
    w16.resize(w16.size() + 512, 0);
    r8.resize(r8.size() + 512, 0);
    s64.resize(s64.size() + 512, 0.0);
    w64.resize(w64.size() + 512, 0.0);
    if (color) rgb8.resize(rgb8.size() + 512 * 3, 0);
    return b;
  }
  int64_t find(int64_t k) const {
    auto it = index.find(k);
    return it == index.end() ? -1 : it->second;
  }
  // twin voxel value (FP32 or widened FP16) and weight
  void twin(int64_t b, int i, float& s, float& w) const {
    const size_t v = (size_t)b * 512 + i;
    if (storage == 2) {
      s = decode_comp(s16[v], r8[v]);
      w = (float)w16[v];
    } else if (storage == 1) {
      s = from_half(s16[v]);
      w = (float)w16[v];
    } else {
      s = s32[v];
      w = w32[v];
    }
  }
};

// ------------------------------------------------------------------ band traversal
// Blocks crossed by the segment of pixel (u,v) between camera depths max(d-mu, min_depth)
// and d+mu. Camera point: (rx*t, ry*t, t) with the ray per unit depth rx = (u-cx)*RN(1/fx),
// ry = (v-cy)*RN(1/fy) (frame2cam, src/tsdf.cu:67-73, with the divisions by the focal lengths
// taken as products by their rounded reciprocals); world point R p + t as cam2base
// (src/tsdf.cu:95-105); block coordinates world * RN(1/(8 vox)).
// Walk: from floor(start) to floor(end), stepping the axis whose next face is crossed
// first among the axes that still have to move (lowest axis on ties); an axis's crossing
// parameter is (face - start) * RN(1/direction) (the reciprocal taken once per pixel).
template <class Emit>
void band_blocks(const Oracle& o, const float* T, int u, int v, float d, Emit emit) {
  float t0 = d - o.mu;
  if (t0 < o.min_depth) t0 = o.min_depth;
  const float t1 = d + o.mu;
  const float ibsz = 1.0f / (o.vox * 8.0f), ifx = 1.0f / o.fx, ify = 1.0f / o.fy;
  const float rx = ((float)u - o.cx) * ifx, ry = ((float)v - o.cy) * ify;
  float cam[2][3];
  const float ts[2] = {t0, t1};
  for (int e = 0; e < 2; ++e) {
    cam[e][0] = rx * ts[e];
    cam[e][1] = ry * ts[e];
    cam[e][2] = ts[e];
  }
  float p[2][3];
  for (int e = 0; e < 2; ++e)
    for (int r = 0; r < 3; ++r)
      p[e][r] = (cam[e][0] * T[4 * r] + cam[e][1] * T[4 * r + 1] + cam[e][2] * T[4 * r + 2] + T[4 * r + 3]) * ibsz;
  for (int e = 0; e < 2; ++e)
    for (int r = 0; r < 3; ++r)
      if (!(std::fabs(p[e][r]) < 1048000.0f)) return;
  int cur[3], end[3];
  float rdir[3];  // 1/direction, 2^64 with its sign outside 2^-64 <= |direction| <= 2^64
  for (int r = 0; r < 3; ++r) {
    cur[r] = (int)std::floor(p[0][r]);
    end[r] = (int)std::floor(p[1][r]);
    const float dir = p[1][r] - p[0][r], m = std::fabs(dir);
    rdir[r] = (m >= 0x1p-64f && m <= 0x1p64f) ? 1.0f / dir : std::copysign(0x1p64f, dir);
  }
  emit(cur[0], cur[1], cur[2]);
  while (cur[0] != end[0] || cur[1] != end[1] || cur[2] != end[2]) {
    int best = -1;
    float bt = 0.0f;
    for (int r = 0; r < 3; ++r) {
      if (cur[r] == end[r]) continue;
      const float face = (float)(end[r] > cur[r] ? cur[r] + 1 : cur[r]);
      const float t = (face - p[0][r]) * rdir[r];
      if (best < 0 || t < bt) {
        best = r;
        bt = t;
      }
    }
    cur[best] += end[best] > cur[best] ? 1 : -1;
    emit(cur[0], cur[1], cur[2]);
  }
}

// ------------------------------------------------------------------ integration
// Carving (every allocated voxel the frame observes is updated; the reference integrates every
// existing block found along each sampled pixel ray from 0.1 m to max_depth inside its frustum
// test, src/tsdf.cu:2128-2177, :2056-2064, so free space in front of the surface is updated).
// A cheap bound first: false only if no voxel of the block can pass the observation test of
// integrate() below (z >= 2^-64 and a rounded pixel inside the image, then d - z > -mu for a
// depth d <= max_depth). Exact FP64 camera coordinates of the 8 corners of the block's lattice
// box bound every voxel's (the map is affine, and for a box wholly in front of the camera the
// image of the box is the hull of its corners' images); the margins are far wider than the
// rounding of the FP32 voxel arithmetic: 1 mm + 1e-5 of the magnitudes in depth, and in the
// image 1 pixel plus 4e-6 of the magnitudes over the nearest depth times (f + |u - c|).
bool may_observe(const Oracle& o, const double* T, int64_t key) {
  int b[3];
  key_xyz(key, b[0], b[1], b[2]);
  double zmin = INFINITY, zmax = -INFINITY, mag = 0.0, P[8][3];
  for (int c = 0; c < 8; ++c) {
    double d[3];
    for (int k = 0; k < 3; ++k) d[k] = ((double)b[k] * 8.0 + ((c >> k) & 1) * 7.0) * (double)o.vox - T[4 * k + 3];
    for (int r = 0; r < 3; ++r) P[c][r] = T[r] * d[0] + T[4 + r] * d[1] + T[8 + r] * d[2];
    mag = std::max(mag, std::fabs(d[0]) + std::fabs(d[1]) + std::fabs(d[2]));
    zmin = std::min(zmin, P[c][2]);
    zmax = std::max(zmax, P[c][2]);
  }
  const double dz = 1e-3 + 1e-5 * mag;
  if (zmax < -dz || zmin > (double)o.max_depth + (double)o.mu + dz) return false;
  if (!(zmin > dz)) return true;
  double umin = INFINITY, umax = -INFINITY, vmin = INFINITY, vmax = -INFINITY;
  for (int c = 0; c < 8; ++c) {
    const double u = o.fx * P[c][0] / P[c][2] + o.cx, v = o.fy * P[c][1] / P[c][2] + o.cy;
    umin = std::min(umin, u); umax = std::max(umax, u);
    vmin = std::min(vmin, v); vmax = std::max(vmax, v);
  }
  const double rel = 4e-6 * mag / zmin;
  const double eu = 1.0 + rel * (o.fx + std::max(std::fabs(umin - o.cx), std::fabs(umax - o.cx)));
  const double ev = 1.0 + rel * (o.fy + std::max(std::fabs(vmin - o.cy), std::fabs(vmax - o.cy)));
  return !(umax < -0.5 - eu || umin > o.W - 0.5 + eu || vmax < -0.5 - ev || vmin > o.H - 0.5 + ev);
}

// Colour of one observation (src/tsdf.cu:743-745): per channel (c * w_old + rgb) / (w_old + 1)
// in FP32 (product and sum exact below 2^24), truncated to 8 bits; w_old is the stored weight
// before the update.
void color_update(uint8_t* c, float w_old, const uint8_t* px) {
  const float wn = w_old + 1.0f;
  for (int k = 0; k < 3; ++k) c[k] = (uint8_t)(std::fmaf((float)c[k], w_old, (float)px[k]) / wn);
}

void integrate(Oracle& o, const uint16_t* dep, float scale, const float* T, const uint8_t* rgb) {
  const int W = o.W, H = o.H;
  std::vector<float> d((size_t)W * H);
  for (size_t i = 0; i < d.size(); ++i) d[i] = (float)dep[i] * scale;
  // pre-scaled depth d / mu of each pixel (FP32 product by the rounded 1/mu), -inf where the
  // pixel has no valid measurement: the observation test below is then one comparison
  const float inv_mu = 1.0f / o.mu;
  std::vector<float> dm(d.size());
  for (size_t i = 0; i < d.size(); ++i) dm[i] = (d[i] > 0.0f && d[i] <= o.max_depth) ? d[i] * inv_mu : -INFINITY;

  // 1. touched set
  int nth = 1;
#ifdef _OPENMP
  nth = omp_get_max_threads();
#endif
  std::vector<std::vector<int64_t>> local((size_t)nth);
#pragma omp parallel for schedule(dynamic, 8)
  for (int v = 0; v < H; ++v) {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    auto& out = local[(size_t)tid];
    for (int u = 0; u < W; ++u) {
      const float dv = d[(size_t)v * W + u];
      if (!(dv > 0.0f && dv <= o.max_depth)) continue;
      band_blocks(o, T, u, v, dv, [&](int x, int y, int z) { out.push_back(key_of(x, y, z)); });
    }
  }
  std::vector<int64_t> t;
  for (auto& l : local) t.insert(t.end(), l.begin(), l.end());
  std::sort(t.begin(), t.end());
  t.erase(std::unique(t.begin(), t.end()), t.end());
  o.touched = t;
  std::vector<int64_t> blk(t.size());
  for (size_t k = 0; k < t.size(); ++k) {
    int64_t b = o.find(t[k]);
    blk[k] = b >= 0 ? b : o.add_block(t[k]);
  }

  // 2. voxels of the touched blocks (carving: of every allocated block the frame may observe,
  //    the ones just allocated included)
  const double T64[12] = {T[0], T[1], T[2], T[3], T[4], T[5], T[6], T[7], T[8], T[9], T[10], T[11]};
  if (o.carve) {
    const long nb = (long)o.keys.size();
    std::vector<uint8_t> keep((size_t)nb);
#pragma omp parallel for schedule(static, 4096)
    for (long b = 0; b < nb; ++b) keep[(size_t)b] = (o.carve == 2 || may_observe(o, T64, o.keys[(size_t)b])) ? 1 : 0;
    blk.clear();
    for (long b = 0; b < nb; ++b)
      if (keep[(size_t)b]) blk.push_back(b);
  }
  const double vox64 = o.vox, mu64 = o.mu, scale64 = scale;
  const float capf = (float)o.cap;
  // homogeneous voxel steps of the frame: the camera step of each lattice axis, R^T e vox (a
  // row of the pose times vox), with the intrinsics folded in as (fx x + cx z, fy y + cy z, z)
  auto hom = [&](const float c[3], float out[3]) {
    out[0] = std::fmaf(o.fx, c[0], o.cx * c[2]);
    out[1] = std::fmaf(o.fy, c[1], o.cy * c[2]);
    out[2] = c[2];
  };
  float HS[3][3];
  for (int ax = 0; ax < 3; ++ax) {
    const float c[3] = {T[4 * ax] * o.vox, T[4 * ax + 1] * o.vox, T[4 * ax + 2] * o.vox};
    hom(c, HS[ax]);
  }
#pragma omp parallel for schedule(dynamic, 4)
  for (long k = 0; k < (long)blk.size(); ++k) {
    int bx, by, bz;
    key_xyz(o.keys[(size_t)blk[(size_t)k]], bx, by, bz);
    const size_t base = (size_t)blk[(size_t)k] * 512;
    float cb[3], hb[3];  // camera coordinates of the block origin, R^T (o - t), and homogeneous
    {
      const float d0 = (float)(bx * 8) * o.vox - T[3], d1 = (float)(by * 8) * o.vox - T[7],
                  d2 = (float)(bz * 8) * o.vox - T[11];
      for (int r = 0; r < 3; ++r) cb[r] = T[r] * d0 + T[4 + r] * d1 + T[8 + r] * d2;
    }
    hom(cb, hb);
    for (int i = 0; i < 512; ++i) {
      const int gx = bx * 8 + (i >> 6), gy = by * 8 + ((i >> 3) & 7), gz = bz * 8 + (i & 7);
      // FP32 geometry: decides observation, pixel and the FP32 distance. The voxel's
      // homogeneous pixel point is formed per block origin and voxel step, H(R^T(o - t)) +
      // H(vox R^T (k, j, i)): the column offset as fma(j, b, i*c) added to the origin, then
      // fma(k, a, column) per voxel; one reciprocal of z feeds both image coordinates X/z, Y/z,
      // each rounded to the nearest pixel (one rounding, ties to even); the truncated distance is
      // fma(-z, 1/mu, d/mu), the voxel observed iff it exceeds -1. (The reference rotates each
      // voxel's p - t, divides x/z, y/z and by mu, and rounds ties away from zero:
      // src/tsdf.cu:75-93,738, compiled with -use_fast_math.)
      const int vx = i >> 6, vy = (i >> 3) & 7, vz = i & 7;
      float P[3];
      for (int r = 0; r < 3; ++r) {
        const float col = hb[r] + std::fmaf((float)vy, HS[1][r], (float)vz * HS[2][r]);
        P[r] = std::fmaf((float)vx, HS[0][r], col);
      }
      const float zc = P[2];
      if (!(zc >= 0x1p-64f && zc <= 0x1p64f)) continue;  // depth range of the projection
      // the image coordinates: X * (1/z) rounded once to the nearest integer (ties to even)
      // by adding M = 1.5 * 2^23 in the same fused multiply-add, fma(X, 1/z, M) - M
      const float rz = 1.0f / zc;
      constexpr float M = 0x1.8p23f;
      const float uf = std::fmaf(P[0], rz, M);
      const float vf = std::fmaf(P[1], rz, M);
      if (!(uf >= M && uf <= M + (float)(W - 1) && vf >= M && vf <= M + (float)(H - 1))) continue;
      const size_t pix = (size_t)(vf - M) * W + (size_t)(uf - M);
      const float t = std::fmaf(-zc, inv_mu, dm[pix]);
      if (!(t > -1.0f)) continue;
      const float dist = std::fmin(1.0f, t);
      if (o.color) {
        float s0, w0;
        o.twin(blk[(size_t)k], i, s0, w0);
        color_update(&o.rgb8[3 * (base + i)], w0, rgb + 3 * pix);
      }
      // twin update: mean with the old weight, s + (dist - s) * RN(1 / (w + 1)), then the cap
      if (o.storage == 2) {
        float s = decode_comp(o.s16[base + i], o.r8[base + i]), w = (float)o.w16[base + i];
        const float wn = w + 1.0f;
        s = std::fmaf(dist - s, 1.0f / wn, s);
        w = (capf > 0.0f && wn > capf) ? capf : wn;
        encode_comp(s, o.s16[base + i], o.r8[base + i]);
        o.w16[base + i] = (uint16_t)w;
      } else if (o.storage == 1) {
        float s = from_half(o.s16[base + i]), w = (float)o.w16[base + i];
        const float wn = w + 1.0f;
        s = std::fmaf(dist - s, 1.0f / wn, s);
        w = (capf > 0.0f && wn > capf) ? capf : wn;
        o.s16[base + i] = to_half(s);
        o.w16[base + i] = (uint16_t)w;
      } else {
        float s = o.s32[base + i], w = o.w32[base + i];
        const float wn = w + 1.0f;
        s = std::fmaf(dist - s, 1.0f / wn, s);
        w = (capf > 0.0f && wn > capf) ? capf : wn;
        o.s32[base + i] = s;
        o.w32[base + i] = w;
      }
      // FP64 shadow: same observation, distance and mean in double
      if (!o.shadow) continue;
      const double X = gx * vox64 - T64[3], Y = gy * vox64 - T64[7], Z = gz * vox64 - T64[11];
      const double z64 = T64[2] * X + T64[6] * Y + T64[10] * Z;
      const double d64 = (double)dep[pix] * scale64;
      const double dist64 = std::min(1.0, (d64 - z64) / mu64);
      double s = o.s64[base + i], w = o.w64[base + i];
      const double wn = w + 1.0;
      s = (s * w + dist64) / wn;
      w = (o.cap > 0 && wn > o.cap) ? (double)o.cap : wn;
      o.s64[base + i] = s;
      o.w64[base + i] = w;
    }
  }
  ++o.frames;
}

// ------------------------------------------------------------------ marching cubes
struct Tri {
  int32_t c[4];
  float p[9];
  uint8_t rgb[9];
};

// VertexInterp(iso = 0) in voxel units (src/tsdf.cu:1640-1660); |x| < 1e-5 compares the
// float magnitude against the double constant, as the reference does.
// Colour (ca, cb, cout: 3 channels each, may be null) as the reference's
// p.r = p1.r + mu*(p2.r - p1.r) (src/tsdf.cu:1655-1657): the channel difference is an int,
// the rest FP32, the result truncated to 8 bits; a snapped vertex takes its corner's colour.
void interp(const float a[3], const float b[3], float va, float vb, float out[3], const uint8_t* ca = nullptr,
            const uint8_t* cb = nullptr, uint8_t* cout = nullptr) {
  auto take = [&](const float* p, const uint8_t* c) {
    std::memcpy(out, p, 12);
    if (cout) std::memcpy(cout, c, 3);
  };
  if ((double)std::fabs(va) < 0.00001) return take(a, ca);
  if ((double)std::fabs(vb) < 0.00001) return take(b, cb);
  if ((double)std::fabs(va - vb) < 0.00001) return take(a, ca);
  const float m = (0.0f - va) / (vb - va);
  for (int k = 0; k < 3; ++k) out[k] = a[k] + m * (b[k] - a[k]);
  if (cout)
    for (int k = 0; k < 3; ++k) cout[k] = (uint8_t)((float)ca[k] + m * (float)((int)cb[k] - (int)ca[k]));
}

// VertexInterp in FP64 (same snaps) for the FP64 reference mesh.
void interp64(const double a[3], const double b[3], double va, double vb, double out[3]) {
  if (std::fabs(va) < 0.00001) { std::memcpy(out, a, 24); return; }
  if (std::fabs(vb) < 0.00001) { std::memcpy(out, b, 24); return; }
  if (std::fabs(va - vb) < 0.00001) { std::memcpy(out, a, 24); return; }
  const double m = (0.0 - va) / (vb - va);
  for (int k = 0; k < 3; ++k) out[k] = a[k] + m * (b[k] - a[k]);
}

// k64 = false: the twin's values, FP32 interpolation (the GPU's mesh); true: the FP64
// shadow's values and FP64 interpolation, positions rounded to FP32 once at the end.
// keep: also the zero-area triangles (the indexed mesh's faces).
template <bool k64>
void cube_triangles(const Oracle& o, int gx, int gy, int gz, std::vector<Tri>& out, bool keep = false) {
  using R = typename std::conditional<k64, double, float>::type;
  R val[8];
  uint8_t col[8][3] = {};
  for (int c = 0; c < 8; ++c) {
    const int x = gx + kCorner[c][0], y = gy + kCorner[c][1], z = gz + kCorner[c][2];
    const int64_t b = o.find(key_of(floor_div8(x), floor_div8(y), floor_div8(z)));
    if (b < 0) return;
    const int i = ((x - floor_div8(x) * 8) * 8 + (y - floor_div8(y) * 8)) * 8 + (z - floor_div8(z) * 8);
    if (k64) {
      const size_t v = (size_t)b * 512 + i;
      if (!(o.w64[v] > 0.0)) return;
      val[c] = (R)o.s64[v];
      continue;
    }
    float s, w;
    o.twin(b, i, s, w);
    if (!(w > 0.0f)) return;  // unobserved corner: no surface (corrects the reference, which has no test)
    val[c] = s;
    if (o.color) std::memcpy(col[c], &o.rgb8[3 * ((size_t)b * 512 + i)], 3);
  }
  int ci = 0;
  for (int c = 0; c < 8; ++c)
    if (val[c] < (R)0) ci |= 1 << c;
  const char* tl = kTri[ci];
  const int n = (int)std::strlen(tl) / 3;
  for (int t = 0; t < n; ++t) {
    Tri tr;
    double q[9];
    for (int j = 0; j < 3; ++j) {
      const int e = hexval(tl[3 * t + j]);
      const int ca = kEdge[e][0], cb = kEdge[e][1];
      if (k64) {
        const double pa[3] = {(double)(gx + kCorner[ca][0]), (double)(gy + kCorner[ca][1]), (double)(gz + kCorner[ca][2])};
        const double pb[3] = {(double)(gx + kCorner[cb][0]), (double)(gy + kCorner[cb][1]), (double)(gz + kCorner[cb][2])};
        interp64(pa, pb, (double)val[ca], (double)val[cb], q + 3 * j);
      } else {
        const float pa[3] = {(float)(gx + kCorner[ca][0]), (float)(gy + kCorner[ca][1]), (float)(gz + kCorner[ca][2])};
        const float pb[3] = {(float)(gx + kCorner[cb][0]), (float)(gy + kCorner[cb][1]), (float)(gz + kCorner[cb][2])};
        interp(pa, pb, (float)val[ca], (float)val[cb], tr.p + 3 * j, col[ca], col[cb], tr.rgb + 3 * j);
      }
    }
    if (k64)
      for (int k = 0; k < 9; ++k) tr.p[k] = (float)q[k];  // only for the degenerate test below
    auto same = [&](int a, int b) {
      if (k64) return q[3 * a] == q[3 * b] && q[3 * a + 1] == q[3 * b + 1] && q[3 * a + 2] == q[3 * b + 2];
      return tr.p[3 * a] == tr.p[3 * b] && tr.p[3 * a + 1] == tr.p[3 * b + 1] && tr.p[3 * a + 2] == tr.p[3 * b + 2];
    };
    if (!keep && (same(0, 1) || same(1, 2) || same(2, 0))) continue;  // degenerate (typo-free :1055-1058)
    if (k64) {
      for (int k = 0; k < 9; ++k) tr.p[k] = (float)(q[k] * (double)o.vox);
      std::memset(tr.rgb, 0, 9);
    } else {
      for (int k = 0; k < 9; ++k) tr.p[k] *= o.vox;
    }
    tr.c[0] = gx; tr.c[1] = gy; tr.c[2] = gz; tr.c[3] = (ci << 8) | t;
    out.push_back(tr);
  }
}

// ------------------------------------------------------------------ mesh distance
// Squared distance from p to triangle (a, b, c): the closest point found by the region test
// of Ericson (2005) sec. 5.1.5 (vertex regions, edge regions, then the face interior).
double point_triangle_dist2(const double p[3], const double a[3], const double b[3], const double c[3]) {
  auto sub = [](const double* x, const double* y, double* o) { for (int k = 0; k < 3; ++k) o[k] = x[k] - y[k]; };
  auto dot = [](const double* x, const double* y) { return x[0] * y[0] + x[1] * y[1] + x[2] * y[2]; };
  double ab[3], ac[3], ap[3], bp[3], cp[3], q[3];
  sub(b, a, ab);
  sub(c, a, ac);
  sub(p, a, ap);
  auto at = [&](double v, double w) {
    for (int k = 0; k < 3; ++k) q[k] = a[k] + v * ab[k] + w * ac[k];
    double d[3];
    sub(p, q, d);
    return dot(d, d);
  };
  const double d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0 && d2 <= 0.0) return at(0.0, 0.0);
  sub(p, b, bp);
  const double d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0 && d4 <= d3) return at(1.0, 0.0);
  const double vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) return at(d1 / (d1 - d3), 0.0);
  sub(p, c, cp);
  const double d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0 && d5 <= d6) return at(0.0, 1.0);
  const double vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) return at(0.0, d2 / (d2 - d6));
  const double va = d3 * d6 - d5 * d4;
  if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
    const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    return at(1.0 - w, w);  // on bc: b + w (c - b)
  }
  const double den = 1.0 / (va + vb + vc);
  return at(vb * den, vc * den);
}

// ------------------------------------------------------------------ trilinear query
// The lattice position q/vox and the cell fraction are formed in FP64 for both precisions;
// R is the precision of the blend (float: twin of the GPU kernels; double: FP64 reference).
template <class R>
bool sample(const Oracle& o, double qx, double qy, double qz, R& sdf, R g[3]) {
  const R vox = (R)o.vox;
  const double vd = (double)o.vox;
  const double gq[3] = {qx / vd, qy / vd, qz / vd};
  int ix[3];
  R f[3];
  for (int k = 0; k < 3; ++k) {
    if (!(std::fabs(gq[k]) < 8.0e6)) return false;
    const double fl = std::floor(gq[k]);
    ix[k] = (int)fl;
    f[k] = (R)(gq[k] - fl);
  }
  R s[8];
  for (int k = 0; k < 8; ++k) {
    const int x = ix[0] + (k >> 2), y = ix[1] + ((k >> 1) & 1), z = ix[2] + (k & 1);
    const int64_t b = o.find(key_of(floor_div8(x), floor_div8(y), floor_div8(z)));
    if (b < 0) return false;
    const int i = ((x & 7) * 8 + (y & 7)) * 8 + (z & 7);
    float sv, wv;
    o.twin(b, i, sv, wv);
    if (!(wv > 0.0f)) return false;
    s[k] = (R)sv;
  }
  const R ax[2] = {(R)1 - f[0], f[0]}, ay[2] = {(R)1 - f[1], f[1]}, az[2] = {(R)1 - f[2], f[2]};
  R acc = 0, gx = 0, gy = 0, gz = 0;
  for (int k = 0; k < 8; ++k) {
    const int dx = k >> 2, dy = (k >> 1) & 1, dz = k & 1;
    const R w = (ax[dx] * ay[dy]) * az[dz];
    acc = acc + w * s[k];
  }
  for (int k = 0; k < 8; ++k) {
    const int dx = k >> 2, dy = (k >> 1) & 1, dz = k & 1;
    const R sx = dx ? (R)1 : (R)-1, sy = dy ? (R)1 : (R)-1, sz = dz ? (R)1 : (R)-1;
    gx = gx + ((sx * ay[dy]) * az[dz]) * s[k];
    gy = gy + ((ax[dx] * sy) * az[dz]) * s[k];
    gz = gz + ((ax[dx] * ay[dy]) * sz) * s[k];
  }
  sdf = acc;
  g[0] = gx / vox;
  g[1] = gy / vox;
  g[2] = gz / vox;
  return true;
}

}  // namespace

extern "C" {

void* vho_create(int W, int H, float fx, float fy, float cx, float cy, float max_depth, float min_depth, float vox,
                 float mu, int storage, int cap) {
  Oracle* o = new Oracle();
  o->W = W; o->H = H; o->fx = fx; o->fy = fy; o->cx = cx; o->cy = cy;
  o->max_depth = max_depth; o->min_depth = min_depth; o->vox = vox; o->mu = mu;
  o->storage = storage; o->cap = cap;
  return o;
}
void vho_destroy(void* p) { delete static_cast<Oracle*>(p); }
void vho_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}
int vho_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

int64_t vho_integrate(void* p, const uint16_t* depth, float scale, const float* c2w) {
  Oracle* o = static_cast<Oracle*>(p);
  if (o->color) return -1;  // a colour volume needs vho_integrate_rgb
  integrate(*o, depth, scale, c2w, nullptr);
  return (int64_t)o->touched.size();
}
// Makes the oracle a colour volume (before the first frame; returns 0, or -1 once blocks exist).
int vho_set_color(void* p, int on) {
  Oracle* o = static_cast<Oracle*>(p);
  if (!o->keys.empty()) return -1;
  o->color = on != 0;
  return 0;
}
// Depth frame with its RGB image (H x W x 3 bytes).
int64_t vho_integrate_rgb(void* p, const uint16_t* depth, const uint8_t* rgb, float scale, const float* c2w) {
  Oracle* o = static_cast<Oracle*>(p);
  if (!o->color || !rgb) return -1;
  integrate(*o, depth, scale, c2w, rgb);
  return (int64_t)o->touched.size();
}
// Colours of all blocks in insertion order (512 x 3 bytes per block).
int64_t vho_export_colors(void* p, uint8_t* rgb, int64_t cap) {
  Oracle* o = static_cast<Oracle*>(p);
  const int64_t n = (int64_t)o->keys.size();
  if (!o->color || cap < n) return n;
  if (rgb) std::memcpy(rgb, o->rgb8.data(), (size_t)n * 512 * 3);
  return n;
}
int64_t vho_num_blocks(void* p) { return (int64_t)static_cast<Oracle*>(p)->keys.size(); }
// FP64 shadow kept (1, the default) or not (0: the FP32 twin alone, for timing).
void vho_set_shadow(void* p, int on) { static_cast<Oracle*>(p)->shadow = on ? 1 : 0; }
// Carving on (1) or off (0): with carving every allocated voxel a frame observes is updated.
// 2: carving with every allocated block tested voxel by voxel (no may_observe bound; the
// brute-force check of that bound).
void vho_set_carve(void* p, int mode) { static_cast<Oracle*>(p)->carve = (mode >= 0 && mode <= 2) ? mode : 0; }

// Last frame's touched blocks, sorted by packed key; returns the count.
int64_t vho_touched(void* p, int32_t* xyz, int64_t cap) {
  Oracle* o = static_cast<Oracle*>(p);
  const int64_t n = (int64_t)o->touched.size();
  if (xyz && cap >= n)
    for (int64_t k = 0; k < n; ++k) key_xyz(o->touched[(size_t)k], xyz[3 * k], xyz[3 * k + 1], xyz[3 * k + 2]);
  return n;
}

// All blocks in insertion order: keys, twin values (FP16 widened) and the FP64 shadow.
int64_t vho_export(void* p, int32_t* xyz, float* s, float* w, double* s64, double* w64, int64_t cap) {
  Oracle* o = static_cast<Oracle*>(p);
  const int64_t n = (int64_t)o->keys.size();
  if (cap < n) return n;
  for (int64_t b = 0; b < n; ++b) {
    if (xyz) key_xyz(o->keys[(size_t)b], xyz[3 * b], xyz[3 * b + 1], xyz[3 * b + 2]);
    for (int i = 0; i < 512; ++i) {
      const size_t v = (size_t)b * 512 + i;
      float sv, wv;
      o->twin(b, i, sv, wv);
      if (s) s[v] = sv;
      if (w) w[v] = wv;
      if (s64) s64[v] = o->s64[v];
      if (w64) w64[v] = o->w64[v];
    }
  }
  return n;
}

// Marching cubes over every block; triangles ordered by cube then by table position.
extern "C++" template <bool k64>
int64_t mc_all(void* p, float* xyz, int32_t* cubes, uint8_t* rgb, int64_t cap, bool keep = false) {
  Oracle* o = static_cast<Oracle*>(p);
  std::vector<int64_t> order(o->keys.begin(), o->keys.end());
  std::sort(order.begin(), order.end());
  std::vector<std::vector<Tri>> per((size_t)order.size());
#pragma omp parallel for schedule(dynamic, 16)
  for (long k = 0; k < (long)order.size(); ++k) {
    int bx, by, bz;
    key_xyz(order[(size_t)k], bx, by, bz);
    for (int i = 0; i < 512; ++i)
      cube_triangles<k64>(*o, bx * 8 + (i >> 6), by * 8 + ((i >> 3) & 7), bz * 8 + (i & 7), per[(size_t)k], keep);
  }
  int64_t n = 0;
  for (auto& v : per) n += (int64_t)v.size();
  if (cap < n) return n;
  int64_t m = 0;
  for (auto& v : per)
    for (const Tri& t : v) {
      if (xyz) std::memcpy(xyz + 9 * m, t.p, 36);
      if (cubes) std::memcpy(cubes + 4 * m, t.c, 16);
      if (rgb) std::memcpy(rgb + 9 * m, t.rgb, 9);
      ++m;
    }
  return n;
}
int64_t vho_mc_rgb(void* p, float* xyz, int32_t* cubes, uint8_t* rgb, int64_t cap) {
  return mc_all<false>(p, xyz, cubes, rgb, cap);
}
int64_t vho_mc(void* p, float* xyz, int32_t* cubes, int64_t cap) { return mc_all<false>(p, xyz, cubes, nullptr, cap); }
// The FP64 reference mesh: marching cubes over the FP64 shadow (same cube records).
int64_t vho_mc64(void* p, float* xyz, int32_t* cubes, int64_t cap) { return mc_all<true>(p, xyz, cubes, nullptr, cap); }

// Indexed mesh of the twin's surface: the faces are every table triangle of every cube whose
// eight corners are observed, in the order of vho_mc, the zero-area ones included (the soup
// of vho_mc drops those; where the surface passes exactly through a lattice point, dropping
// them would open the edge-shared mesh); one vertex per lattice edge used by a face, numbered
// in order of first use, interpolated from the edge's lower lattice point to its upper one
// with the twin's values (VertexInterp with that orientation). Returns the vertex count and
// the face count in *n_faces; arrays are written when their capacities suffice (vertices:
// vcap x 3 floats in metres, rgb: vcap x 3 bytes, faces: fcap x 3 ints, face_cubes: fcap x 4
// cube records as vho_mc writes them).
int64_t vho_mesh_indexed(void* p, float* vertices, uint8_t* rgb, int64_t vcap, int32_t* faces, int32_t* face_cubes,
                         int64_t fcap, int64_t* n_faces) {
  Oracle* o = static_cast<Oracle*>(p);
  const int64_t nt = mc_all<false>(p, nullptr, nullptr, nullptr, 0, true);
  std::vector<float> xyz((size_t)nt * 9);
  std::vector<int32_t> cubes((size_t)nt * 4);
  mc_all<false>(p, xyz.data(), cubes.data(), nullptr, nt, true);
  if (n_faces) *n_faces = nt;
  std::map<std::tuple<int, int, int, int>, int32_t> id;
  std::vector<std::array<int, 4>> edges;  // lower lattice point and axis, by vertex id
  std::vector<int32_t> f((size_t)nt * 3);
  for (int64_t t = 0; t < nt; ++t) {
    const int32_t* c = &cubes[(size_t)t * 4];
    const int ci = c[3] >> 8, tt = c[3] & 255;
    for (int j = 0; j < 3; ++j) {
      const int e = hexval(kTri[ci][3 * tt + j]);
      const int* a = kCorner[kEdge[e][0]];
      const int* b = kCorner[kEdge[e][1]];
      const int lx = c[0] + std::min(a[0], b[0]), ly = c[1] + std::min(a[1], b[1]), lz = c[2] + std::min(a[2], b[2]);
      const int axis = a[0] != b[0] ? 0 : (a[1] != b[1] ? 1 : 2);
      auto ins = id.emplace(std::make_tuple(lx, ly, lz, axis), (int32_t)edges.size());
      if (ins.second) edges.push_back({lx, ly, lz, axis});
      f[(size_t)t * 3 + j] = ins.first->second;
    }
  }
  const int64_t nv = (int64_t)edges.size();
  auto value = [&](int x, int y, int z, float& s, const uint8_t*& col) {
    const int64_t b = o->find(key_of(floor_div8(x), floor_div8(y), floor_div8(z)));
    const int i = ((x - floor_div8(x) * 8) * 8 + (y - floor_div8(y) * 8)) * 8 + (z - floor_div8(z) * 8);
    float w;
    o->twin(b, i, s, w);
    col = o->color ? &o->rgb8[3 * ((size_t)b * 512 + i)] : nullptr;
  };
  if (vertices && vcap >= nv)
    for (int64_t k = 0; k < nv; ++k) {
      const auto& E = edges[(size_t)k];
      const float pa[3] = {(float)E[0], (float)E[1], (float)E[2]};
      const float pb[3] = {(float)(E[0] + (E[3] == 0)), (float)(E[1] + (E[3] == 1)), (float)(E[2] + (E[3] == 2))};
      float va, vb;
      const uint8_t *ca, *cb;
      value(E[0], E[1], E[2], va, ca);
      value((int)pb[0], (int)pb[1], (int)pb[2], vb, cb);
      float q[3];
      uint8_t col[3] = {0, 0, 0};
      interp(pa, pb, va, vb, q, ca, cb, (rgb && ca) ? col : nullptr);
      for (int r = 0; r < 3; ++r) vertices[3 * k + r] = q[r] * o->vox;
      if (rgb) std::memcpy(rgb + 3 * k, col, 3);
    }
  if (faces && fcap >= nt) std::memcpy(faces, f.data(), f.size() * sizeof(int32_t));
  if (face_cubes && fcap >= nt) std::memcpy(face_cubes, cubes.data(), cubes.size() * sizeof(int32_t));
  return nv;
}

double vho_point_triangle_dist(const double* p, const double* a, const double* b, const double* c) {
  return std::sqrt(point_triangle_dist2(p, a, b, c));
}

// One-sided distances from mesh A to mesh B (metres): the samples are every A triangle's
// three vertices and centroid; each is measured against B's triangles in the 27 lattice
// cubes around its triangle's cube (cube records as vho_mc writes them). out[0] = max,
// out[1] = mean over samples that found a B triangle there, out[2] = samples that found none,
// out[3] = samples. dist (may be null): 4 * na per-sample distances, -1 where none was found.
void vho_mesh_distance(const float* a_xyz, const int32_t* a_cubes, int64_t na, const float* b_xyz,
                       const int32_t* b_cubes, int64_t nb, double* out, float* dist) {
  std::vector<int64_t> order((size_t)nb);
  for (int64_t i = 0; i < nb; ++i) order[(size_t)i] = i;
  auto key = [](const int32_t* c) { return key_of(c[0], c[1], c[2]); };
  std::sort(order.begin(), order.end(),
            [&](int64_t x, int64_t y) { return key(b_cubes + 4 * x) < key(b_cubes + 4 * y); });
  std::unordered_map<int64_t, std::pair<int64_t, int64_t>> range;  // cube -> [first, end) in order
  range.reserve((size_t)nb);
  for (int64_t i = 0; i < nb;) {
    const int64_t k = key(b_cubes + 4 * order[(size_t)i]);
    int64_t j = i + 1;
    while (j < nb && key(b_cubes + 4 * order[(size_t)j]) == k) ++j;
    range.emplace(k, std::make_pair(i, j));
    i = j;
  }
  double mx = 0.0, sum = 0.0;
  int64_t far = 0, found = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(max : mx) reduction(+ : sum, far, found)
  for (long t = 0; t < (long)na; ++t) {
    const float* P = a_xyz + 9 * t;
    const int32_t* C = a_cubes + 4 * t;
    double pts[4][3];
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) pts[j][k] = P[3 * j + k];
    for (int k = 0; k < 3; ++k) pts[3][k] = (pts[0][k] + pts[1][k] + pts[2][k]) / 3.0;
    double best[4] = {1e300, 1e300, 1e300, 1e300};
    for (int dx = -1; dx <= 1; ++dx)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dz = -1; dz <= 1; ++dz) {
          auto it = range.find(key_of(C[0] + dx, C[1] + dy, C[2] + dz));
          if (it == range.end()) continue;
          for (int64_t i = it->second.first; i < it->second.second; ++i) {
            const float* Q = b_xyz + 9 * order[(size_t)i];
            const double a[3] = {Q[0], Q[1], Q[2]}, b[3] = {Q[3], Q[4], Q[5]}, c[3] = {Q[6], Q[7], Q[8]};
            for (int s2 = 0; s2 < 4; ++s2) best[s2] = std::min(best[s2], point_triangle_dist2(pts[s2], a, b, c));
          }
        }
    for (int s2 = 0; s2 < 4; ++s2) {
      if (best[s2] >= 1e300) {
        ++far;
        if (dist) dist[4 * t + s2] = -1.0f;
        continue;
      }
      const double d = std::sqrt(best[s2]);
      if (dist) dist[4 * t + s2] = (float)d;
      mx = std::max(mx, d);
      sum += d;
      ++found;
    }
  }
  out[0] = mx;
  out[1] = found ? sum / (double)found : 0.0;
  out[2] = (double)far;
  out[3] = (double)(4 * na);
}

// Triangles of one cube from 8 given corner values (all observed), lattice origin (0,0,0),
// vox = 1: for exhaustive checks of the 256 cases.
// Places n blocks (keys: n x 3 block coordinates) with the given FP32 values and weights
// (n x 512 each, in-block order) into an FP32 twin and its FP64 shadow, replacing blocks
// already there: a volume built voxel by voxel, for tests that need chosen values. Returns
// -1 unless the twin stores FP32.
int vho_import_f32(void* p, const int32_t* keys, const float* tsdf, const float* weight, int64_t n) {
  Oracle* o = static_cast<Oracle*>(p);
  if (o->storage != 0) return -1;
  for (int64_t i = 0; i < n; ++i) {
    const int64_t k = key_of(keys[3 * i], keys[3 * i + 1], keys[3 * i + 2]);
    int64_t b = o->find(k);
    if (b < 0) b = o->add_block(k);
    for (int v = 0; v < 512; ++v) {
      const size_t d = (size_t)b * 512 + v, s = (size_t)i * 512 + v;
      o->s32[d] = tsdf[s];
      o->w32[d] = weight[s];
      o->s64[d] = tsdf[s];
      o->w64[d] = weight[s];
    }
  }
  return 0;
}

int vho_mc_cube(const float* val, float* xyz) {
  int ci = 0;
  for (int c = 0; c < 8; ++c)
    if (val[c] < 0.0f) ci |= 1 << c;
  const char* tl = kTri[ci];
  const int n = (int)std::strlen(tl) / 3;
  for (int t = 0; t < n; ++t)
    for (int j = 0; j < 3; ++j) {
      const int e = hexval(tl[3 * t + j]);
      const int ca = kEdge[e][0], cb = kEdge[e][1];
      const float pa[3] = {(float)kCorner[ca][0], (float)kCorner[ca][1], (float)kCorner[ca][2]};
      const float pb[3] = {(float)kCorner[cb][0], (float)kCorner[cb][1], (float)kCorner[cb][2]};
      interp(pa, pb, val[ca], val[cb], xyz + 9 * t + 3 * j);
    }
  return n;
}
// The triangle table in this file's form (edge digits), for cross-checks.
const char* vho_tri_case(int ci) { return (ci >= 0 && ci < 256) ? kTri[ci] : ""; }

// Trilinear samples: FP32 twin and FP64 evaluation on the same stored values.
void vho_query(void* p, const float* pts, int64_t n, float* sdf32, double* sdf64, double* grad64, uint8_t* valid) {
  Oracle* o = static_cast<Oracle*>(p);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)n; ++i) {
    float s = 0.0f, g[3];
    double s6 = 0.0, g6[3] = {0, 0, 0};
    const bool ok = sample<float>(*o, (double)pts[3 * i], (double)pts[3 * i + 1], (double)pts[3 * i + 2], s, g);
    if (ok) sample<double>(*o, (double)pts[3 * i], (double)pts[3 * i + 1], (double)pts[3 * i + 2], s6, g6);
    if (sdf32) sdf32[i] = ok ? s : 0.0f;
    if (sdf64) sdf64[i] = ok ? s6 : 0.0;
    if (grad64)
      for (int k = 0; k < 3; ++k) grad64[3 * i + k] = ok ? g6[k] : 0.0;
    if (valid) valid[i] = ok ? 1 : 0;
  }
}
// FP32 twin of the kernels' sample: FP32 value and FP32 gradient w.r.t. q at FP64 points.
void vho_query_twin(void* p, const double* pts, int64_t n, float* sdf, float* grad, uint8_t* valid) {
  Oracle* o = static_cast<Oracle*>(p);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)n; ++i) {
    float s = 0.0f, g[3] = {0, 0, 0};
    const bool ok = sample<float>(*o, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], s, g);
    if (sdf) sdf[i] = ok ? s : 0.0f;
    if (grad)
      for (int k = 0; k < 3; ++k) grad[3 * i + k] = ok ? g[k] : 0.0f;
    if (valid) valid[i] = ok ? 1 : 0;
  }
}

// FP64 trilinear at double-precision world points (for finite differences).
void vho_query64(void* p, const double* pts, int64_t n, double* sdf, double* grad, uint8_t* valid) {
  Oracle* o = static_cast<Oracle*>(p);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)n; ++i) {
    double s = 0.0, g[3] = {0, 0, 0};
    const bool ok = sample<double>(*o, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], s, g);
    if (sdf) sdf[i] = ok ? s : 0.0;
    if (grad)
      for (int k = 0; k < 3; ++k) grad[3 * i + k] = ok ? g[k] : 0.0;
    if (valid) valid[i] = ok ? 1 : 0;
  }
}

// Voxel gradient of sum_i up_i r_i with r_i the trilinear sample at pts[i]: each point with a
// valid sample adds up_i w_k to its 8 corner voxels, w_k = (ax ay) az from the same cell
// fraction as sample(): in FP64 (twin = 0), or as the kernels form it (twin = 1: FP32
// weights and the FP32 product up_i w_k, so that only the device's FP32 accumulation
// differs). grad: blocks in insertion order (vho_export's), 512 voxels each, accumulated in
// FP64 point by point in input order (a CPU scatter with no atomics). Returns the number of
// points with a valid sample.
extern "C++" template <class R>
int64_t scatter_grad(Oracle* o, const double* pts, const float* up, int64_t n, double* grad) {
  int64_t used = 0;
  for (int64_t i = 0; i < n; ++i) {
    double s = 0.0, g[3];
    if (!sample<double>(*o, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], s, g)) continue;
    ++used;
    const double vd = (double)o->vox;
    int ix[3];
    R f[3];
    for (int k = 0; k < 3; ++k) {
      const double gq = pts[3 * i + k] / vd, fl = std::floor(gq);
      ix[k] = (int)fl;
      f[k] = (R)(gq - fl);
    }
    const R ax[2] = {(R)1 - f[0], f[0]}, ay[2] = {(R)1 - f[1], f[1]}, az[2] = {(R)1 - f[2], f[2]};
    for (int k = 0; k < 8; ++k) {
      const int dx = k >> 2, dy = (k >> 1) & 1, dz = k & 1;
      const int x = ix[0] + dx, y = ix[1] + dy, z = ix[2] + dz;
      const int64_t b = o->find(key_of(floor_div8(x), floor_div8(y), floor_div8(z)));
      const int v = ((x & 7) * 8 + (y & 7)) * 8 + (z & 7);
      grad[(size_t)b * 512 + v] += (double)((R)up[i] * ((ax[dx] * ay[dy]) * az[dz]));
    }
  }
  return used;
}
int64_t vho_scatter_grad(void* p, const double* pts, const float* up, int64_t n, double* grad, int twin) {
  Oracle* o = static_cast<Oracle*>(p);
  return twin ? scatter_grad<float>(o, pts, up, n, grad) : scatter_grad<double>(o, pts, up, n, grad);
}

// Blocks crossed by one pixel's truncation band, in walk order; returns the count.
int vho_band(void* p, const float* c2w, int u, int v, float d, int32_t* xyz, int cap) {
  Oracle* o = static_cast<Oracle*>(p);
  int n = 0;
  band_blocks(*o, c2w, u, v, d, [&](int x, int y, int z) {
    if (n < cap) { xyz[3 * n] = x; xyz[3 * n + 1] = y; xyz[3 * n + 2] = z; }
    ++n;
  });
  return n;
}

uint16_t vho_f32_to_f16(float f) { return to_half(f); }
float vho_f16_to_f32(uint16_t h) { return from_half(h); }
void vho_vertex_interp(const float* a, const float* b, float va, float vb, float* out) { interp(a, b, va, vb, out); }
void vho_vertex_interp_rgb(const float* a, const float* b, float va, float vb, const uint8_t* ca, const uint8_t* cb,
                           float* out, uint8_t* cout) {
  interp(a, b, va, vb, out, ca, cb, cout);
}
void vho_color_update(uint8_t* c, float w_old, const uint8_t* px) { color_update(c, w_old, px); }

}  // extern "C"
