// vh_math.h - scalar math shared by the gfx950 kernels and the host side of libvh.
//
// Everything here is written so that a CPU build with -ffp-contract=off and the HIP
// build with -ffp-contract=off evaluate the same IEEE-754 operations in the same order:
// no FMA contraction, correctly-rounded division, floorf/roundf (both exact). That is
// what makes the allocated block set and the FP32 voxel values reproducible bit-for-bit
// by a CPU twin of the algorithm.
//
// Lattice conventions (reference src/tsdf.cu:24-64):
//   voxel (i,j,k) sits at world (i*vox, j*vox, k*vox) - corner lattice, no origin offset;
//   block = floor(voxel / 8); in-block linear index = (x*8 + y)*8 + z, z fastest.
// Camera conventions (reference src/tsdf.cu:66-116, PointCloudGenerator.cpp:118-125):
//   c2w is a row-major 4x4 camera-to-world; world->camera is R^T (p - t);
//   projection rounds to the nearest pixel with roundf.
#pragma once
#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define VH_HD __host__ __device__ __forceinline__
#else
#define VH_HD static inline
#endif

namespace vh {

constexpr int kB = 8;          // voxels per block edge
constexpr int kB3 = 512;       // voxels per block
constexpr int kSlotsPerBucket = 4;
constexpr int kKeyBits = 21;   // bits per axis in a packed key
constexpr int kKeyBias = 1 << 20;
constexpr uint64_t kEmptyKey = ~0ull;
constexpr uint32_t kNoBlock = 0xffffffffu;

// ---------------------------------------------------------------- keys and hashing
// A block coordinate is packed as three biased 21-bit fields; bit 63 is always 0, so the
// all-ones word can never be a valid key and serves as the empty-slot sentinel.
// (Replaces the reference's int3 key with the {999999,999999,999999} empty value,
// src/tsdf.cu:1488, and BlockEqual, include/tsdf.cuh:158-165.)
VH_HD bool key_in_range(int x, int y, int z) {
  return x >= -kKeyBias && x < kKeyBias && y >= -kKeyBias && y < kKeyBias &&
         z >= -kKeyBias && z < kKeyBias;
}
VH_HD uint64_t pack_key(int x, int y, int z) {
  return ((uint64_t)(uint32_t)(x + kKeyBias) << 42) | ((uint64_t)(uint32_t)(y + kKeyBias) << 21) |
         (uint64_t)(uint32_t)(z + kKeyBias);
}
VH_HD void unpack_key(uint64_t k, int& x, int& y, int& z) {
  const uint64_t m = (1ull << kKeyBits) - 1;
  x = (int)((k >> 42) & m) - kKeyBias;
  y = (int)((k >> 21) & m) - kKeyBias;
  z = (int)(k & m) - kKeyBias;
}
// 64-bit finaliser (MurmurHash3 fmix64). The bucket is the top log2_buckets bits, so the
// table size is a power of two and no modulo is needed (the reference reduces a
// Teschner xor-hash modulo 100001, include/tsdf.cuh:144-156, include/vhashing.h:151-152).
VH_HD uint64_t mix_key(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}
VH_HD uint64_t bucket_of(uint64_t key, int log2_buckets) {
  return mix_key(key) >> (64 - log2_buckets);
}

// ---------------------------------------------------------------- intrinsics / pose
struct Intr {
  float fx, fy, cx, cy;
};

// world -> camera, R^T (p - t) with c2w row-major (reference base2cam, src/tsdf.cu:81-93).
VH_HD void world_to_cam(const float* T, float px, float py, float pz, float& x, float& y,
                        float& z) {
  const float dx = px - T[3], dy = py - T[7], dz = pz - T[11];
  x = T[0] * dx + T[4] * dy + T[8] * dz;
  y = T[1] * dx + T[5] * dy + T[9] * dz;
  z = T[2] * dx + T[6] * dy + T[10] * dz;
}
// camera -> world, R p + t (reference cam2base, src/tsdf.cu:95-105).
VH_HD void cam_to_world(const float* T, float x, float y, float z, float& px, float& py,
                        float& pz) {
  px = x * T[0] + y * T[1] + z * T[2] + T[3];
  py = x * T[4] + y * T[5] + z * T[6] + T[7];
  pz = x * T[8] + y * T[9] + z * T[10] + T[11];
}

// ---------------------------------------------------------------- integration (one voxel)
// Camera coordinates of the voxels of one block, in the reference's subtract-then-rotate
// form (base2cam, src/tsdf.cu:81-93) evaluated once per block origin o = 8*b*vox:
//   c = R^T (o - t),  then voxel (k, j, i) of the block: c + vox*R^T (k, j, i).
// With a = R^T e_x vox, b = R^T e_y vox, c = R^T e_z vox (per frame) a lane holding one
// (j, i) column adds its offset once: c_l = c + fma(j, b, i*c), and voxel k along x is
// fma(k, a, c_l) (one rounding per coordinate).
struct BlockFrame {
  float ax[3], ay[3], az[3];  // per-frame step vectors (camera coordinates per voxel step)
};
VH_HD void block_frame(const float* T, float vox, BlockFrame& f) {
  for (int i = 0; i < 3; ++i) {
    f.ax[i] = T[i] * vox;
    f.ay[i] = T[4 + i] * vox;
    f.az[i] = T[8 + i] * vox;
  }
}
// Camera coordinates of the block origin (block coordinates bx, by, bz).
VH_HD void block_origin_cam(const float* T, float vox, int bx, int by, int bz, float c[3]) {
  const float d0 = (float)(bx * kB) * vox - T[3];
  const float d1 = (float)(by * kB) * vox - T[7];
  const float d2 = (float)(bz * kB) * vox - T[11];
  for (int i = 0; i < 3; ++i) c[i] = T[i] * d0 + T[4 + i] * d1 + T[8 + i] * d2;
}

// Homogeneous pixel coordinates of a camera-frame vector (x, y, z): (fx x + cx z, fy y + cy z,
// z), each sum one fused multiply-add over the rounded product c z. The intrinsics are folded
// into the per-frame steps and the block origin this way, so a voxel's pixel is (X/z, Y/z)
// and its three coordinates cost three fused multiply-adds, as its camera coordinates did.
VH_HD void to_pixel_h(const Intr& K, const float c[3], float out[3]) {
  out[0] = fmaf(K.fx, c[0], K.cx * c[2]);
  out[1] = fmaf(K.fy, c[1], K.cy * c[2]);
  out[2] = c[2];
}
// Homogeneous steps of a frame (to_pixel_h of the camera steps a, b, c).
VH_HD void block_frame_h(const float* T, float vox, const Intr& K, BlockFrame& f) {
  BlockFrame c;
  block_frame(T, vox, c);
  to_pixel_h(K, c.ax, f.ax);
  to_pixel_h(K, c.ay, f.ay);
  to_pixel_h(K, c.az, f.az);
}
// Homogeneous coordinates of the column (j, i) of a block whose origin is at cb (homogeneous):
// cb + fma(j, b, i*c); voxel k of the column is then fma(k, a, column).
VH_HD void column_h(const BlockFrame& f, const float cb[3], int j, int i, float out[3]) {
  const float fj = (float)j, fi = (float)i;
  for (int r = 0; r < 3; ++r) out[r] = cb[r] + fmaf(fj, f.ay[r], fi * f.az[r]);
}

// Observation of the voxel at homogeneous point (X, Y, z) (reference IntegrateHashKernel skip
// rules, src/tsdf.cu:706-738), in two steps so that a caller can issue all depth fetches of a
// block before using any of them:
//   pixel: one correctly rounded reciprocal iz = 1/z (passed in by the caller) serves both
//     image coordinates; each is the exact product X * iz rounded once to the nearest integer
//     (ties to even) by adding 1.5 * 2^23 in the same fused multiply-add: fma(X, iz, M) - M
//     (the reference divides per coordinate and rounds ties away from zero, cam2frame,
//     src/tsdf.cu:75-79). Within the image the sum lies in [2^23, 2^24), where the float
//     spacing is 1, so the sum is M plus the rounded coordinate, and its bits give the pixel
//     index without a conversion. -1 when z is outside 2^-64 <= z <= 2^64 or the pixel outside
//     the image. The range bound keeps z and 1/z normal, where the kernels' three-instruction
//     reciprocal equals the IEEE quotient.
//   truncation: the pixel's pre-scaled depth m = d * (1/mu) for a depth 0 < d <= max_depth
//     (-inf otherwise) gives t = fma(-z, 1/mu, m) = (d - z) / mu with one rounding; the voxel
//     is observed iff t > -1 (the reference's d - z > -mu, with d > 0 and d <= max_depth), and
//     dist = min(1, t).
constexpr float kMinZ = 0x1p-64f, kMaxZ = 0x1p64f;
constexpr float kPixMagic = 0x1.8p23f;  // M = 1.5 * 2^23, bits 0x4B400000
constexpr uint32_t kPixMagicBits = 0x4B400000u;
VH_HD int pixel_of(float X, float Y, float z, float iz, int W, int H) {
  const float uf = fmaf(X, iz, kPixMagic);
  const float vf = fmaf(Y, iz, kPixMagic);
  const bool inside = (z >= kMinZ) & (z <= kMaxZ) & (uf >= kPixMagic) & (uf <= kPixMagic + (float)(W - 1)) &
                      (vf >= kPixMagic) & (vf <= kPixMagic + (float)(H - 1));
  // both factors are below 2^24 (W, H <= 16384): the masks let the compiler use a 24-bit
  // multiply-add instead of a 64-bit one
  return inside ? (int)(((uint32_t)(vf - kPixMagic) & 0xffffffu) * ((uint32_t)W & 0xffffffu) + (uint32_t)(uf - kPixMagic))
                : -1;
}
// Byte offset 4 (v W + u) of pixel (u, v) from the sums uf = M + u, vf = M + v of pixel_of
// (both inside the image): the low 24 bits of vf are 0x400000 + v and the bits of uf are
// 0x4B400000 + u, so low24(vf) * 4W + 4 bits(uf) exceeds the offset by
// pix_bias(W) = 4 (0x400000 W + 0x4B400000) modulo 2^32 (one shift-add and one 24-bit
// multiply-add).
VH_HD uint32_t pix_bias(int P) { return 4u * (0x400000u * (uint32_t)P + kPixMagicBits); }
VH_HD float scaled_depth(float d, float max_depth, float inv_mu) {
  return (d > 0.0f && d <= max_depth) ? d * inv_mu : -INFINITY;
}

// Running average with the old weight (the intent of src/tsdf.cu:855-858; the hash kernel
// at :739-742 assigns the new weight first, which is a defect and is not reproduced), in the
// increment form s + (dist - s) * RN(1/(w + 1)): one rounded difference and one fused
// multiply-add by the correctly rounded reciprocal of the new weight. cap > 0 bounds the
// stored weight, turning the mean into an exponential average.
VH_HD void tsdf_update(float& s, float& w, float dist, float cap) {
  const float w_new = w + 1.0f;
  s = fmaf(dist - s, 1.0f / w_new, s);
  w = (cap > 0.0f && w_new > cap) ? cap : w_new;
}

// ---------------------------------------------------------------- binary16 conversion
// Round-to-nearest-even float -> binary16 bits, and exact binary16 -> float. Used on the
// host (exports, tests); kernels use the hardware conversion, which rounds the same way.
VH_HD uint16_t f32_to_f16_bits(float f) {
  union { float f; uint32_t u; } c;
  c.f = f;
  const uint32_t u = c.u;
  const uint32_t sign = (u >> 16) & 0x8000u;
  const uint32_t a = u & 0x7fffffffu;
  if (a >= 0x7f800000u) return (uint16_t)(sign | (a > 0x7f800000u ? 0x7e00u : 0x7c00u));
  if (a >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);  // rounds to >= 65520 -> inf
  if (a < 0x38800000u) {                                      // result is subnormal or zero
    if (a < 0x33000000u) return (uint16_t)sign;               // < 2^-25 rounds to 0
    const uint32_t m = (a & 0x7fffffu) | 0x800000u;
    const int e = (int)(a >> 23);                             // 102..112
    const int sh = 126 - e;                                   // 14..24
    uint32_t r = m >> sh;
    const uint32_t rem = m & ((1u << sh) - 1u);
    const uint32_t half = 1u << (sh - 1);
    if (rem > half || (rem == half && (r & 1u))) ++r;
    return (uint16_t)(sign | r);
  }
  uint32_t r = ((a >> 13) - (112u << 10));
  const uint32_t rem = a & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (r & 1u))) ++r;
  return (uint16_t)(sign | r);
}
VH_HD float f16_bits_to_f32(uint16_t h) {
  const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  const uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
  union { float f; uint32_t u; } c;
  if (e == 0) {
    c.f = (float)m * 5.9604644775390625e-08f;  // m * 2^-24, exact
    c.u |= sign;
    return c.f;
  }
  if (e == 31) {
    c.u = sign | 0x7f800000u | (m << 13);
    return c.f;
  }
  c.u = sign | ((e + 112u) << 23) | (m << 13);
  return c.f;
}

// Compensated binary16 (storage F16R): value = binary16 h + r * ulp(h)/256 with r an int8.
// ulp(h)/256 is a power of two, so decoding is exact in FP32 (11 + 8 significant bits) and the
// encoder's residual is exact before its final rounding to an integer.
VH_HD float f16r_scale_from_bits(uint16_t h) {  // ulp(h) / 256
  const uint32_t e = (h >> 10) & 0x1fu;
  union { uint32_t u; float f; } c;
  c.u = (e == 0 ? 95u : e + 94u) << 23;  // 2^-32 for subnormals, else 2^(e-33)
  return c.f;
}
VH_HD float f16r_inv_scale_from_bits(uint16_t h) {
  const uint32_t e = (h >> 10) & 0x1fu;
  union { uint32_t u; float f; } c;
  c.u = (e == 0 ? 159u : 160u - e) << 23;
  return c.f;
}
VH_HD float f16r_decode(uint16_t h, int r) {
  return f16_bits_to_f32(h) + (float)r * f16r_scale_from_bits(h);
}
// Residual of s against its binary16 rounding h, in units of ulp(h)/256 (an exact quotient:
// s - h is exact by Sterbenz's lemma, the scale a power of two), rounded to the nearest
// integer with ties to even and clamped to int8.
VH_HD int f16r_residual(float s, uint16_t h) {
  const float d = (s - f16_bits_to_f32(h)) * f16r_inv_scale_from_bits(h);
  float r = rintf(d);
  r = r > 127.0f ? 127.0f : (r < -128.0f ? -128.0f : r);
  return (int)r;
}
// f16r_scale_from_bits and its inverse from the FP32 value hf of the binary16 number instead of
// its bits (what a kernel holds after the conversion): a normal h with biased exponent e has
// FP32 exponent e + 112, so ulp(h)/256 = 2^(e - 33) has the FP32 exponent field of hf minus 18;
// subnormal halves and zero have an FP32 exponent field below 113, and the maximum with the
// exponent field 95 (2^-32) covers them. Equal to the bit forms for every finite binary16
// (tests/host_checks/f16r_scale.cpp).
VH_HD int32_t f16r_scale_field(float hf) {
  union { float f; int32_t i; } c;
  c.f = hf;
  const int32_t e = (c.i & 0x7f800000) - (18 << 23);
  return e > (95 << 23) ? e : (95 << 23);
}
VH_HD float f16r_scale_of(float hf) {
  union { int32_t i; float f; } c;
  c.i = f16r_scale_field(hf);
  return c.f;
}
VH_HD float f16r_inv_scale_of(float hf) {
  union { int32_t i; float f; } c;
  c.i = (254 << 23) - f16r_scale_field(hf);
  return c.f;
}

// ---------------------------------------------------------------- truncation-band DDA
// Enumerates every block whose cube is crossed by the ray segment of pixel (u,v) between
// camera depths max(d-mu, min_depth) and d+mu (Amanatides-Woo traversal in block units).
// This replaces the reference's whole-ray DDA over every 10th pixel (src/tsdf.cu:2088-2238),
// which ignores the truncation band. Steps are chosen among the axes that still have to
// move, smallest crossing parameter first, ties to the lower axis; the walk therefore ends
// exactly at the end block after |e - i|_1 steps. band_setup returns false if the band
// leaves the key range.
struct BandDDA {
  float ax, ay, az;  // start point, block units
  float rx, ry, rz;  // band_recip of the direction (end - start)
  int ix, iy, iz;    // current block
  int ex, ey, ez;    // last block
};

// Reciprocal of a direction component for the crossing parameters: RN(1/d) for 2^-64 <= |d|
// <= 2^64, else 2^64 with d's sign (an axis that moves has |d| >= 2^-64 unless its segment
// straddles the face 0 within 2^-64 of it; a zero component's axis does not move). rn: a
// correctly rounded reciprocal valid on that range (the kernels pass their three-instruction
// one).
template <class Rn>
VH_HD float band_recip(float d, Rn rn) {
  const float m = fabsf(d);
  return (m >= 0x1p-64f && m <= 0x1p64f) ? rn(d) : copysignf(0x1p64f, d);
}

// Reciprocals of the intrinsics' focal lengths and of the block size, RN(1/x) of each: the band
// setup multiplies by them instead of dividing.
struct BandScale {
  float ifx, ify, ibsz;
};
VH_HD BandScale band_scale(const Intr& K, float bsz) { return BandScale{1.0f / K.fx, 1.0f / K.fy, 1.0f / bsz}; }

// Segment of pixel (u, v) between camera depths t0 = max(d - mu, min_depth) and t1 = d + mu,
// in block units: the pixel's ray per unit depth r = ((u - cx) * RN(1/fx), (v - cy) * RN(1/fy)),
// camera points (r t, t), world points R p + t (cam_to_world), block coordinates
// world * RN(1/bsz).
template <class Rn>
VH_HD bool band_setup(const float* T, const Intr& K, const BandScale& sc, float mu, float min_depth, int u,
                      int v, float d, BandDDA& s, Rn rn) {
  float t0 = d - mu;
  if (t0 < min_depth) t0 = min_depth;
  const float t1 = d + mu;
  const float rx = ((float)u - K.cx) * sc.ifx, ry = ((float)v - K.cy) * sc.ify;
  float wax, way, waz, wbx, wby, wbz;
  cam_to_world(T, rx * t0, ry * t0, t0, wax, way, waz);
  cam_to_world(T, rx * t1, ry * t1, t1, wbx, wby, wbz);
  s.ax = wax * sc.ibsz; s.ay = way * sc.ibsz; s.az = waz * sc.ibsz;
  const float bx = wbx * sc.ibsz, by = wby * sc.ibsz, bz = wbz * sc.ibsz;
  const float lim = 1048000.0f;  // keeps every block inside the 21-bit key range (NaN fails)
  const int in_range = (int)(fabsf(s.ax) < lim) & (int)(fabsf(s.ay) < lim) & (int)(fabsf(s.az) < lim) &
                       (int)(fabsf(bx) < lim) & (int)(fabsf(by) < lim) & (int)(fabsf(bz) < lim);
  if (!in_range) return false;
  s.rx = band_recip(bx - s.ax, rn); s.ry = band_recip(by - s.ay, rn); s.rz = band_recip(bz - s.az, rn);
  s.ix = (int)floorf(s.ax); s.iy = (int)floorf(s.ay); s.iz = (int)floorf(s.az);
  s.ex = (int)floorf(bx); s.ey = (int)floorf(by); s.ez = (int)floorf(bz);
  return true;
}
// Crossing parameter of the next block face along one axis (only for axes still moving):
// (face - a) * band_recip(direction).
VH_HD float band_t(int i, int e, float a, float rdir) {
  const float bound = (float)(e > i ? i + 1 : i);
  return (bound - a) * rdir;
}
// Advances to the next block; returns false when the current block was the last one.
VH_HD bool band_next(BandDDA& s) {
  const bool mx = s.ix != s.ex, my = s.iy != s.ey, mz = s.iz != s.ez;
  if (!(mx || my || mz)) return false;
  const float tx = mx ? band_t(s.ix, s.ex, s.ax, s.rx) : 0.0f;
  const float ty = my ? band_t(s.iy, s.ey, s.ay, s.ry) : 0.0f;
  const float tz = mz ? band_t(s.iz, s.ez, s.az, s.rz) : 0.0f;
  // smallest crossing parameter first; ties go to the lower axis
  const bool take_x = mx && (!my || !(ty < tx)) && (!mz || !(tz < tx));
  const bool take_y = !take_x && my && (!mz || !(tz < ty));
  if (take_x) s.ix += s.ex > s.ix ? 1 : -1;
  else if (take_y) s.iy += s.ey > s.iy ? 1 : -1;
  else s.iz += s.ez > s.iz ? 1 : -1;
  return true;
}
VH_HD bool depth_valid(float d, float max_depth) { return d > 0.0f && d <= max_depth; }

// ---------------------------------------------------------------- marching cubes
// Corner order of the reference (include/tsdf.cuh:234-241): x/y swapped against Bourke's
// numbering, which fixes the triangle winding. Edge e joins corners kEdgeA[e], kEdgeB[e]
// (reference VertexInterp calls, src/tsdf.cu:1001-1037).
VH_HD constexpr int corner_dx(int c) { return (c == 2 || c == 3 || c == 6 || c == 7) ? 1 : 0; }
VH_HD constexpr int corner_dy(int c) { return (c == 1 || c == 2 || c == 5 || c == 6) ? 1 : 0; }
VH_HD constexpr int corner_dz(int c) { return c >= 4 ? 1 : 0; }
VH_HD constexpr int edge_a(int e) { return e < 8 ? e : e - 8; }
VH_HD constexpr int edge_b(int e) { return e < 8 ? ((e & 3) == 3 ? e - 3 : e + 1) : e - 4; }

// |v| < 1e-5 evaluated as the reference does it (float |v| promoted to double against the
// double literal, src/tsdf.cu:1645-1650): equivalent to |v| <= 1e-5f in float.
VH_HD bool near_iso(float v) { return fabsf(v) <= 1e-5f; }

// Interpolates along one cube edge in voxel units (reference VertexInterp, iso = 0,
// src/tsdf.cu:1640-1660). p1/p2 are integer lattice coordinates held as floats.
VH_HD void vertex_interp(float p1x, float p1y, float p1z, float p2x, float p2y, float p2z,
                         float v1, float v2, float& x, float& y, float& z) {
  if (near_iso(v1)) {
    x = p1x; y = p1y; z = p1z;
    return;
  }
  if (near_iso(v2)) {
    x = p2x; y = p2y; z = p2z;
    return;
  }
  if (near_iso(v1 - v2)) {
    x = p1x; y = p1y; z = p1z;
    return;
  }
  const float mu = (0.0f - v1) / (v2 - v1);
  x = p1x + mu * (p2x - p1x);
  y = p1y + mu * (p2y - p1y);
  z = p1z + mu * (p2z - p1z);
}

// Which end point of its edge the vertex_interp vertex lands on: 0 on p1, 1 on p2, 2 on
// neither. p1a, p2a are the end points' coordinates along the edge's axis, computed as
// vertex_interp computes that coordinate; the other two coordinates of the vertex equal the
// end points' exactly (mu is finite after the snaps, and mu * 0 adds zero). Vertices on
// distinct edges of a cube coincide exactly when both land on the corner the edges share.
VH_HD int vertex_end(float p1a, float p2a, float v1, float v2) {
  if (near_iso(v1)) return 0;
  if (near_iso(v2)) return 1;
  if (near_iso(v1 - v2)) return 0;
  const float mu = (0.0f - v1) / (v2 - v1);
  const float t = p1a + mu * (p2a - p1a);
  return t == p1a ? 0 : (t == p2a ? 1 : 2);
}

// VertexInterp with the colour lerp of the reference (src/tsdf.cu:1655-1657): a snapped vertex
// takes its corner's colour; otherwise each channel is c1 + mu*(c2 - c1) (the difference is
// an exact integer, the rest FP32), truncated to 8 bits.
VH_HD void vertex_interp_rgb(float p1x, float p1y, float p1z, float p2x, float p2y, float p2z, float v1,
                             float v2, uint32_t c1, uint32_t c2, float& x, float& y, float& z, uint32_t& c) {
  if (near_iso(v1)) {
    x = p1x; y = p1y; z = p1z; c = c1;
    return;
  }
  if (near_iso(v2)) {
    x = p2x; y = p2y; z = p2z; c = c2;
    return;
  }
  if (near_iso(v1 - v2)) {
    x = p1x; y = p1y; z = p1z; c = c1;
    return;
  }
  const float mu = (0.0f - v1) / (v2 - v1);
  x = p1x + mu * (p2x - p1x);
  y = p1y + mu * (p2y - p1y);
  z = p1z + mu * (p2z - p1z);
  c = 0;
  for (int k = 0; k < 3; ++k) {
    const int a = (int)((c1 >> (8 * k)) & 255u), b = (int)((c2 >> (8 * k)) & 255u);
    const float v = (float)a + mu * (float)(b - a);
    c |= (uint32_t)(uint8_t)v << (8 * k);
  }
}

// ---------------------------------------------------------------- trilinear weights
// Corner k = (dx<<2)|(dy<<1)|dz of the lattice cell containing g; weight products are
// formed as (wx*wy)*wz and sums run in k order, so CPU and GPU agree bit-for-bit.
VH_HD void trilinear_weights(float fx, float fy, float fz, float w[8]) {
  const float ax[2] = {1.0f - fx, fx}, ay[2] = {1.0f - fy, fy}, az[2] = {1.0f - fz, fz};
  for (int k = 0; k < 8; ++k) w[k] = (ax[k >> 2] * ay[(k >> 1) & 1]) * az[k & 1];
}
// d(sum_k w_k s_k)/d(fx,fy,fz) for the cell values s.
VH_HD void trilinear_dfrac(float fx, float fy, float fz, const float s[8], float g[3]) {
  const float ax[2] = {1.0f - fx, fx}, ay[2] = {1.0f - fy, fy}, az[2] = {1.0f - fz, fz};
  float gx = 0.0f, gy = 0.0f, gz = 0.0f;
  for (int k = 0; k < 8; ++k) {
    const int dx = k >> 2, dy = (k >> 1) & 1, dz = k & 1;
    const float sx = dx ? 1.0f : -1.0f, sy = dy ? 1.0f : -1.0f, sz = dz ? 1.0f : -1.0f;
    gx = gx + ((sx * ay[dy]) * az[dz]) * s[k];
    gy = gy + ((ax[dx] * sy) * az[dz]) * s[k];
    gz = gz + ((ax[dx] * ay[dy]) * sz) * s[k];
  }
  g[0] = gx; g[1] = gy; g[2] = gz;
}

}  // namespace vh
