// vh_synth.cpp - seeded synthetic depth sequences (host code, compiled into libvh).
// The reference loads recorded frames from disk (src/SaveFrame.cpp:154-218); the workloads
// here are generated analytically so that any frame can be reproduced from (seed, index).
#include "vh_synth.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <thread>
#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846;

inline uint64_t splitmix64(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}
// Uniform double in [0,1) from the top 53 bits.
inline double unit(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }

struct Sphere {
  double c[3], r;
};
struct Box {
  double lo[3], hi[3];
};

struct Rng {  // sequential draws for scene layout
  uint64_t state;
  double next() { return unit(splitmix64(state++)); }
  double range(double a, double b) { return a + (b - a) * next(); }
};

}  // namespace

struct vh_scene {
  int preset = 0;
  uint64_t seed = 0;
  int W = 640, H = 480, n_frames = 0;
  double fx = 525.0, fy = 525.0, cx = 319.5, cy = 239.5;
  double ext[3] = {0, 0, 0};
  double dropout = 0.01;
  std::vector<Sphere> spheres;
  std::vector<Box> boxes;  // primitives and walls, all solid
  int n_prim_boxes = 0;
  bool loop_path = false;    // false: orbit about the room centre; true: square loop
  double orbit_r = 0.0;      // metres
  double height = 1.5;       // camera height
  double pitch = -25.0 * kPi / 180.0;
};

namespace {

// Distance from point p to the camera path (at camera height).
double path_distance(const vh_scene& s, const double p[3]) {
  if (!s.loop_path) {
    const double dx = p[0] - 0.5 * s.ext[0], dy = p[1] - 0.5 * s.ext[1];
    const double rho = std::sqrt(dx * dx + dy * dy);
    const double a = rho - s.orbit_r, b = p[2] - s.height;
    return std::sqrt(a * a + b * b);
  }
  const double c[4][2] = {{5, 5}, {15, 5}, {15, 15}, {5, 15}};
  double best = 1e30;
  for (int k = 0; k < 4; ++k) {
    const double* a = c[k];
    const double* b = c[(k + 1) & 3];
    const double ux = b[0] - a[0], uy = b[1] - a[1];
    double t = ((p[0] - a[0]) * ux + (p[1] - a[1]) * uy) / (ux * ux + uy * uy);
    t = std::min(1.0, std::max(0.0, t));
    const double qx = a[0] + t * ux - p[0], qy = a[1] + t * uy - p[1], qz = s.height - p[2];
    best = std::min(best, std::sqrt(qx * qx + qy * qy + qz * qz));
  }
  return best;
}

bool overlaps_wall(const vh_scene& s, const double c[3], double r, size_t n_walls) {
  for (size_t k = 0; k < n_walls; ++k) {
    const Box& b = s.boxes[k];
    double d2 = 0.0;
    for (int a = 0; a < 3; ++a) {
      const double v = std::max(b.lo[a] - c[a], std::max(0.0, c[a] - b.hi[a]));
      d2 += v * v;
    }
    if (d2 < (r + 0.05) * (r + 0.05)) return true;
  }
  return false;
}

void add_walls(vh_scene& s) {
  const double t0 = 9.95, t1 = 10.05, top = 4.0, door = 2.1;
  auto box = [&](double x0, double x1, double y0, double y1, double z0, double z1) {
    s.boxes.push_back(Box{{x0, y0, z0}, {x1, y1, z1}});
  };
  // wall x = 10 with doorways at y in [4.5,5.5] and [14.5,15.5]
  box(t0, t1, 0.0, 4.5, 0.0, top);
  box(t0, t1, 5.5, 14.5, 0.0, top);
  box(t0, t1, 15.5, 20.0, 0.0, top);
  box(t0, t1, 4.5, 5.5, door, top);
  box(t0, t1, 14.5, 15.5, door, top);
  // wall y = 10 with doorways at x in [4.5,5.5] and [14.5,15.5]
  box(0.0, 4.5, t0, t1, 0.0, top);
  box(5.5, t0, t0, t1, 0.0, top);
  box(t1, 14.5, t0, t1, 0.0, top);
  box(15.5, 20.0, t0, t1, 0.0, top);
  box(4.5, 5.5, t0, t1, door, top);
  box(14.5, 15.5, t0, t1, door, top);
}

void place_primitives(vh_scene& s, int n_sph, int n_box, double clearance) {
  Rng rng{splitmix64(s.seed)};
  const size_t n_walls = s.boxes.size();
  int placed_s = 0, placed_b = 0;
  for (int attempt = 0; attempt < 100000 && (placed_s < n_sph || placed_b < n_box); ++attempt) {
    const bool want_sphere = placed_s < n_sph && (placed_b >= n_box || (attempt & 1) == 0);
    if (want_sphere) {
      const double r = rng.range(0.2, 0.5);
      double c[3];
      for (int a = 0; a < 3; ++a) c[a] = rng.range(r, s.ext[a] - r);
      if (path_distance(s, c) < r + clearance || overlaps_wall(s, c, r, n_walls)) continue;
      s.spheres.push_back(Sphere{{c[0], c[1], c[2]}, r});
      ++placed_s;
    } else {
      double h[3], c[3];
      for (int a = 0; a < 3; ++a) h[a] = rng.range(0.1, 0.4);
      for (int a = 0; a < 3; ++a) c[a] = rng.range(h[a], s.ext[a] - h[a]);
      const double rb = std::sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
      if (path_distance(s, c) < rb + clearance || overlaps_wall(s, c, rb, n_walls)) continue;
      s.boxes.push_back(Box{{c[0] - h[0], c[1] - h[1], c[2] - h[2]}, {c[0] + h[0], c[1] + h[1], c[2] + h[2]}});
      ++placed_b;
    }
  }
  s.n_prim_boxes = placed_b;
}

void pose_of(const vh_scene& s, int f, float T[16]) {
  const double yaw = (double)f * kPi / 180.0;
  double px, py;
  double heading;
  if (!s.loop_path) {
    px = 0.5 * s.ext[0] + s.orbit_r * std::cos(yaw);
    py = 0.5 * s.ext[1] + s.orbit_r * std::sin(yaw);
    heading = yaw;  // looking outward
  } else {
    const double c[4][2] = {{5, 5}, {15, 5}, {15, 15}, {5, 15}};
    double arc = std::fmod(0.02 * (double)f, 40.0);
    const int seg = std::min(3, (int)(arc / 10.0));
    const double t = (arc - 10.0 * seg) / 10.0;
    const double* a = c[seg];
    const double* b = c[(seg + 1) & 3];
    px = a[0] + t * (b[0] - a[0]);
    py = a[1] + t * (b[1] - a[1]);
    heading = yaw;
  }
  const double cp = std::cos(s.pitch), sp = std::sin(s.pitch);
  const double fwd[3] = {std::cos(heading) * cp, std::sin(heading) * cp, sp};
  const double right[3] = {std::sin(heading), -std::cos(heading), 0.0};
  const double down[3] = {fwd[1] * right[2] - fwd[2] * right[1], fwd[2] * right[0] - fwd[0] * right[2],
                          fwd[0] * right[1] - fwd[1] * right[0]};
  // columns: camera x (right), y (down), z (forward)
  const double R[3][3] = {{right[0], down[0], fwd[0]}, {right[1], down[1], fwd[1]}, {right[2], down[2], fwd[2]}};
  const double t[3] = {px, py, s.height};
  for (int r = 0; r < 3; ++r) {
    for (int c2 = 0; c2 < 3; ++c2) T[r * 4 + c2] = (float)R[r][c2];
    T[r * 4 + 3] = (float)t[r];
  }
  T[12] = T[13] = T[14] = 0.0f;
  T[15] = 1.0f;
}

// Nearest positive hit along o + t d (d has unit camera-z, so t is camera depth). *id (may
// be null) receives the surface hit: 2a + side for the room's faces (axis a, side 1 = far),
// 6 + k for sphere k, 6 + spheres + k for box k (walls first), -1 for none.
double cast(const vh_scene& s, const double o[3], const double d[3], int* id = nullptr) {
  double best = 1e30;
  int hit = -1;
  // room interior: exit distance
  for (int a = 0; a < 3; ++a) {
    double t = 1e30;
    if (d[a] > 0.0) t = (s.ext[a] - o[a]) / d[a];
    else if (d[a] < 0.0) t = (0.0 - o[a]) / d[a];
    if (t < best) {
      best = t;
      hit = 2 * a + (d[a] > 0.0);
    }
  }
  int k = 6;
  for (const Sphere& sp : s.spheres) {
    const double oc[3] = {o[0] - sp.c[0], o[1] - sp.c[1], o[2] - sp.c[2]};
    const double A = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const double B = oc[0] * d[0] + oc[1] * d[1] + oc[2] * d[2];
    const double C = oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2] - sp.r * sp.r;
    const double disc = B * B - A * C;
    if (disc < 0.0) continue;
    const double t = (-B - std::sqrt(disc)) / A;
    if (t > 1e-9 && t < best) {
      best = t;
      hit = k;
    }
    ++k;
  }
  for (const Box& b : s.boxes) {
    double tmin = -1e30, tmax = 1e30;
    bool miss = false;
    for (int a = 0; a < 3; ++a) {
      if (d[a] == 0.0) {
        if (o[a] < b.lo[a] || o[a] > b.hi[a]) { miss = true; break; }
        continue;
      }
      double t1 = (b.lo[a] - o[a]) / d[a], t2 = (b.hi[a] - o[a]) / d[a];
      if (t1 > t2) std::swap(t1, t2);
      tmin = std::max(tmin, t1);
      tmax = std::min(tmax, t2);
    }
    if (!(miss || tmax < tmin || tmin <= 1e-9) && tmin < best) {
      best = tmin;
      hit = k;
    }
    ++k;
  }
  if (id) *id = best < 1e29 ? hit : -1;
  return best;
}

// Colour of surface id at world point p: a per-surface base colour (channels 48..239 from
// the seed) on a 10 cm world checkerboard whose dark cells carry 5/8 of it. The room's faces
// lie on cell planes (extents are multiples of 10 cm): the 1e-6 cell nudge puts their hits,
// which land within rounding of the plane on either side, into one cell.
void surface_rgb(const vh_scene& s, int id, const double p[3], uint8_t* out) {
  if (id < 0) {
    out[0] = out[1] = out[2] = 0;
    return;
  }
  const uint64_t h = splitmix64(splitmix64(s.seed + 7ull) + (uint64_t)id);
  long cell = 0;
  for (int a = 0; a < 3; ++a) cell += (long)std::floor(p[a] * 10.0 + 1e-6);
  for (int c = 0; c < 3; ++c) {
    const int base = 48 + (int)((h >> (8 * c)) % 192u);
    out[c] = (uint8_t)((cell & 1) ? base * 5 / 8 : base);
  }
}

// out(pixel, depth, surface id, world hit point) per pixel of rows [r0, r1).
template <class Out>
void render_rows(const vh_scene& s, int f, const float* T, int r0, int r1, Out out) {
  double R[3][3], t[3];
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) R[r][c] = (double)T[r * 4 + c];
    t[r] = (double)T[r * 4 + 3];
  }
  for (int v = r0; v < r1; ++v) {
    for (int u = 0; u < s.W; ++u) {
      const double xc = ((double)u - s.cx) / s.fx, yc = ((double)v - s.cy) / s.fy;
      double d[3];
      for (int r = 0; r < 3; ++r) d[r] = R[r][0] * xc + R[r][1] * yc + R[r][2];
      int id = -1;
      const double z = cast(s, t, d, &id);
      const double p[3] = {t[0] + z * d[0], t[1] + z * d[1], t[2] + z * d[2]};
      out(v * s.W + u, z, id, p);
    }
  }
}

template <class Fn>
void parallel_rows(int H, Fn fn) {
  unsigned hw = std::thread::hardware_concurrency();
  int n = (int)std::max(1u, std::min(hw, 16u));
  if (n == 1) { fn(0, H); return; }
  std::vector<std::thread> th;
  const int chunk = (H + n - 1) / n;
  for (int k = 0; k < n; ++k) {
    const int r0 = k * chunk, r1 = std::min(H, r0 + chunk);
    if (r0 >= r1) break;
    th.emplace_back([=] { fn(r0, r1); });
  }
  for (auto& x : th) x.join();
}

uint16_t quantise(const vh_scene& s, int f, int pix, double z) {
  if (!(z < 1e29)) return 0;
  const uint64_t base = splitmix64(splitmix64(s.seed + 1ull) + (uint64_t)f);
  const uint64_t h = splitmix64(base + (uint64_t)pix * 4ull);
  if (unit(splitmix64(h + 1)) < s.dropout) return 0;
  double u1 = unit(splitmix64(h + 2));
  const double u2 = unit(splitmix64(h + 3));
  u1 = 1.0 - u1;  // (0,1]
  const double g = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
  const double sigma = 0.0012 + 0.0019 * (z - 0.4) * (z - 0.4);
  const double mm = std::nearbyint((z + sigma * g) * 1000.0);
  if (mm <= 0.0) return 0;
  if (mm >= 65535.0) return 65535;
  return (uint16_t)mm;
}

}  // namespace

extern "C" {

int vh_scene_create(int32_t preset, uint64_t seed, vh_scene** out) {
  if (!out) return 1;
  *out = nullptr;
  vh_scene* s = new (std::nothrow) vh_scene();
  if (!s) return 2;
  s->preset = preset;
  s->seed = seed;
  s->orbit_r = 0.02 / (kPi / 180.0);
  switch (preset) {
    case 1:
      s->ext[0] = 4.0; s->ext[1] = 3.0; s->ext[2] = 2.5;
      s->n_frames = 40;
      s->dropout = 0.01;
      place_primitives(*s, 3, 2, 0.3);
      break;
    case 2:
      s->ext[0] = 8.0; s->ext[1] = 6.0; s->ext[2] = 3.0;
      s->n_frames = 500;
      s->dropout = 0.01;
      place_primitives(*s, 6, 6, 0.3);
      break;
    case 5:
      s->ext[0] = 20.0; s->ext[1] = 20.0; s->ext[2] = 4.0;
      s->n_frames = 2000;
      s->dropout = 0.015;
      s->loop_path = true;
      add_walls(*s);
      place_primitives(*s, 20, 20, 0.4);
      break;
    default:
      delete s;
      return 1;
  }
  *out = s;
  return 0;
}

int vh_scene_destroy(vh_scene* s) {
  delete s;
  return 0;
}

int vh_scene_info_get(const vh_scene* s, vh_scene_info* o) {
  if (!s || !o) return 1;
  o->width = s->W; o->height = s->H; o->n_frames = s->n_frames;
  o->fx = (float)s->fx; o->fy = (float)s->fy; o->cx = (float)s->cx; o->cy = (float)s->cy;
  for (int a = 0; a < 3; ++a) o->extent[a] = (float)s->ext[a];
  o->n_spheres = (int)s->spheres.size();
  o->n_boxes = s->n_prim_boxes;
  o->dropout = (float)s->dropout;
  return 0;
}

int vh_scene_pose(const vh_scene* s, int32_t f, float c2w[16]) {
  if (!s || !c2w || f < 0) return 1;
  pose_of(*s, f, c2w);
  return 0;
}

int vh_scene_render_u16(const vh_scene* s, int32_t f, const float* c2w, uint16_t* depth) {
  if (!s || !depth || f < 0) return 1;
  float T[16];
  if (c2w) std::memcpy(T, c2w, sizeof T); else pose_of(*s, f, T);
  parallel_rows(s->H, [&](int r0, int r1) {
    render_rows(*s, f, T, r0, r1, [&](int pix, double z, int, const double*) { depth[pix] = quantise(*s, f, pix, z); });
  });
  return 0;
}

int vh_scene_render_clean(const vh_scene* s, int32_t f, const float* c2w, float* depth) {
  if (!s || !depth || f < 0) return 1;
  float T[16];
  if (c2w) std::memcpy(T, c2w, sizeof T); else pose_of(*s, f, T);
  parallel_rows(s->H, [&](int r0, int r1) {
    render_rows(*s, f, T, r0, r1,
                [&](int pix, double z, int, const double*) { depth[pix] = z < 1e29 ? (float)z : 0.0f; });
  });
  return 0;
}

int vh_scene_render_rgb(const vh_scene* s, int32_t f, const float* c2w, uint8_t* rgb) {
  if (!s || !rgb || f < 0) return 1;
  float T[16];
  if (c2w) std::memcpy(T, c2w, sizeof T); else pose_of(*s, f, T);
  parallel_rows(s->H, [&](int r0, int r1) {
    render_rows(*s, f, T, r0, r1,
                [&](int pix, double, int id, const double* p) { surface_rgb(*s, id, p, rgb + 3 * (size_t)pix); });
  });
  return 0;
}

int vh_scene_render_range(const vh_scene* s, int32_t f0, int32_t n, uint16_t* depth, float* c2w) {
  if (!s || !depth || f0 < 0 || n < 0) return 1;
  for (int k = 0; k < n; ++k) {
    float T[16];
    pose_of(*s, f0 + k, T);
    if (c2w) std::memcpy(c2w + 16 * (size_t)k, T, sizeof T);
    int rc = vh_scene_render_u16(s, f0 + k, T, depth + (size_t)k * s->W * s->H);
    if (rc) return rc;
  }
  return 0;
}

}  // extern "C"
