// Host check of vertex_end (csrc/vh_math.h): for random cubes (lattice origins up to 2^20,
// corner values spread over the snap threshold, exact zeros, near-equal pairs and tiny
// values), every pair of cube edges gets equal vertex_interp positions exactly when
// vertex_end puts both vertices on the same shared corner. Prints "ok <pairs> <coincident>"
// or the first counterexample.
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "vh_math.h"

using namespace vh;

int main() {
  std::mt19937_64 rng(20240611);
  std::uniform_real_distribution<float> uni(-1.0f, 1.0f);
  const float scales[] = {1.0f, 1e-3f, 2e-5f, 1e-5f, 1.2e-5f, 1e-7f, 1e-30f};
  int64_t pairs = 0, coincident = 0;
  for (int it = 0; it < 400000; ++it) {
    const int span = 1 << (int)(rng() % 21);
    const int gx = (int)(rng() % span) - span / 2, gy = (int)(rng() % span) - span / 2, gz = (int)(rng() % span);
    float val[8];
    for (int c = 0; c < 8; ++c) {
      const int kind = (int)(rng() % 8);
      val[c] = kind == 0 ? 0.0f : uni(rng) * scales[rng() % 7];
      if (kind == 1 && c > 0) val[c] = val[c - 1] + uni(rng) * 1e-5f;  // near-equal neighbour
    }
    float P[12][3];
    int end[12], corner[12];
    for (int e = 0; e < 12; ++e) {
      const int a = edge_a(e), b = edge_b(e);
      const float pa[3] = {(float)(gx + corner_dx(a)), (float)(gy + corner_dy(a)), (float)(gz + corner_dz(a))};
      const float pb[3] = {(float)(gx + corner_dx(b)), (float)(gy + corner_dy(b)), (float)(gz + corner_dz(b))};
      vertex_interp(pa[0], pa[1], pa[2], pb[0], pb[1], pb[2], val[a], val[b], P[e][0], P[e][1], P[e][2]);
      const int ax = corner_dx(a) != corner_dx(b) ? 0 : (corner_dy(a) != corner_dy(b) ? 1 : 2);
      end[e] = vertex_end(pa[ax], pb[ax], val[a], val[b]);
      corner[e] = end[e] == 0 ? a : (end[e] == 1 ? b : -1);
    }
    for (int e = 0; e < 12; ++e)
      for (int f = e + 1; f < 12; ++f) {
        const bool same = P[e][0] == P[f][0] && P[e][1] == P[f][1] && P[e][2] == P[f][2];
        const bool pred = corner[e] >= 0 && corner[e] == corner[f];
        ++pairs;
        coincident += same;
        if (same != pred) {
          std::printf("mismatch it %d edges %d %d origin %d %d %d\n", it, e, f, gx, gy, gz);
          return 1;
        }
      }
  }
  std::printf("ok %" PRId64 " %" PRId64 "\n", pairs, coincident);
  return 0;
}
