// vh_tables.cpp - the marching-cubes tables in the reference's unpacked form, for callers of
// MarchingCubeParam (include/tsdf.cuh:243-534), built from this library's packed table.
#include <cstring>

#include "vh.h"
#include "vh_mc_table.h"

extern "C" int vh_mc_tables(int32_t* edge_table, int32_t* tri_table) {
  for (int c = 0; c < 256; ++c) {
    const uint64_t w = vh::kTriTable[c];
    const int n = 3 * vh::tri_count(w);
    int32_t mask = 0;
    for (int j = 0; j < 16; ++j) {
      const int e = j < n ? (int)((w >> (4 * j)) & 15u) : -1;
      if (tri_table) tri_table[16 * c + j] = e;
      if (e >= 0) mask |= 1 << e;
    }
    // an edge is listed iff its end corners differ in sign, so the listed edges are the mask
    if (edge_table) edge_table[c] = mask;
  }
  return VH_OK;
}
