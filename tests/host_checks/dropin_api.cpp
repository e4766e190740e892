// Caller of every public member and type of the C++ drop-in (host/GpuTsdfGenerator.hpp), as a
// user of the reference's ark::GpuTsdfGenerator would write it. Without a GPU it checks what
// needs none (the types, the marching-cubes tables) and exits 3 when the constructor reports
// that no device is present; with a GPU it integrates three frames of a plane 1 m in front of
// the camera and writes <dir>/plane.ply and <dir>/plane.tsdf.
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "GpuTsdfGenerator.hpp"

struct GridSize {
  unsigned x, y, z;
};

static int fail(const char* what) {
  std::fprintf(stderr, "check failed: %s\n", what);
  return 1;
}

int main(int argc, char** argv) {
  const std::string dir = argc > 1 ? argv[1] : ".";
  int32_t edge[256], tri[256 * 16];
  if (vh_mc_tables(edge, tri) != VH_OK) return fail("vh_mc_tables");
  if (edge[1] != 0x109 || edge[255] != 0 || tri[16] != 0 || tri[17] != 8 || tri[18] != 3 || tri[19] != -1)
    return fail("table entries");
  ark::Vertex a, b(1.0f, 2.0f, 3.0f);
  a = b;
  ark::VertexEqual eq;
  ark::VertexHasher hs;
  if (!eq(a, b) || !(a == b) || b.r != 0) return fail("Vertex");
  if (hs(b) != (((size_t)1 * 73856093) ^ ((size_t)2 * 19349669) ^ ((size_t)3 * 83492791))) return fail("VertexHasher");
  ark::Triangle t;
  if (t.valid) return fail("Triangle");
  ark::Face face{{0, 1, 2}};
  (void)face;
  try {
    ark::GpuTsdfGenerator gen(640, 480, 525.0f, 525.0f, 319.5f, 239.5f, 3.0f, -1.5f, -1.5f, 0.5f, 0.01f, 0.04f, 256,
                              256, 256, true);
    ark::MarchingCubeParam* p = gen.getMarchingCubeParam();
    if (std::memcmp(p->edgeTable, edge, sizeof edge) || std::memcmp(&p->triTable[0][0], tri, sizeof tri))
      return fail("MarchingCubeParam tables");
    if (p->vox_dim.x != 256 || p->total_vox != 256 * 256 * 256 || p->vox_origin.z != 0.5f || p->block_size != 0.08f ||
        p->im_width != 640 || p->fx != 525.0f || p->idxMap[2][0] != 1 || p->idxMap[2][1] != 1)
      return fail("MarchingCubeParam fields");
    std::vector<float> depth(640 * 480, 1.0f);
    std::vector<unsigned char> rgb(640 * 480 * 3);
    for (size_t i = 0; i < rgb.size(); ++i) rgb[i] = (unsigned char)(i * 7 % 251);
    float c2w[16] = {1, 0, 0, 1.28f, 0, 1, 0, 1.28f, 0, 0, 1, 0, 0, 0, 0, 1};
    gen.getLocalGrid();
    gen.insert_tri();
    gen.render();
    gen.setStreaming(false);
    for (int k = 0; k < 3; ++k) gen.processFrame(depth.data(), rgb.data(), c2w);
    gen.SavePLY(dir + "/plane.ply");
    gen.SaveTSDF(dir + "/plane.tsdf");
    gen.extract();
    std::vector<ark::Vertex>* V = gen.getVertices();
    std::vector<ark::Face>* F = gen.getFaces();
    if (!gen.volume() || V->empty() || F->empty()) return fail("mesh");
    const GridSize g{256, 256, 256};
    if (!gen.getHashMap()->empty()) return fail("hash map filled before buildHashMap");
    gen.buildHashMap(g, 0.01f);
    size_t found = 0;
    for (size_t i = 0; i < V->size(); ++i) {
      const int j = gen.find_vertex((*V)[i], g, 0.01f, *gen.getHashMap());
      if (j >= 0) {
        if (!((*V)[(size_t)j] == (*V)[i])) return fail("find_vertex returned another position");
        ++found;
      }
    }
    if (gen.find_vertex(ark::Vertex(-1.0f, 0.5f, 0.5f), g, 0.01f, *gen.getHashMap()) != -1)
      return fail("find_vertex outside the grid");
    std::printf("ok vertices %zu faces %zu found %zu\n", V->size(), F->size(), found);
    gen.Shutdown();
  } catch (const std::runtime_error& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return std::strstr(e.what(), "no HIP device") ? 3 : 2;
  }
  return 0;
}
