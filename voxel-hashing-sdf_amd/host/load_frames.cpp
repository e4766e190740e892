// load_frames.cpp - example caller of the drop-in ark::GpuTsdfGenerator, shaped like the
// reference's application loop (src/main.cpp:236-283: load a frame, PushFrame -> Reproject ->
// processFrame, then SavePly; src/PointCloudGenerator.cpp:93-143). Frames come from the
// synthetic generator (include/vh_synth.h) instead of RGB/, depth/ and tcw/ directories.
//
// usage: load_frames <out.ply> [n_frames]
// exit status: 0 ok, 2 usage, 3 no GPU, 1 other failure
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "GpuTsdfGenerator.hpp"
#include "vh_synth.h"

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <out.ply> [n_frames]\n", argv[0]);
    return 2;
  }
  vh_scene* scene = nullptr;
  if (vh_scene_create(VH_SCENE_ROOM_SMALL, 20240611ull, &scene) != 0) return 1;
  vh_scene_info info;
  vh_scene_info_get(scene, &info);
  const int n = argc > 2 ? std::atoi(argv[2]) : info.n_frames;
  try {
    // voxel 1 cm, truncation 4 cm, as the small-room workload
    ark::GpuTsdfGenerator gen(info.width, info.height, info.fx, info.fy, info.cx, info.cy, 5.0f, -1.5f, -1.5f,
                              0.5f, 0.01f, 0.04f, 500, 500, 500);
    std::vector<uint16_t> mm((size_t)info.width * info.height);
    std::vector<float> depth(mm.size());
    for (int f = 0; f < n; ++f) {
      float c2w[16];
      vh_scene_pose(scene, f, c2w);
      vh_scene_render_u16(scene, f, c2w, mm.data());
      for (size_t i = 0; i < mm.size(); ++i) depth[i] = (float)mm[i] * 0.001f;  // as SaveFrame.cpp:179-180
      gen.processFrame(depth.data(), nullptr, c2w);
    }
    gen.SavePLY(argv[1]);
    std::printf("frames %d vertices %zu faces %zu\n", n, gen.getVertices()->size(), gen.getFaces()->size());
  } catch (const std::exception& e) {
    std::fprintf(stderr, "load_frames: %s\n", e.what());
    vh_scene_destroy(scene);
    return std::strstr(e.what(), "no HIP device") ? 3 : 1;
  }
  vh_scene_destroy(scene);
  return 0;
}
