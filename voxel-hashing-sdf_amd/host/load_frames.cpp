// load_frames.cpp - example caller of the drop-in ark::GpuTsdfGenerator, shaped like the
// reference's application loop (src/main.cpp:236-283: load a frame, PushFrame -> Reproject ->
// processFrame, then SavePly; src/PointCloudGenerator.cpp:19-49,93-143).
//
// usage: load_frames <out.ply> [n_frames] [--max-depth <m>] [--stream]
//          frames (depth and colour) from the synthetic small-room generator
//          (include/vh_synth.h); --max-depth sets the depth range (default 5 m), --stream
//          turns on the generator's out-of-core streaming (setStreaming, stream-out every
//          frame with no margin) and prints the blocks streamed
//        load_frames <out.ply> --dataset <dir> <settings.yaml> [first] [last]
//          frames <dir>/depth/<id>.png (uint16) and <dir>/tcw/<id>.txt (camera-to-world) for
//          id = first.. until last or the first missing depth image (SaveFrame::frameLoad,
//          src/SaveFrame.cpp:154-218); camera and voxel parameters from the settings file
//          (the keys PointCloudGenerator reads); depth in metres = value / DepthMapFactor
//          (the reference multiplies by 0.001, its DepthMapFactor of 1000); the colour
//          images (<dir>/RGB/<id>.jpg) are not decoded, so this mode integrates depth only
// exit status: 0 ok, 2 usage, 3 no GPU, 1 other failure
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "GpuTsdfGenerator.hpp"
#include "vh_io.h"
#include "vh_synth.h"

namespace {

int usage(const char* argv0) {
  std::fprintf(stderr,
               "usage: %s <out.ply> [n_frames] [--max-depth <m>] [--stream]\n"
               "       %s <out.ply> --dataset <dir> <settings.yaml> [first] [last]\n",
               argv0, argv0);
  return 2;
}

int run_synthetic(const char* out, int n_arg, float max_depth, bool stream) {
  vh_scene* scene = nullptr;
  if (vh_scene_create(VH_SCENE_ROOM_SMALL, 20240611ull, &scene) != 0) return 1;
  vh_scene_info info;
  vh_scene_info_get(scene, &info);
  const int n = n_arg > 0 ? n_arg : info.n_frames;
  int rc = 0;
  try {
    // voxel 1 cm, truncation 4 cm, as the small-room workload
    ark::GpuTsdfGenerator gen(info.width, info.height, info.fx, info.fy, info.cx, info.cy, max_depth, -1.5f, -1.5f,
                              0.5f, 0.01f, 0.04f, 500, 500, 500);
    if (stream) gen.setStreaming(true, 1, 0.0f);
    std::vector<uint16_t> mm((size_t)info.width * info.height);
    std::vector<float> depth(mm.size());
    std::vector<unsigned char> rgb(mm.size() * 3);
    for (int f = 0; f < n; ++f) {
      float c2w[16];
      vh_scene_pose(scene, f, c2w);
      vh_scene_render_u16(scene, f, c2w, mm.data());
      vh_scene_render_rgb(scene, f, c2w, rgb.data());
      for (size_t i = 0; i < mm.size(); ++i) depth[i] = (float)mm[i] * 0.001f;  // as SaveFrame.cpp:179-180
      gen.processFrame(depth.data(), rgb.data(), c2w);
    }
    vh_stats st;
    if (vh_get_stats(gen.volume(), &st) != 0) throw std::runtime_error(vh_last_error());
    gen.SavePLY(out);
    std::printf("frames %d vertices %zu faces %zu streamed_out %lld\n", n, gen.getVertices()->size(),
                gen.getFaces()->size(), (long long)st.streamed_out_total);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "load_frames: %s\n", e.what());
    rc = std::strstr(e.what(), "no HIP device") ? 3 : 1;
  }
  vh_scene_destroy(scene);
  return rc;
}

int run_dataset(const char* out, const char* dir, const char* settings, int first, int last) {
  vh_settings s;
  if (vh_settings_load(settings, &s) != VH_OK) {
    std::fprintf(stderr, "load_frames: %s\n", vh_last_error());
    return 1;
  }
  const uint32_t need = VH_SET_FX | VH_SET_FY | VH_SET_CX | VH_SET_CY | VH_SET_WIDTH | VH_SET_HEIGHT |
                        VH_SET_MAX_DEPTH | VH_SET_VOX_SIZE | VH_SET_TRUNC;
  if ((s.found & need) != need || s.width <= 0 || s.height <= 0) {
    std::fprintf(stderr, "load_frames: %s lacks a camera, MaxDepth or voxel key\n", settings);
    return 1;
  }
  const float scale = (s.found & VH_SET_DEPTH_FACTOR) && s.depth_factor > 0.0f ? 1.0f / s.depth_factor : 0.001f;
  try {
    // the reference's constructor call (PointCloudGenerator.cpp:43-45)
    ark::GpuTsdfGenerator gen(s.width, s.height, s.fx, s.fy, s.cx, s.cy, s.max_depth, s.origin[0], s.origin[1],
                              s.origin[2], s.vox_size, s.trunc, s.vox_dim[0], s.vox_dim[1], s.vox_dim[2],
                              /*color=*/false);
    std::vector<uint16_t> mm((size_t)s.width * s.height);
    std::vector<float> depth(mm.size());
    int n = 0;
    for (int id = first; last < 0 || id <= last; ++id) {
      float c2w[16];
      const int rc = vh_frame_load(dir, id, mm.data(), s.width, s.height, c2w);
      if (rc == VH_ERR_IO && id > first) break;  // end of the sequence
      if (rc != VH_OK) throw std::runtime_error(vh_last_error());
      for (size_t i = 0; i < mm.size(); ++i) depth[i] = (float)mm[i] * scale;
      gen.processFrame(depth.data(), nullptr, c2w);
      ++n;
    }
    gen.SavePLY(out);
    std::printf("frames %d vertices %zu faces %zu\n", n, gen.getVertices()->size(), gen.getFaces()->size());
  } catch (const std::exception& e) {
    std::fprintf(stderr, "load_frames: %s\n", e.what());
    return std::strstr(e.what(), "no HIP device") ? 3 : 1;
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) return usage(argv[0]);
  if (argc >= 3 && std::strcmp(argv[2], "--dataset") == 0) {
    if (argc < 5) return usage(argv[0]);
    const int first = argc > 5 ? std::atoi(argv[5]) : 0;
    const int last = argc > 6 ? std::atoi(argv[6]) : -1;
    return run_dataset(argv[1], argv[3], argv[4], first, last);
  }
  int n = 0;
  float max_depth = 5.0f;
  bool stream = false;
  for (int i = 2; i < argc; ++i) {
    if (std::strcmp(argv[i], "--stream") == 0) stream = true;
    else if (std::strcmp(argv[i], "--max-depth") == 0 && i + 1 < argc) max_depth = (float)std::atof(argv[++i]);
    else if (argv[i][0] != '-') n = std::atoi(argv[i]);
    else return usage(argv[0]);
  }
  return run_synthetic(argv[1], n, max_depth, stream);
}
