// GpuTsdfGenerator.hpp - header-only C++ drop-in for the reference's ark::GpuTsdfGenerator
// (include/tsdf.cuh:548-683), implemented over the C ABI of libvh (include/vh.h).
//
// Kept from the reference: the namespace, the 15-argument constructor (tsdf.cuh:605-607;
// its defaults, which the reference only gives on the out-of-line definition at
// src/tsdf.cu:1249-1252, are on the declaration here), processFrame(depth, rgb, c2w)
// (tsdf.cuh:609-610) with metric float depth and a row-major camera-to-world matrix,
// SavePLY(filename) writing an ASCII PLY of shared vertices and triangle faces
// (tsdf2mesh, src/tsdf.cu:1760-1888), render(), Shutdown(), getVertices()/getFaces().
// Colour is integrated as the reference does (each frame's rgb, vertex colours in the PLY);
// a trailing constructor argument color = false turns it off for depth-only callers.
// Changed: errors throw std::runtime_error (the reference exits the process, tsdf.cuh:46-61),
// which includes processFrame with a null rgb on a colour volume (the reference reads it);
// render() draws nothing (no GL in this library); the mesh comes from marching cubes over
// every allocated block at the time of the call, indexed on the GPU: one vertex per crossed
// lattice edge, shared by every face that uses the edge (the reference welds float-identical
// coordinates on the host instead), and the faces keep the zero-area triangles that appear
// where the surface passes exactly through a lattice point, so closed surfaces stay closed.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "vh.h"

namespace ark {

struct Vertex {
  float x, y, z;
  unsigned char r, g, b;
};

struct Face {
  int vIdx[3];
};

class GpuTsdfGenerator {
 public:
  GpuTsdfGenerator(int width, int height, float fx, float fy, float cx, float cy, float max_depth,
                   float origin_x = -1.5f, float origin_y = -1.5f, float origin_z = 0.5f, float vox_size = 0.006f,
                   float trunc_m = 0.03f, int vox_dim_x = 500, int vox_dim_y = 500, int vox_dim_z = 500,
                   bool color = true) {
    vh_config c;
    vh_config_default(&c);
    c.width = width;
    c.height = height;
    c.fx = fx;
    c.fy = fy;
    c.cx = cx;
    c.cy = cy;
    c.max_depth = max_depth;
    c.origin[0] = origin_x;
    c.origin[1] = origin_y;
    c.origin[2] = origin_z;
    c.vox_size = vox_size;
    c.trunc = trunc_m;
    c.vox_dim[0] = vox_dim_x;
    c.vox_dim[1] = vox_dim_y;
    c.vox_dim[2] = vox_dim_z;
    c.log2_buckets = 20;
    c.pool_blocks = 1 << 19;
    c.color = color ? 1 : 0;
    open(c);
  }
  // Full control over hash and pool sizes and the voxel storage format.
  explicit GpuTsdfGenerator(const vh_config& cfg) { open(cfg); }

  GpuTsdfGenerator(const GpuTsdfGenerator&) = delete;
  GpuTsdfGenerator& operator=(const GpuTsdfGenerator&) = delete;
  ~GpuTsdfGenerator() { Shutdown(); }

  // Out-of-core operation, the role of the reference's per-frame chunk streaming
  // (streamInCPU2GPU / streamOutGPU2CPU, src/tsdf.cu:277-596), off by default. When on,
  // every processFrame first brings into the GPU pool every host-store block within reach of
  // its camera (the farthest point its truncation bands can touch: (max_depth + trunc) times
  // the image corner's ray length per unit depth, plus a block diagonal), and every
  // `every_frames` frames moves the blocks farther than reach + margin_m to the host store.
  // extract() and SavePLY bring everything back first. The mesh is the one the volume would
  // have without streaming.
  void setStreaming(bool on, int every_frames = 8, float margin_m = 0.5f) {
    vh_config c;
    check(vh_get_config(vol_, &c), "setStreaming");
    const double dx = std::fmax(c.cx, c.width - 1 - c.cx) / c.fx, dy = std::fmax(c.cy, c.height - 1 - c.cy) / c.fy;
    reach_ = (float)((c.max_depth + c.trunc) * std::sqrt(1.0 + dx * dx + dy * dy) + 8.0 * c.vox_size * std::sqrt(3.0));
    stream_ = on;
    stream_every_ = every_frames > 0 ? every_frames : 1;
    stream_margin_ = margin_m > 0.0f ? margin_m : 0.0f;
    frames_since_out_ = 0;
  }

  // depth: height x width metres (0 = invalid); rgb: height x width x 3 bytes (ignored, and
  // may be null, when colour is off); c2w: row-major 4x4 camera-to-world.
  void processFrame(float* depth, unsigned char* rgb, float* c2w) {
    if (stream_) {
      const float cam[3] = {c2w[3], c2w[7], c2w[11]};
      int64_t n = 0;
      if (++frames_since_out_ >= stream_every_) {
        check(vh_stream_out(vol_, cam, reach_ + stream_margin_, &n), "processFrame (stream out)");
        frames_since_out_ = 0;
      }
      check(vh_stream_in(vol_, cam, reach_, &n), "processFrame (stream in)");
    }
    vh_frame f;
    std::memset(&f, 0, sizeof f);
    f.depth = depth;
    f.depth_format = VH_DEPTH_F32;
    f.depth_scale = 1.0f;
    f.pitch = 0;
    f.rgb = rgb;
    std::memcpy(f.c2w, c2w, sizeof f.c2w);
    check(vh_integrate(vol_, &f), "processFrame");
  }

  // ASCII PLY: "x y z r g b" vertices (metres; colour 0 when colour is off), then
  // "3 i j k" faces.
  void SavePLY(std::string filename) {
    extract();
    FILE* fp = std::fopen(filename.c_str(), "w");
    if (!fp) throw std::runtime_error("SavePLY: cannot open " + filename);
    std::fprintf(fp,
                 "ply\nformat ascii 1.0\ncomment vh marching cubes\nelement vertex %zu\nproperty float x\n"
                 "property float y\nproperty float z\nproperty uchar red\nproperty uchar green\n"
                 "property uchar blue\nelement face %zu\nproperty list uchar int vertex_index\nend_header\n",
                 vertices_.size(), faces_.size());
    for (const Vertex& v : vertices_)
      std::fprintf(fp, "%.9g %.9g %.9g %d %d %d\n", v.x, v.y, v.z, (int)v.r, (int)v.g, (int)v.b);
    for (const Face& f : faces_) std::fprintf(fp, "3 %d %d %d\n", f.vIdx[0], f.vIdx[1], f.vIdx[2]);
    if (std::fclose(fp) != 0) throw std::runtime_error("SavePLY: write failed");
  }

  void render() {}

  void Shutdown() {
    if (vol_) {
      vh_destroy(vol_);
      vol_ = nullptr;
    }
  }

  // The mesh of the last SavePLY/extract call.
  std::vector<Vertex>* getVertices() { return &vertices_; }
  std::vector<Face>* getFaces() { return &faces_; }

  // Recomputes the indexed mesh from the current volume (all of it: streamed-out blocks
  // come back first).
  void extract() {
    if (stream_) {
      const float origin[3] = {0.0f, 0.0f, 0.0f};
      int64_t n = 0;
      check(vh_stream_in(vol_, origin, INFINITY, &n), "extract (stream in)");
    }
    vh_indexed_mesh m;
    check(vh_extract_mesh_indexed(vol_, &m), "extract mesh");
    vertices_.resize((size_t)m.n_vertices);
    faces_.resize((size_t)m.n_faces);
    for (int64_t i = 0; i < m.n_vertices; ++i) {
      const float* p = m.vertices + 3 * i;
      const uint8_t* c = m.rgb ? m.rgb + 3 * i : nullptr;
      vertices_[(size_t)i] = Vertex{p[0], p[1], p[2], c ? c[0] : (unsigned char)0, c ? c[1] : (unsigned char)0,
                                    c ? c[2] : (unsigned char)0};
    }
    for (int64_t t = 0; t < m.n_faces; ++t)
      for (int j = 0; j < 3; ++j) faces_[(size_t)t].vIdx[j] = m.faces[3 * t + j];
    vh_indexed_mesh_free(&m);
  }

  vh_volume* volume() { return vol_; }

 private:
  void open(const vh_config& c) { check(vh_create(&c, &vol_), "GpuTsdfGenerator"); }
  static void check(int rc, const char* what) {
    if (rc != VH_OK) throw std::runtime_error(std::string(what) + ": " + vh_last_error());
  }

  vh_volume* vol_ = nullptr;
  bool stream_ = false;
  int stream_every_ = 8, frames_since_out_ = 0;
  float reach_ = 0.0f, stream_margin_ = 0.5f;
  std::vector<Vertex> vertices_;
  std::vector<Face> faces_;
};

}  // namespace ark
