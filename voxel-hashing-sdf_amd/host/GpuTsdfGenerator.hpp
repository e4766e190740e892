// GpuTsdfGenerator.hpp - header-only C++ drop-in for the reference's ark::GpuTsdfGenerator
// (include/tsdf.cuh:548-683), implemented over the C ABI of libvh (include/vh.h).
//
// Kept from the reference: the namespace, the types Vertex (both constructors), VertexEqual,
// VertexHasher, Triangle, Face and MarchingCubeParam (tsdf.cuh:65-113, 220-534), the
// 15-argument constructor (tsdf.cuh:605-607; its defaults, which the reference only gives on
// the out-of-line definition at src/tsdf.cu:1249-1252, are on the declaration here),
// processFrame(depth, rgb, c2w) (tsdf.cuh:609-610) with metric float depth and a row-major
// camera-to-world matrix, SavePLY(filename) writing an ASCII PLY of shared vertices and
// triangle faces (tsdf2mesh, src/tsdf.cu:1760-1888), SaveTSDF(filename) (src/tsdf.cu:1663-1695),
// render(), Shutdown(), getVertices()/getFaces(), getHashMap(), getMarchingCubeParam(),
// find_vertex() (src/tsdf.cu:1891-1929), getLocalGrid() and insert_tri(), whose bodies in the
// reference are commented out (src/tsdf.cu:1416-1482, 1601-1617), as no-ops.
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
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <list>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "vh.h"

namespace ark {

// A mesh vertex in metres with its colour. The (x, y, z) constructor sets the colour to 0
// (the reference leaves it unset).
struct Vertex {
  float x, y, z;
  unsigned char r, g, b;
  Vertex() {}
  Vertex(float xi, float yi, float zi) : x(xi), y(yi), z(zi), r(0), g(0), b(0) {}
};

inline bool operator==(const Vertex& a, const Vertex& b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

struct VertexEqual {
  bool operator()(Vertex v1, Vertex v2) const { return v1 == v2; }
};

// The reference's spatial hash of a vertex (include/tsdf.cuh:87-99): integer parts of the
// coordinates times three primes, combined by xor.
struct VertexHasher {
  size_t operator()(Vertex v) const {
    return ((size_t)v.x * (size_t)73856093) ^ ((size_t)v.y * (size_t)19349669) ^ ((size_t)v.z * (size_t)83492791);
  }
};

struct Triangle {
  Vertex p[3];
  bool valid;
  Triangle() : valid(false) {}
};

struct Face {
  int vIdx[3];
};

struct Float3 {
  float x, y, z;
};
struct Int3 {
  int x, y, z;
};

// The grid and camera parameters and the marching-cubes tables (tsdf.cuh:220-534). The
// tables come from the library (vh_mc_tables), the corner order of idxMap is the one every
// table here uses. vox_dim / total_vox describe the dense grid of SaveTSDF.
class MarchingCubeParam {
 public:
  Float3 vox_origin;
  float vox_size;
  float trunc_margin;
  Int3 vox_dim;
  int total_vox;
  float max_depth;
  float min_depth;
  float block_size;
  int im_width;
  int im_height;
  float fx, fy, cx, cy;
  int idxMap[8][3] = {{0, 0, 0}, {0, 1, 0}, {1, 1, 0}, {1, 0, 0}, {0, 0, 1}, {0, 1, 1}, {1, 1, 1}, {1, 0, 1}};
  int edgeTable[256];
  int triTable[256][16];
};

// Vertex lists per cell, as getHashMap() returns them.
using VertexHashMap = std::vector<std::list<std::pair<Vertex, int>>>;

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
    c.carve = 1;  // the reference integrates every allocated block along its pixel rays
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

  // The dense voxel grid of the constructor's vox_dim, origin and vox_size, in the
  // reference's file layout (src/tsdf.cu:1663-1695): 8 floats (vox_dim x, y, z as floats,
  // origin x, y, z, vox_size, trunc_m), then vox_dim.x * y * z TSDF floats (x fastest, then y,
  // then z), then 3 bytes of colour per voxel in the same order. Dense voxel (i, j, k) is the
  // volume's lattice voxel round(origin / vox_size) + (i, j, k); voxels that are not
  // allocated or never observed hold 1 (free space) and colour 0. Streamed-out blocks come
  // back first.
  void SaveTSDF(std::string filename) {
    stream_all_in("SaveTSDF");
    const MarchingCubeParam& p = param_;
    const int64_t nx = p.vox_dim.x, ny = p.vox_dim.y, nz = p.vox_dim.z;
    if (nx <= 0 || ny <= 0 || nz <= 0) throw std::runtime_error("SaveTSDF: empty grid");
    int64_t nb = 0;
    int rc = vh_export_blocks(vol_, nullptr, nullptr, nullptr, 0, &nb);
    if (rc != VH_OK && rc != VH_ERR_TRUNCATED) check(rc, "SaveTSDF");
    std::vector<int32_t> keys((size_t)nb * 3);
    std::vector<float> tsdf((size_t)nb * 512), wt((size_t)nb * 512);
    check(vh_export_blocks(vol_, keys.data(), tsdf.data(), wt.data(), nb, &nb), "SaveTSDF");
    std::vector<uint8_t> rgb;
    if (color_) {
      rgb.resize((size_t)nb * 512 * 3);
      check(vh_export_block_colors(vol_, rgb.data(), nb, &nb), "SaveTSDF");
    }
    const int64_t n = nx * ny * nz;
    std::vector<float> grid((size_t)n, 1.0f);
    std::vector<uint8_t> cgrid((size_t)n * 3, 0);
    const int64_t o[3] = {(int64_t)std::lround(p.vox_origin.x / p.vox_size),
                          (int64_t)std::lround(p.vox_origin.y / p.vox_size),
                          (int64_t)std::lround(p.vox_origin.z / p.vox_size)};
    for (int64_t b = 0; b < nb; ++b)
      for (int i = 0; i < 512; ++i) {
        const int64_t x = (int64_t)keys[3 * b] * 8 + (i >> 6) - o[0], y = (int64_t)keys[3 * b + 1] * 8 + ((i >> 3) & 7) - o[1],
                      z = (int64_t)keys[3 * b + 2] * 8 + (i & 7) - o[2];
        if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz || !(wt[(size_t)b * 512 + i] > 0.0f)) continue;
        const int64_t d = (z * ny + y) * nx + x;
        grid[(size_t)d] = tsdf[(size_t)b * 512 + i];
        if (color_) std::memcpy(&cgrid[(size_t)d * 3], &rgb[((size_t)b * 512 + i) * 3], 3);
      }
    FILE* fp = std::fopen(filename.c_str(), "wb");
    if (!fp) throw std::runtime_error("SaveTSDF: cannot open " + filename);
    const float head[8] = {(float)nx, (float)ny, (float)nz, p.vox_origin.x, p.vox_origin.y, p.vox_origin.z,
                           p.vox_size, p.trunc_margin};
    bool ok = std::fwrite(head, sizeof(float), 8, fp) == 8;
    ok = ok && std::fwrite(grid.data(), sizeof(float), (size_t)n, fp) == (size_t)n;
    ok = ok && std::fwrite(cgrid.data(), 1, (size_t)n * 3, fp) == (size_t)n * 3;
    ok = (std::fclose(fp) == 0) && ok;
    if (!ok) throw std::runtime_error("SaveTSDF: write failed");
  }

  void render() {}
  // No-ops: the reference's bodies are commented out.
  void getLocalGrid() {}
  void insert_tri() {}

  void Shutdown() {
    if (vol_) {
      vh_destroy(vol_);
      vol_ = nullptr;
    }
  }

  // The mesh of the last SavePLY/extract call.
  std::vector<Vertex>* getVertices() { return &vertices_; }
  std::vector<Face>* getFaces() { return &faces_; }

  // Vertex lists per cell for find_vertex. The reference returns a member that nothing fills;
  // here it stays empty until buildHashMap fills it from the last extracted mesh.
  VertexHashMap* getHashMap() { return &hash_map_; }
  MarchingCubeParam* getMarchingCubeParam() { return &param_; }

  // Fills getHashMap()'s table with every vertex of the last extracted mesh, in the cell of
  // calc_cell_hash(floor(p / cell_size), grid_size) (src/tsdf.cu:1891-1914: cells outside
  // the grid are dropped; coordinates are masked by grid_size - 1, so grid sizes should be
  // powers of two). The table has grid_size.x * y * z lists.
  template <class U3>
  void buildHashMap(U3 grid_size, float cell_size) {
    hash_map_.assign((size_t)grid_size.x * grid_size.y * grid_size.z, {});
    for (size_t i = 0; i < vertices_.size(); ++i) {
      const unsigned key = cell_hash(vertices_[i], grid_size, cell_size);
      if (key != 0xffffffffu) hash_map_[key].push_back(std::make_pair(vertices_[i], (int)i));
    }
  }

  // Index of the vertex with exactly p's coordinates in hash_table (filled as buildHashMap
  // does), or -1 (src/tsdf.cu:1916-1929). A cell beyond the table's size gives -1 (the
  // reference indexes past its end).
  template <class U3>
  int find_vertex(Vertex p, U3 grid_size, float cell_size, VertexHashMap& hash_table) {
    const unsigned key = cell_hash(p, grid_size, cell_size);
    if (key == 0xffffffffu || key >= hash_table.size()) return -1;
    for (const auto& e : hash_table[key])
      if (e.first == p) return e.second;
    return -1;
  }

  // Recomputes the indexed mesh from the current volume (all of it: streamed-out blocks
  // come back first).
  void extract() {
    stream_all_in("extract (stream in)");
    vh_indexed_mesh m;
    check(vh_extract_mesh_indexed(vol_, &m), "extract mesh");
    vertices_.resize((size_t)m.n_vertices);
    faces_.resize((size_t)m.n_faces);
    for (int64_t i = 0; i < m.n_vertices; ++i) {
      const float* p = m.vertices + 3 * i;
      const uint8_t* c = m.rgb ? m.rgb + 3 * i : nullptr;
      Vertex v(p[0], p[1], p[2]);
      if (c) {
        v.r = c[0];
        v.g = c[1];
        v.b = c[2];
      }
      vertices_[(size_t)i] = v;
    }
    for (int64_t t = 0; t < m.n_faces; ++t)
      for (int j = 0; j < 3; ++j) faces_[(size_t)t].vIdx[j] = m.faces[3 * t + j];
    vh_indexed_mesh_free(&m);
  }

  vh_volume* volume() { return vol_; }

 private:
  void open(const vh_config& c) {
    check(vh_create(&c, &vol_), "GpuTsdfGenerator");
    color_ = c.color != 0;
    param_.vox_origin = Float3{c.origin[0], c.origin[1], c.origin[2]};
    param_.vox_size = c.vox_size;
    param_.trunc_margin = c.trunc;
    param_.vox_dim = Int3{c.vox_dim[0], c.vox_dim[1], c.vox_dim[2]};
    param_.total_vox = c.vox_dim[0] * c.vox_dim[1] * c.vox_dim[2];
    param_.max_depth = c.max_depth;
    param_.min_depth = c.min_depth;
    param_.block_size = 8 * c.vox_size;
    param_.im_width = c.width;
    param_.im_height = c.height;
    param_.fx = c.fx;
    param_.fy = c.fy;
    param_.cx = c.cx;
    param_.cy = c.cy;
    check(vh_mc_tables(param_.edgeTable, &param_.triTable[0][0]), "GpuTsdfGenerator");
  }
  void stream_all_in(const char* what) {
    if (!stream_) return;
    const float origin[3] = {0.0f, 0.0f, 0.0f};
    int64_t n = 0;
    check(vh_stream_in(vol_, origin, INFINITY, &n), what);
  }
  // calc_cell_pos + calc_cell_hash of src/tsdf.cu:1891-1914
  template <class U3>
  static unsigned cell_hash(const Vertex& p, U3 g, float cell_size) {
    const int cx = (int)std::floor(p.x / cell_size), cy = (int)std::floor(p.y / cell_size),
              cz = (int)std::floor(p.z / cell_size);
    if (cx < 0 || cx >= (int)g.x || cy < 0 || cy >= (int)g.y || cz < 0 || cz >= (int)g.z) return 0xffffffffu;
    const unsigned x = (unsigned)cx & (g.x - 1), y = (unsigned)cy & (g.y - 1), z = (unsigned)cz & (g.z - 1);
    return z * g.y * g.x + y * g.x + x;
  }
  static void check(int rc, const char* what) {
    if (rc != VH_OK) throw std::runtime_error(std::string(what) + ": " + vh_last_error());
  }

  vh_volume* vol_ = nullptr;
  bool color_ = false;
  MarchingCubeParam param_;
  VertexHashMap hash_map_;
  bool stream_ = false;
  int stream_every_ = 8, frames_since_out_ = 0;
  float reach_ = 0.0f, stream_margin_ = 0.5f;
  std::vector<Vertex> vertices_;
  std::vector<Face> faces_;
};

}  // namespace ark
