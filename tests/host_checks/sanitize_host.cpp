// Host code under AddressSanitizer and UndefinedBehaviorSanitizer (built by
// tests/test_sanitize.py with g++ from the sources of the synthetic frame generator, the
// settings / PNG / pose readers and the CPU oracle): every entry point on valid input, and the
// readers on malformed files (truncated and corrupted PNGs, absurd sizes, other PNG formats,
// broken YAML and pose files). Readers must return an error code, never crash or read out of
// bounds. Prints "ok" at the end; any sanitizer report aborts with a non-zero status.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "vh.h"
#include "vh_io.h"
#include "vh_synth.h"

// The error text store of the C ABI lives in the device library; the readers only call this.
namespace vh {
int set_last_error(int code, const std::string&) { return code; }
}  // namespace vh

extern "C" {
void* vho_create(int W, int H, float fx, float fy, float cx, float cy, float max_depth, float min_depth, float vox,
                 float mu, int storage, int cap);
void vho_destroy(void* p);
int64_t vho_integrate(void* p, const uint16_t* depth, float scale, const float* c2w);
int vho_set_color(void* p, int on);
int64_t vho_integrate_rgb(void* p, const uint16_t* depth, const uint8_t* rgb, float scale, const float* c2w);
int64_t vho_num_blocks(void* p);
void vho_set_carve(void* p, int mode);
int64_t vho_export(void* p, int32_t* xyz, float* s, float* w, double* s64, double* w64, int64_t cap);
int64_t vho_mc(void* p, float* xyz, int32_t* cubes, int64_t cap);
int64_t vho_mc64(void* p, float* xyz, int32_t* cubes, int64_t cap);
int64_t vho_mesh_indexed(void* p, float* vertices, uint8_t* rgb, int64_t vcap, int32_t* faces, int32_t* face_cubes,
                         int64_t fcap, int64_t* n_faces);
void vho_mesh_distance(const float* a_xyz, const int32_t* a_cubes, int64_t na, const float* b_xyz,
                       const int32_t* b_cubes, int64_t nb, double* out, float* dist);
int vho_import_f32(void* p, const int32_t* keys, const float* tsdf, const float* weight, int64_t n);
void vho_query64(void* p, const double* pts, int64_t n, double* sdf, double* grad, uint8_t* valid);
void vho_query_twin(void* p, const double* pts, int64_t n, float* sdf, float* grad, uint8_t* valid);
int64_t vho_scatter_grad(void* p, const double* pts, const float* up, int64_t n, double* grad, int twin);
int vho_band(void* p, const float* c2w, int u, int v, float d, int32_t* xyz, int cap);
}

static int failures = 0;
#define CHECK(c)                                                     \
  do {                                                               \
    if (!(c)) {                                                      \
      std::fprintf(stderr, "check failed line %d: %s\n", __LINE__, #c); \
      ++failures;                                                    \
    }                                                                \
  } while (0)

static std::vector<uint8_t> read_file(const std::string& p) {
  std::vector<uint8_t> b;
  FILE* f = std::fopen(p.c_str(), "rb");
  if (!f) return b;
  int c;
  while ((c = std::fgetc(f)) != EOF) b.push_back((uint8_t)c);
  std::fclose(f);
  return b;
}
static void write_file(const std::string& p, const uint8_t* d, size_t n) {
  FILE* f = std::fopen(p.c_str(), "wb");
  if (n) std::fwrite(d, 1, n, f);
  std::fclose(f);
}
static void write_text(const std::string& p, const std::string& s) {
  write_file(p, reinterpret_cast<const uint8_t*>(s.data()), s.size());
}
static uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

// A PNG must load completely or fail with an error; the pixel buffer is the size the
// caller states.
static int try_png(const std::string& path, int W, int H) {
  int32_t w = 0, h = 0;
  const int rc = vh_png16_info(path.c_str(), &w, &h);
  std::vector<uint16_t> px((size_t)W * H);
  const int rc2 = vh_png16_load(path.c_str(), px.data(), W, H);
  return rc != VH_OK ? rc : rc2;
}

int main(int argc, char** argv) {
  const std::string dir = argc > 1 ? argv[1] : "/tmp";
  const std::string yaml_ok = argc > 2 ? argv[2] : "";

  // synthetic frames: every preset, every renderer
  std::vector<uint16_t> depth(640 * 480 * 3);
  std::vector<float> poses(16 * 3), clean(640 * 480);
  std::vector<uint8_t> rgb(640 * 480 * 3);
  for (int preset : {1, 2, 5}) {
    vh_scene* s = nullptr;
    CHECK(vh_scene_create(preset, 12345, &s) == VH_OK);
    vh_scene_info info;
    CHECK(vh_scene_info_get(s, &info) == VH_OK && info.width == 640);
    float c2w[16];
    CHECK(vh_scene_pose(s, 3, c2w) == VH_OK);
    CHECK(vh_scene_render_u16(s, 3, nullptr, depth.data()) == VH_OK);
    CHECK(vh_scene_render_clean(s, 3, c2w, clean.data()) == VH_OK);
    CHECK(vh_scene_render_rgb(s, 3, nullptr, rgb.data()) == VH_OK);
    CHECK(vh_scene_render_range(s, 0, 3, depth.data(), poses.data()) == VH_OK);
    CHECK(vh_scene_render_u16(s, -1, nullptr, depth.data()) != VH_OK);
    vh_scene_destroy(s);
  }
  vh_scene* bad = nullptr;
  CHECK(vh_scene_create(9, 1, &bad) != VH_OK && !bad);

  // PNG: a round trip, then every truncation and corruption of the file
  const std::string png = dir + "/d.png";
  CHECK(vh_png16_save(png.c_str(), depth.data(), 640, 480) == VH_OK);
  {
    std::vector<uint16_t> back(640 * 480);
    CHECK(vh_png16_load(png.c_str(), back.data(), 640, 480) == VH_OK);
    CHECK(std::memcmp(back.data(), depth.data(), back.size() * 2) == 0);
    CHECK(vh_png16_load(png.c_str(), back.data(), 320, 480) != VH_OK);  // the caller's size disagrees
  }
  const std::vector<uint8_t> good = read_file(png);
  CHECK(good.size() > 100);
  const std::string bad_png = dir + "/bad.png";
  int rejected = 0;
  for (size_t n = 0; n < good.size(); n += 1 + good.size() / 97) {
    write_file(bad_png, good.data(), n);
    rejected += try_png(bad_png, 640, 480) != VH_OK;
  }
  CHECK(rejected > 90);
  uint32_t seed = 1;
  for (int t = 0; t < 200; ++t) {
    std::vector<uint8_t> b = good;
    for (int k = 0; k < 4; ++k) {
      seed = seed * 1664525u + 1013904223u;
      b[8 + seed % (b.size() - 8)] ^= (uint8_t)(1 + (seed >> 24) % 255);
    }
    write_file(bad_png, b.data(), b.size());
    (void)try_png(bad_png, 640, 480);  // any status; the sanitizers judge
  }
  // IHDR fields: absurd size, other bit depths and colour types, interlace
  for (int field = 0; field < 5; ++field) {
    std::vector<uint8_t> b = good;
    uint8_t* ihdr = b.data() + 16;  // width, height, depth, colour, compression, filter, interlace
    if (field == 0) { ihdr[0] = 0x7f; ihdr[1] = 0xff; }
    if (field == 1) { ihdr[4] = 0x40; ihdr[5] = 0; ihdr[6] = 0; ihdr[7] = 0; }
    if (field == 2) ihdr[8] = 8;
    if (field == 3) ihdr[9] = 2;
    if (field == 4) ihdr[12] = 1;
    write_file(bad_png, b.data(), b.size());
    CHECK(try_png(bad_png, 640, 480) != VH_OK);
  }
  // a chunk length beyond the file
  {
    std::vector<uint8_t> b = good;
    size_t i = 8 + 25;  // the chunk after IHDR
    CHECK(be32(b.data() + 8) == 13);
    b[i] = 0x7f;
    write_file(bad_png, b.data(), b.size());
    CHECK(try_png(bad_png, 640, 480) != VH_OK);
  }
  CHECK(try_png(dir + "/missing.png", 640, 480) != VH_OK);

  // settings files
  vh_settings st;
  if (!yaml_ok.empty()) {
    CHECK(vh_settings_load(yaml_ok.c_str(), &st) == VH_OK);
    vh_config cfg;
    std::memset(&cfg, 0, sizeof cfg);
    CHECK(vh_settings_to_config(&st, &cfg) == VH_OK && cfg.width == 640);
  }
  const std::string y = dir + "/s.yaml";
  const char* yamls[] = {
      "",
      "%YAML:1.0\n",
      "%YAML:1.0\nCamera.fx: abc\n",
      "%YAML:1.0\nCamera.fx:\n",
      "%YAML:1.0\nCamera.fx: 1e999999\n",
      "%YAML:1.0\n: 5\n:::\n#\n\n",
      "%YAML:1.0\nCamera.width: 99999999999999999999\n",
      "no header\nCamera.fx: 1\n",
      "%YAML:1.0\nCamera.fx: 577\nCamera.fx: 578\nMaxDepth: -1\n",
  };
  for (const char* t : yamls) {
    write_text(y, t);
    if (vh_settings_load(y.c_str(), &st) == VH_OK) {
      vh_config cfg;
      std::memset(&cfg, 0, sizeof cfg);
      (void)vh_settings_to_config(&st, &cfg);
    }
  }
  {
    std::string longline = "%YAML:1.0\nCamera.fx: ";
    longline.append(1 << 20, '9');
    longline += "\n";
    write_text(y, longline);
    (void)vh_settings_load(y.c_str(), &st);
    std::vector<uint8_t> junk(4096);
    for (size_t i = 0; i < junk.size(); ++i) junk[i] = (uint8_t)(i * 131 + 7);
    write_file(y, junk.data(), junk.size());
    (void)vh_settings_load(y.c_str(), &st);
  }
  CHECK(vh_settings_load((dir + "/missing.yaml").c_str(), &st) != VH_OK);

  // pose files
  float c2w[16];
  const std::string pf = dir + "/p.txt";
  write_text(pf, "1 0 0 0 0 1 0 0 0 0 1 0 0 0 0 1\n");
  CHECK(vh_pose_load(pf.c_str(), c2w) == VH_OK && c2w[15] == 1.0f);
  for (const char* t : {"", "1 2 3", "a b c d e f g h i j k l m n o p", "1 0 0 0 0 1 0 0 0 0 1 0 0 0 0 nan"}) {
    write_text(pf, t);
    (void)vh_pose_load(pf.c_str(), c2w);
  }
  std::vector<uint16_t> fd(640 * 480);
  CHECK(vh_frame_load((dir + "/nowhere").c_str(), 0, fd.data(), 640, 480, c2w) != VH_OK);

  // the oracle on three frames of the small room, FP32 and F16R, with colour and carving
  vh_scene* s = nullptr;
  CHECK(vh_scene_create(1, 2024, &s) == VH_OK);
  CHECK(vh_scene_render_range(s, 0, 3, depth.data(), poses.data()) == VH_OK);
  for (int storage : {0, 2}) {
    void* o = vho_create(640, 480, 525, 525, 319.5f, 239.5f, 5.0f, 0.1f, 0.01f, 0.04f, storage, storage ? 255 : 0);
    vho_set_color(o, 1);
    vho_set_carve(o, 1);
    for (int k = 0; k < 3; ++k) {
      CHECK(vh_scene_render_rgb(s, k, &poses[16 * k], rgb.data()) == VH_OK);
      CHECK(vho_integrate_rgb(o, depth.data() + (size_t)k * 640 * 480, rgb.data(), 0.001f, &poses[16 * k]) > 0);
    }
    const int64_t nb = vho_num_blocks(o);
    std::vector<int32_t> keys((size_t)nb * 3);
    std::vector<float> ts((size_t)nb * 512), wt((size_t)nb * 512);
    std::vector<double> s64((size_t)nb * 512), w64((size_t)nb * 512);
    CHECK(vho_export(o, keys.data(), ts.data(), wt.data(), s64.data(), w64.data(), nb) == nb);
    const int64_t nt = vho_mc(o, nullptr, nullptr, 0);
    std::vector<float> xyz((size_t)nt * 9);
    std::vector<int32_t> cubes((size_t)nt * 4);
    CHECK(vho_mc(o, xyz.data(), cubes.data(), nt) == nt && nt > 1000);
    const int64_t nt64 = vho_mc64(o, nullptr, nullptr, 0);
    std::vector<float> xyz64((size_t)nt64 * 9);
    std::vector<int32_t> cubes64((size_t)nt64 * 4);
    CHECK(vho_mc64(o, xyz64.data(), cubes64.data(), nt64) == nt64);
    double out[16];
    std::vector<float> dist((size_t)nt * 4);
    vho_mesh_distance(xyz.data(), cubes.data(), nt, xyz64.data(), cubes64.data(), nt64, out, dist.data());
    int64_t nf = 0;
    const int64_t nv = vho_mesh_indexed(o, nullptr, nullptr, 0, nullptr, nullptr, 0, &nf);
    std::vector<float> verts((size_t)nv * 3);
    std::vector<uint8_t> vrgb((size_t)nv * 3);
    std::vector<int32_t> faces((size_t)nf * 3), fcubes((size_t)nf * 4);
    CHECK(vho_mesh_indexed(o, verts.data(), vrgb.data(), nv, faces.data(), fcubes.data(), nf, &nf) == nv);
    std::vector<double> pts(3000);
    for (size_t i = 0; i < pts.size(); ++i) pts[i] = xyz[(i % (size_t)(nt * 9))];
    std::vector<double> sdf(1000), grad(3000), g((size_t)nb * 512);
    std::vector<float> sdf32(1000), grad32(3000), up(1000, 0.5f);
    std::vector<uint8_t> valid(1000);
    vho_query64(o, pts.data(), 1000, sdf.data(), grad.data(), valid.data());
    vho_query_twin(o, pts.data(), 1000, sdf32.data(), grad32.data(), valid.data());
    CHECK(vho_scatter_grad(o, pts.data(), up.data(), 1000, g.data(), 1) >= 0);
    int32_t band[3 * 64];
    CHECK(vho_band(o, &poses[0], 320, 240, 1.5f, band, 64) >= 0);
    void* o2 = vho_create(640, 480, 525, 525, 319.5f, 239.5f, 5.0f, 0.1f, 0.01f, 0.04f, 0, 0);
    CHECK(vho_import_f32(o2, keys.data(), ts.data(), wt.data(), nb) == 0);
    vho_destroy(o2);
    vho_destroy(o);
  }
  vh_scene_destroy(s);
  if (failures) {
    std::fprintf(stderr, "%d checks failed\n", failures);
    return 1;
  }
  std::printf("ok\n");
  return 0;
}
