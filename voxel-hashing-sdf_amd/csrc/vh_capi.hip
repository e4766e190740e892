// vh_capi.hip - extern "C" entry points of libvh (see include/vh.h).
//
// One volume owns one HIP stream (default priority) and all of its device memory. The
// per-frame pipeline (k_alloc -> k_integrate, preceded for tracked sequences by the
// back-projection and the Gauss-Newton iterations) is captured once into a HIP graph per
// input kind (host-staged frames, bound device sequence, tracked sequence); its kernels read
// the frame index and counters from device memory, so a launch needs no host-side parameter
// updates.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <tuple>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "vh.h"
#include "vh_error.h"
#include "vh_internal.h"
#include "vh_math.h"

using namespace vh;

namespace {
thread_local std::string g_last_error;
}  // namespace

int vh::set_last_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

namespace {

int fail(int code, const std::string& msg) { return set_last_error(code, msg); }

#define VH_HIP(call)                                                                           \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess)                                                                      \
      return fail(VH_ERR_DEVICE, std::string(#call " failed: ") + hipGetErrorString(e_));      \
  } while (0)

// Integration workers (waves, four per workgroup), a multiple of kStripes: 5,120 is one
// resident round at 5 waves per SIMD (256 CUs x 4 SIMDs). room_medium_5mm FP32, batches of 8,
// average launch: 8,192 workers 91.4 us, 5,120 84.8 us, 4,096 85.2 us.
constexpr int kIntegrateGrid = 5120;
// Integration workers of the fused batch launch (k_integrate_alloc), whose allocation
// workgroups are dispatched after the integration ones: fewer workers leave the allocation
// more of the GPU. room_medium_5mm, integration-only frames/s by worker count (3,072 /
// 4,096 / 8,192): FP32 63.8k / 62.6k / 59.6k, FP16 63.9k / 63.8k / 62.2k, F16R 51.1k /
// 52.2k / 51.8k, FP32 + colour 40.6k / 39.6k / 38.9k, F16R + colour 35.4k / 36.5k / 36.0k;
// FP32 with 2,048 workers 56.4k. Average fused launch after the depth-window and load-order
// work (2,048 / 3,072 / 4,096 / 5,120 workers): FP32 115.1 / 100.8 / 101.2 / 102.6 us, FP16
// 124.8 / 106.9 / 105.1 / 106.2, F16R 142.4 / 126.1 / 121.7 / 123.3, FP32 + colour 195.9 /
// 167.6 / 164.1 / 171.5. Plain FP32 takes 3,072; every costlier voxel update takes 4,096.
constexpr int kFusedGrid = 3072;
constexpr int kFusedGridHeavy = 4096;
constexpr int kGnGrid = 512;  // most workgroups of a solver or backward reduction

// Solver grid for n points: one point per lane, up to kGnGrid workgroups (a partial per
// workgroup is summed by the solve step). The accumulation is latency-bound, so the shorter
// chain per lane wins over the extra partials: tracked floor_large_4mm (19,200 points per
// evaluation) 4,532 frames/s with two points per lane, 4,783 with one, 3,766 with four.
int gn_grid(int64_t n) {
  const int64_t g = (n + 255) / 256;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, kGnGrid));
}

// Launch-shape tuning: VH_INTEGRATE_GRID and VH_FUSED_GRID override the integration grids of
// the integration launch and of the fused batch launch (rounded to a multiple of kStripes).
int grid_setting(const char* name, long dflt) {
  const char* e = std::getenv(name);
  long v = e ? std::strtol(e, nullptr, 10) : dflt;
  if (v < kStripes) v = kStripes;
  if (v > (1 << 20)) v = 1 << 20;
  return (int)(v / kStripes * kStripes);
}
int integrate_grid() {
  static const int g = grid_setting("VH_INTEGRATE_GRID", kIntegrateGrid);
  return g;
}
int fused_grid(int storage, bool color) {
  static const int g = grid_setting("VH_FUSED_GRID", kFusedGrid);
  static const int heavy = grid_setting("VH_FUSED_GRID", kFusedGridHeavy);
  return storage == VH_STORAGE_F32 && !color ? g : heavy;
}

}  // namespace

struct vh_volume {
  vh_config cfg{};
  int device = 0;
  hipStream_t stream = nullptr;
  uint64_t n_slots = 0;
  // device tables
  unsigned long long* keys = nullptr;
  unsigned long long* vals = nullptr;
  uint32_t* marks = nullptr;  // 2 x n_slots
  unsigned long long* block_keys = nullptr;
  unsigned long long* block_info = nullptr;  // per pool block: hash slot | birth frame << 32
  float* tilemax = nullptr;    // carving volumes: 2 x kMaxBatch frames of max-depth tiles
  float* dm = nullptr;         // 2 x kMaxBatch frames of pre-scaled depth (depth / mu)
  void* voxels = nullptr;
  uint32_t* colors = nullptr;  // colour volumes: one u32 per voxel
  uint32_t* rgba = nullptr;    // colour volumes: 2 packed frames of W*H words
  uint32_t* touched = nullptr;
  uint32_t stripe_cap = 0;
  FrameCtl* ctl = nullptr;
  // staged host input (one frame, both formats) and its pose
  void* stage_depth = nullptr;
  float* stage_pose = nullptr;
  uint8_t* stage_rgb = nullptr;  // colour volumes: W*H*3
  const uint8_t* seq_rgb = nullptr;  // RGB of the bound sequence (vh_bind_sequence_rgb)
  // bound device sequence
  const void* seq_depth = nullptr;
  const float* seq_pose = nullptr;
  int32_t seq_n = 0, seq_format = VH_DEPTH_U16;
  float seq_scale = 0.001f;
  // tracking of the bound sequence (vh_bind_tracking)
  bool track = false;
  const float* trk_prior = nullptr;
  float* trk_out = nullptr;
  double* trk_stats = nullptr;
  int32_t trk_iters = 0, trk_stride = 1, trk_increment = 0;
  float trk_max_depth = 0.0f;  // depth limit of the tracking points
  float trk_sigma_t = 0.0f, trk_sigma_r = 0.0f, trk_points = 0.0f;  // motion prior, depth-term scale
  float* trk_pts = nullptr;
  size_t trk_pts_cap = 0;  // points the buffer holds
  // host frames queued by vh_integrate_async: copied on copy_stream into kRing device slots;
  // a frame's kernels wait for its copy, a slot's next copy waits for the frame that read it
  static constexpr int kRing = 4;
  hipStream_t copy_stream = nullptr;
  void* ring_depth = nullptr;       // kRing frames of W*H elements
  int32_t ring_format = -1;         // depth format of the queued frames (slot stride W*H*element)
  uint8_t* ring_rgb = nullptr;      // colour volumes: kRing frames of W*H*3 bytes
  float* ring_pose = nullptr;       // kRing x 16 (device)
  float* ring_pose_host = nullptr;  // kRing x 16 (page-locked)
  hipEvent_t ring_copied[kRing] = {}, ring_done[kRing] = {};
  uint64_t ring_next = 0;           // frames queued since create/reset (= device cursor 2)
  // graphs by (kind, frames integrated, frames allocated): kinds [0] staged u16, [1] staged
  // f32, [2] bound sequence step (integrate batch b + allocate batch b+1), [3] tracked
  // sequence, [4] bound sequence first allocation, [5] bound sequence last integration,
  // [6] ring u16, [7] ring f32
  static constexpr int kGraphs = 8;
  std::map<std::tuple<int, int, int>, hipGraphExec_t> graphs;
  int batch = kMaxBatch;  // frames per launch on bound sequences
  float ring_scale = -1.0f;  // depth scale baked into the ring uint16 graph
  // profiling: 1 = eager launches with events around each kernel, 2 = the captured graphs with
  // events around each fused batch launch
  int profiling = 0;
  std::vector<hipEvent_t> ev_pool;
  std::vector<hipEvent_t> ev_pending;  // groups of 5: start, after tracking, after alloc, after carve, after integrate
  std::vector<int> ev_frames;          // frames of each group
  std::vector<hipEvent_t> ev_fused;    // pairs around fused batch launches (profiling 2)
  double alloc_ms = 0.0, integrate_ms = 0.0, track_ms = 0.0, carve_ms = 0.0, fused_ms = 0.0;
  int64_t profiled_frames = 0, profiled_launches = 0, fused_launches = 0;
  // marching-cubes output buffer (grown on demand) and scan scratch
  float* mc_tris = nullptr;
  int32_t* mc_cubes = nullptr;
  uint8_t* mc_rgb = nullptr;       // colour volumes
  // indexed mesh: per pool block used-edge words, their prefixes and vertex counts; the
  // host variant's device output (ix_fcap faces, ix_vcap vertices)
  uint32_t* ix_bits = nullptr;     // pool_blocks x kEdgeWords
  uint16_t* ix_pre = nullptr;      // pool_blocks x kEdgeWords
  uint32_t* ix_vcounts = nullptr;  // pool_blocks
  uint32_t* ix_vchunks = nullptr;  // pool_blocks / kScanChunk + 1
  uint32_t* ix_vtotal = nullptr;
  int32_t* ix_faces = nullptr;     // ix_fcap x 3
  float* ix_verts = nullptr;       // ix_vcap x 3
  uint8_t* ix_vrgb = nullptr;      // ix_vcap x 3 (colour volumes)
  int64_t ix_fcap = 0, ix_vcap = 0;
  // out-of-core streaming: the host store holds blocks moved off the GPU (voxel words and
  // colours byte for byte) in pinned slabs of kSlabBlocks blocks that the kernels address
  // directly; slot s is block s % kSlabBlocks of slab s / kSlabBlocks
  std::vector<uint32_t*> slab_w, slab_c;        // host addresses (colour slabs: colour volumes)
  uint32_t** d_slab_w = nullptr;                // device addresses of the slabs
  uint32_t** d_slab_c = nullptr;
  size_t d_slab_cap = 0;                        // entries of the device address tables
  bool slab_table_dirty = false;
  std::vector<unsigned long long> slot_key;     // key per slot, kEmptyKey when free
  std::vector<uint32_t> free_slots;
  std::unordered_map<unsigned long long, uint32_t> store_index;  // key -> slot
  uint32_t* st_idx = nullptr;                   // 2 x st_cap index words
  unsigned long long* st_keys = nullptr;        // st_cap keys
  uint32_t st_cap = 0;
  int64_t streamed_out = 0, streamed_in = 0, stream_merged = 0;
  int64_t mc_cap = 0;
  uint32_t* mc_counts = nullptr;   // pool_blocks
  uint32_t* mc_nbrs = nullptr;     // pool_blocks x 8
  uint32_t* pin = nullptr;         // page-locked words for small device-to-host reads (kPinWords)
  uint32_t* mc_chunks = nullptr;   // pool_blocks / kScanChunk + 1
  uint32_t* mc_total = nullptr;
  // scratch for queries / Gauss-Newton
  double* gn_partials = nullptr;
  double* gn_pose = nullptr;
  double* gn_stats = nullptr;
  GnState* gn_state = nullptr;
  float* bw_pose = nullptr;
  float graph_scale = -1.0f;  // depth scale baked into the staged uint16 graph
  std::mutex mu;
};

namespace {

int set_device(vh_volume* v) {
  VH_HIP(hipSetDevice(v->device));
  return VH_OK;
}

size_t voxel_bytes(int storage) { return storage == VH_STORAGE_F32 ? 8 : 4; }

FrameArgs frame_args(vh_volume* v, int which) {
  FrameArgs a{};
  const vh_config& c = v->cfg;
  if (which == 6 || which == 7) {
    a.depth = v->ring_depth;
    a.poses = v->ring_pose;
    a.n_seq = vh_volume::kRing;
    a.cursor_id = 2;
    a.depth_scale = which == 6 ? v->ring_scale : 1.0f;
    a.rgb = v->ring_rgb;
  } else if (which >= 2) {
    a.depth = v->seq_depth;
    a.poses = which == 3 ? v->trk_out : v->seq_pose;
    a.n_seq = (uint32_t)v->seq_n;
    a.cursor_id = 1;
    a.depth_scale = v->seq_format == VH_DEPTH_U16 ? v->seq_scale : 1.0f;
    a.rgb = v->seq_rgb;
  } else {
    a.depth = v->stage_depth;
    a.poses = v->stage_pose;
    a.n_seq = 1;
    a.cursor_id = 0;
    a.depth_scale = 1.0f;  // the caller sets the uint16 scale of a staged frame
    a.rgb = v->stage_rgb;
  }
  a.W = c.width;
  a.H = c.height;
  a.K = Intr{c.fx, c.fy, c.cx, c.cy};
  a.max_depth = c.max_depth;
  a.min_depth = c.min_depth;
  a.vox = c.vox_size;
  a.bsz = c.vox_size * (float)kB;
  a.band = band_scale(a.K, a.bsz);
  a.mu = c.trunc;
  a.inv_mu = 1.0f / c.trunc;
  a.cap = c.weight_cap > 0 ? (float)c.weight_cap : INFINITY;
  a.hash = HashView{v->keys, v->vals, c.log2_buckets, ((uint64_t)1 << c.log2_buckets) - 1};
  a.marks = v->marks;
  a.n_slots = v->n_slots;
  a.block_keys = v->block_keys;
  a.block_info = v->block_info;
  a.pool_cap = (uint32_t)c.pool_blocks;
  a.touched = v->touched;
  a.stripe_cap = v->stripe_cap;
  a.voxels = v->voxels;
  a.colors = v->colors;
  a.rgba = v->rgba;
  if (!a.colors) a.rgb = nullptr;
  a.ctl = v->ctl;
  a.tilemax = v->tilemax;  // null unless the volume carves
  a.dm = v->dm;
  a.tiles_x = (c.width + kTile - 1) / kTile;
  a.tiles_y = (c.height + kTile - 1) / kTile;
  return a;
}

QueryArgs query_args(vh_volume* v) {
  QueryArgs q{};
  q.hash = HashView{v->keys, v->vals, v->cfg.log2_buckets, ((uint64_t)1 << v->cfg.log2_buckets) - 1};
  q.voxels = v->voxels;
  q.pool_cap = (uint32_t)v->cfg.pool_blocks;
  q.vox = v->cfg.vox_size;
  return q;
}

TrackArgs track_args(vh_volume* v) {
  const vh_config& c = v->cfg;
  TrackArgs t{};
  t.depth = v->seq_depth;
  t.depth_u16 = v->seq_format == VH_DEPTH_U16;
  t.depth_scale = t.depth_u16 ? v->seq_scale : 1.0f;
  t.prior = v->trk_prior;
  t.increment = v->trk_increment;
  t.out = v->trk_out;
  t.stats = v->trk_stats;
  t.eval_stats = v->gn_stats;
  t.n_seq = (uint32_t)v->seq_n;
  t.ctl = v->ctl;
  t.W = c.width;
  t.H = c.height;
  t.stride = v->trk_stride;
  t.pw = (c.width + v->trk_stride - 1) / v->trk_stride;
  t.ph = (c.height + v->trk_stride - 1) / v->trk_stride;
  t.K = Intr{c.fx, c.fy, c.cx, c.cy};
  t.max_depth = v->trk_max_depth;
  t.prior_sigma_t = v->trk_sigma_t;
  t.prior_sigma_r = v->trk_sigma_r;
  t.n_eff = v->trk_points;
  t.pts = v->trk_pts;
  t.pose = v->gn_pose;
  t.state = v->gn_state;
  t.iters = v->trk_iters;
  return t;
}

// Solver launches after the start pose and state are in place (cand, state[0]): iters + 1
// evaluations, each one launch that first takes the step from the previous evaluation (the
// partial sums and states alternate between two buffers by evaluation parity), then the final
// step, which leaves the best accepted pose in cand. partials: 2 x kGnGrid x kGnAcc doubles,
// state: 2, stats: (iters + 1) rows.
hipError_t enqueue_gn_evaluations(vh_volume* v, GnArgs g, int iters, double* partials, GnState* state, double* cand,
                                  double* stats) {
  const int grid = gn_grid(g.n);
  const size_t pstride = (size_t)kGnGrid * kGnAcc;
  g.pose = cand;
  for (int it = 0; it <= iters; ++it) {
    g.partials = partials + (it & 1) * pstride;
    g.prev_grid = it > 0 ? grid : 0;
    g.prev_partials = partials + ((it + 1) & 1) * pstride;
    g.st_in = state + ((it + 1) & 1);
    g.st_out = state + (it & 1);
    g.cand_out = nullptr;
    g.stats_row = it > 0 ? stats + kGnStats * (it - 1) : nullptr;
    hipError_t e = launch_gn_accum(g, v->cfg.storage, grid, v->stream);
    if (e != hipSuccess) return e;
  }
  return launch_gn_solve(partials + (iters & 1) * pstride, grid, state + (iters & 1), state + ((iters + 1) & 1), cand,
                         1, stats + kGnStats * iters, v->stream);
}

// Tracking part of a frame of the tracked sequence: back-projection, iters + 1 solver
// evaluations and the final step, the pose write-out.
int enqueue_track_kernels(vh_volume* v) {
  const TrackArgs t = track_args(v);
  VH_HIP(launch_track_begin(t, v->stream));
  GnArgs g{};
  g.q = query_args(v);
  g.mu = v->cfg.trunc;
  g.pts = t.pts;
  g.n = (int64_t)t.pw * t.ph;
  VH_HIP(enqueue_gn_evaluations(v, g, t.iters, v->gn_partials, v->gn_state, v->gn_pose, v->gn_stats));
  VH_HIP(launch_track_end(t, v->stream));
  return VH_OK;
}

int frame_u16(vh_volume* v, int which) {
  if (which == 6 || which == 7) return which == 6;
  return which >= 2 ? v->seq_format == VH_DEPTH_U16 : which == 0;
}

// Kinds 0, 1, 3, 6, 7: allocation then integration of one frame. Bound sequences run in
// batches of up to `batch` frames: kind 4 (allocate the first batch, ka frames), kind 2 per
// further batch (integrate the previous batch, ki frames, and allocate this one, ka frames,
// in one launch) and kind 5 (integrate the last batch, ki frames).
int enqueue_frame_kernels(vh_volume* v, int which, float staged_scale, int ki, int ka) {
  const int fa = which == 4 || which == 5 ? 2 : which;
  FrameArgs a = frame_args(v, fa);
  const int u16 = frame_u16(v, fa);
  if (fa < 2 || fa >= 6) a.depth_scale = u16 ? staged_scale : 1.0f;
  const bool carve = v->cfg.carve != 0;
  if (which == 2) {
    VH_HIP(launch_integrate_alloc(a, v->cfg.storage, u16, fused_grid(v->cfg.storage, v->cfg.color != 0), ki, ka, v->stream));
    if (carve) VH_HIP(launch_carve(a, ka, v->stream));  // the batch just allocated
    return VH_OK;
  }
  if (which == 3) {
    int rc = enqueue_track_kernels(v);
    if (rc) return rc;
  }
  if (which != 5) {
    const int k = which == 4 ? ka : 1;
    VH_HIP(launch_alloc(a, u16, k, v->stream));
    if (carve) VH_HIP(launch_carve(a, k, v->stream));
  }
  if (which != 4) VH_HIP(launch_integrate(a, v->cfg.storage, integrate_grid(), which == 5 ? ki : 1, v->stream));
  return VH_OK;
}

hipEvent_t take_event(vh_volume* v) {
  if (!v->ev_pool.empty()) {
    hipEvent_t e = v->ev_pool.back();
    v->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// One launch of kind `which` (ki frames integrated, ka allocated; single frames: 1, 1). In
// profiling mode a batch of k = max(ki, ka) frames runs eagerly as (tracking,) allocation of k
// frames and integration of k frames, with events around each.
int launch_frame(vh_volume* v, int which, float staged_scale, int ki = 1, int ka = 1) {
  if (v->profiling == 1) {
    hipEvent_t e[5] = {take_event(v), take_event(v), take_event(v), take_event(v), take_event(v)};
    if (!e[0] || !e[1] || !e[2] || !e[3] || !e[4]) return fail(VH_ERR_DEVICE, "hipEventCreate failed");
    FrameArgs a = frame_args(v, which);
    const int u16 = frame_u16(v, which);
    if (which < 2 || which >= 6) a.depth_scale = u16 ? staged_scale : 1.0f;
    const int k = std::max(ki, ka);
    VH_HIP(hipEventRecord(e[0], v->stream));
    if (which == 3) {
      int rc = enqueue_track_kernels(v);
      if (rc) return rc;
    }
    VH_HIP(hipEventRecord(e[1], v->stream));
    VH_HIP(launch_alloc(a, u16, k, v->stream));
    VH_HIP(hipEventRecord(e[2], v->stream));
    if (v->cfg.carve) VH_HIP(launch_carve(a, k, v->stream));
    VH_HIP(hipEventRecord(e[3], v->stream));
    VH_HIP(launch_integrate(a, v->cfg.storage, integrate_grid(), k, v->stream));
    VH_HIP(hipEventRecord(e[4], v->stream));
    for (hipEvent_t x : e) v->ev_pending.push_back(x);
    v->ev_frames.push_back(k);
    return VH_OK;
  }
  hipGraphExec_t& g = v->graphs[std::make_tuple(which, ki, ka)];
  if (!g) {
    hipGraph_t graph = nullptr;
    VH_HIP(hipStreamBeginCapture(v->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_frame_kernels(v, which, staged_scale, ki, ka);
    hipError_t e = hipStreamEndCapture(v->stream, &graph);
    if (rc != VH_OK) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    VH_HIP(e);
    e = hipGraphInstantiate(&g, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    VH_HIP(e);
  }
  if (v->profiling == 2 && which == 2) {
    hipEvent_t e0 = take_event(v), e1 = take_event(v);
    if (!e0 || !e1) return fail(VH_ERR_DEVICE, "hipEventCreate failed");
    VH_HIP(hipEventRecord(e0, v->stream));
    VH_HIP(hipGraphLaunch(g, v->stream));
    VH_HIP(hipEventRecord(e1, v->stream));
    v->ev_fused.push_back(e0);
    v->ev_fused.push_back(e1);
    return VH_OK;
  }
  VH_HIP(hipGraphLaunch(g, v->stream));
  return VH_OK;
}

void drop_graphs(vh_volume* v, int which_mask) {
  for (auto it = v->graphs.begin(); it != v->graphs.end();) {
    if ((which_mask >> std::get<0>(it->first)) & 1) {
      if (it->second) (void)hipGraphExecDestroy(it->second);
      it = v->graphs.erase(it);
    } else {
      ++it;
    }
  }
}

int collect_profile(vh_volume* v) {
  for (size_t k = 0; k + 4 < v->ev_pending.size(); k += 5) {
    float t = 0.0f, a = 0.0f, c = 0.0f, b = 0.0f;
    VH_HIP(hipEventElapsedTime(&t, v->ev_pending[k], v->ev_pending[k + 1]));
    VH_HIP(hipEventElapsedTime(&a, v->ev_pending[k + 1], v->ev_pending[k + 2]));
    VH_HIP(hipEventElapsedTime(&c, v->ev_pending[k + 2], v->ev_pending[k + 3]));
    VH_HIP(hipEventElapsedTime(&b, v->ev_pending[k + 3], v->ev_pending[k + 4]));
    v->track_ms += t;
    v->alloc_ms += a;
    v->carve_ms += c;
    v->integrate_ms += b;
    v->profiled_frames += v->ev_frames[k / 5];
    v->profiled_launches += 1;
  }
  for (hipEvent_t e : v->ev_pending) v->ev_pool.push_back(e);
  v->ev_pending.clear();
  v->ev_frames.clear();
  for (size_t k = 0; k + 1 < v->ev_fused.size(); k += 2) {
    float t = 0.0f;
    VH_HIP(hipEventElapsedTime(&t, v->ev_fused[k], v->ev_fused[k + 1]));
    v->fused_ms += t;
    v->fused_launches += 1;
  }
  for (hipEvent_t e : v->ev_fused) v->ev_pool.push_back(e);
  v->ev_fused.clear();
  return VH_OK;
}

// Device words read through the volume's page-locked words (a pageable destination makes
// the copy go through a staging buffer): n <= kPinWords words from dev into out.
constexpr int kPinWords = 16;
int read_words(vh_volume* v, const uint32_t* dev, uint32_t* out, int n) {
  VH_HIP(hipMemcpyAsync(v->pin, dev, (size_t)n * 4, hipMemcpyDeviceToHost, v->stream));
  VH_HIP(hipStreamSynchronize(v->stream));
  for (int i = 0; i < n; ++i) out[i] = v->pin[i];
  return VH_OK;
}

int read_ctl(vh_volume* v, FrameCtl* out) {
  VH_HIP(hipMemcpyAsync(out, v->ctl, sizeof(FrameCtl), hipMemcpyDeviceToHost, v->stream));
  VH_HIP(hipStreamSynchronize(v->stream));
  return VH_OK;
}

// Waits for the stream, folds profiling events, reports and clears overflow flags.
int sync_and_check(vh_volume* v) {
  uint32_t flags = 0;
  int rc = read_words(v, &v->ctl->flags, &flags, 1);  // after the stream's work
  if (rc || (rc = collect_profile(v))) return rc;
  if (!flags) return VH_OK;
  const uint32_t zero = 0;
  VH_HIP(hipMemcpy(&v->ctl->flags, &zero, sizeof zero, hipMemcpyHostToDevice));
  if (flags & kFlagPoolFull)
    return fail(VH_ERR_POOL_FULL, "voxel-block pool is full: blocks were not integrated");
  if (flags & kFlagTouchedFull)
    return fail(VH_ERR_POOL_FULL, "a frame touched more blocks than the per-frame list holds (sized from pool_blocks)");
  if (flags & kFlagHashFull)
    return fail(VH_ERR_HASH_FULL, "hash table probe limit reached: blocks were not allocated");
  return fail(VH_ERR_RANGE, "a truncation band left the representable block range");
}

int reset_tables(vh_volume* v) {
  v->ring_next = 0;  // the device cursor of the ring restarts with the control block
  v->store_index.clear();  // the slabs stay allocated for reuse
  v->slot_key.clear();
  v->free_slots.clear();
  v->streamed_out = v->streamed_in = v->stream_merged = 0;
  VH_HIP(hipMemsetAsync(v->keys, 0xff, v->n_slots * sizeof(unsigned long long), v->stream));
  VH_HIP(hipMemsetAsync(v->marks, 0, 2 * v->n_slots * 4, v->stream));
  FrameCtl c{};
  VH_HIP(hipMemcpyAsync(&c, v->ctl, sizeof c, hipMemcpyDeviceToHost, v->stream));
  VH_HIP(hipStreamSynchronize(v->stream));
  const uint32_t frame_id = c.frame_id ? c.frame_id : 1u;
  std::memset(&c, 0, sizeof c);
  c.frame_id = frame_id;  // stamps and birth frames stay comparable across resets
  VH_HIP(hipMemcpyAsync(v->ctl, &c, sizeof c, hipMemcpyHostToDevice, v->stream));
  VH_HIP(hipStreamSynchronize(v->stream));
  return VH_OK;
}

int validate(const vh_config* c) {
  if (c->width <= 0 || c->height <= 0 || c->width > 16384 || c->height > 16384)
    return fail(VH_ERR_INVALID_ARG, "width/height out of range");
  if (!(c->fx > 0.0f) || !(c->fy > 0.0f)) return fail(VH_ERR_INVALID_ARG, "fx, fy must be > 0");
  if (!(c->vox_size > 0.0f) || !(c->trunc > 0.0f)) return fail(VH_ERR_INVALID_ARG, "vox_size, trunc must be > 0");
  if (!(c->max_depth > 0.0f) || !(c->min_depth >= 0.0f)) return fail(VH_ERR_INVALID_ARG, "bad depth range");
  if (c->log2_buckets < 1 || c->log2_buckets > 28) return fail(VH_ERR_INVALID_ARG, "log2_buckets must be in 1..28");
  if (c->pool_blocks <= 0 || c->pool_blocks > (1 << 27)) return fail(VH_ERR_INVALID_ARG, "pool_blocks must be in 1..2^27");
  if (c->storage != VH_STORAGE_F32 && c->storage != VH_STORAGE_F16 && c->storage != VH_STORAGE_F16R)
    return fail(VH_ERR_INVALID_ARG, "unknown storage");
  if (c->weight_cap < 0) return fail(VH_ERR_INVALID_ARG, "weight_cap must be >= 0");
  if (c->storage == VH_STORAGE_F16 && (c->weight_cap < 1 || c->weight_cap > 65535))
    return fail(VH_ERR_INVALID_ARG, "FP16 storage needs weight_cap in 1..65535");
  if (c->storage == VH_STORAGE_F16R && (c->weight_cap < 1 || c->weight_cap > 255))
    return fail(VH_ERR_INVALID_ARG, "F16R storage needs weight_cap in 1..255");
  if (c->color != 0 && c->color != 1) return fail(VH_ERR_INVALID_ARG, "color must be 0 or 1");
  if (c->seq_batch < 0 || c->seq_batch > kMaxBatch)
    return fail(VH_ERR_INVALID_ARG, "seq_batch must be in 0..8");
  if (c->carve != 0 && c->carve != 1) return fail(VH_ERR_INVALID_ARG, "carve must be 0 or 1");
  for (int32_t r : c->reserved)
    if (r) return fail(VH_ERR_INVALID_ARG, "reserved fields must be zero");
  return VH_OK;
}

template <class T>
int dev_alloc(T** p, size_t bytes) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, bytes ? bytes : 1);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(VH_ERR_OOM, std::string("hipMalloc of ") + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
  }
  *p = static_cast<T*>(q);
  return VH_OK;
}

void free_index_buffers(vh_volume* v);

void free_volume(vh_volume* v) {
  if (!v) return;
  (void)hipSetDevice(v->device);
  if (v->stream) (void)hipStreamSynchronize(v->stream);
  drop_graphs(v, (1 << vh_volume::kGraphs) - 1);
  if (v->copy_stream) (void)hipStreamSynchronize(v->copy_stream);
  for (hipEvent_t e : v->ev_pool) (void)hipEventDestroy(e);
  for (int k = 0; k < vh_volume::kRing; ++k) {
    if (v->ring_copied[k]) (void)hipEventDestroy(v->ring_copied[k]);
    if (v->ring_done[k]) (void)hipEventDestroy(v->ring_done[k]);
  }
  if (v->ring_pose_host) (void)hipHostFree(v->ring_pose_host);
  if (v->pin) (void)hipHostFree(v->pin);
  if (v->copy_stream) (void)hipStreamDestroy(v->copy_stream);
  for (hipEvent_t e : v->ev_pending) (void)hipEventDestroy(e);
  free_index_buffers(v);
  void* ptrs[] = {v->keys, v->vals, v->marks, v->block_keys, v->block_info, v->tilemax, v->dm, v->voxels, v->touched, v->ctl,
                  v->stage_depth, v->stage_pose, v->mc_tris, v->mc_cubes, v->mc_counts, v->mc_chunks, v->mc_total, v->mc_nbrs,
                  v->gn_partials, v->gn_pose, v->gn_stats, v->bw_pose, v->gn_state, v->trk_pts,
                  v->ring_depth, v->ring_pose, v->colors, v->stage_rgb, v->ring_rgb, v->mc_rgb, v->rgba,
                  v->st_idx, v->st_keys, v->d_slab_w, v->d_slab_c};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (uint32_t* p : v->slab_w) (void)hipHostFree(p);
  for (uint32_t* p : v->slab_c) (void)hipHostFree(p);
  if (v->stream) (void)hipStreamDestroy(v->stream);
  delete v;
}

int ensure_mc_scratch(vh_volume* v) {
  if (v->mc_counts) return VH_OK;
  int rc;
  if ((rc = dev_alloc(&v->mc_counts, (size_t)v->cfg.pool_blocks * 4)) ||
      (rc = dev_alloc(&v->mc_chunks, ((size_t)v->cfg.pool_blocks / kScanChunk + 1) * 4)) ||
      (rc = dev_alloc(&v->mc_nbrs, (size_t)v->cfg.pool_blocks * 8 * 4)) || (rc = dev_alloc(&v->mc_total, 4)))
    return rc;
  return VH_OK;
}

// Marching-cubes arguments over the volume's current blocks (reads the frame control block).
int mc_args(vh_volume* v, McArgs& a) {
  uint32_t pool_count = 0;
  int rc = ensure_mc_scratch(v);
  if (rc || (rc = read_words(v, &v->ctl->pool_count, &pool_count, 1))) return rc;
  a = McArgs{};
  a.hash = HashView{v->keys, v->vals, v->cfg.log2_buckets, ((uint64_t)1 << v->cfg.log2_buckets) - 1};
  a.block_keys = v->block_keys;
  a.voxels = v->voxels;
  a.pool_cap = (uint32_t)v->cfg.pool_blocks;
  a.n_blocks = std::min<uint32_t>(pool_count, (uint32_t)v->cfg.pool_blocks);
  a.vox = v->cfg.vox_size;
  a.counts = v->mc_counts;
  a.nbrs = v->mc_nbrs;
  a.chunk_sums = v->mc_chunks;
  a.total = v->mc_total;
  a.colors = v->colors;
  return VH_OK;
}

int run_mc(vh_volume* v, float* d_xyz, int32_t* d_cubes, uint8_t* d_rgb, int64_t capacity, int64_t* n_tris) {
  McArgs a;
  int rc = mc_args(v, a);
  if (rc) return rc;
  a.tris = d_xyz;
  a.cubes = d_cubes;
  a.rgb = d_rgb;
  a.capacity = (uint32_t)std::min<int64_t>(capacity, 0xffffffffll);
  uint32_t cnt = 0;
  if (a.n_blocks) {
    VH_HIP(launch_mc(a, v->cfg.storage, v->stream));
    if ((rc = read_words(v, v->mc_total, &cnt, 1))) return rc;
  } else {
    VH_HIP(hipStreamSynchronize(v->stream));
  }
  *n_tris = cnt;
  if ((int64_t)cnt > capacity) return fail(VH_ERR_TRUNCATED, "mesh larger than the output buffer");
  return VH_OK;
}
int ensure_mc_buffer(vh_volume* v, int64_t cap) {
  if (cap <= v->mc_cap) return VH_OK;
  if (v->mc_tris) (void)hipFree(v->mc_tris);
  if (v->mc_cubes) (void)hipFree(v->mc_cubes);
  if (v->mc_rgb) (void)hipFree(v->mc_rgb);
  v->mc_tris = nullptr;
  v->mc_cubes = nullptr;
  v->mc_rgb = nullptr;
  v->mc_cap = 0;
  int rc = dev_alloc(&v->mc_tris, (size_t)cap * 9 * sizeof(float));
  if (rc) return rc;
  rc = dev_alloc(&v->mc_cubes, (size_t)cap * 4 * sizeof(int32_t));
  if (rc) return rc;
  if (v->colors && (rc = dev_alloc(&v->mc_rgb, (size_t)cap * 9))) return rc;
  v->mc_cap = cap;
  return VH_OK;
}

void free_index_buffers(vh_volume* v) {
  void* ptrs[] = {v->ix_bits, v->ix_pre, v->ix_vcounts, v->ix_vchunks, v->ix_vtotal, v->ix_faces, v->ix_verts, v->ix_vrgb};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  v->ix_bits = nullptr;
  v->ix_pre = nullptr;
  v->ix_vcounts = nullptr;
  v->ix_vchunks = nullptr;
  v->ix_vtotal = nullptr;
  v->ix_faces = nullptr;
  v->ix_verts = nullptr;
  v->ix_vrgb = nullptr;
  v->ix_fcap = v->ix_vcap = 0;
}

// Indexed mesh, first half: used edges, face and vertex counts and offsets. *a receives the
// arguments for index_write; *n_faces and *n_vertices the totals.
int index_count(vh_volume* v, McArgs& a, int64_t* n_faces, int64_t* n_vertices) {
  int rc = mc_args(v, a);
  if (rc) return rc;
  if (!v->ix_bits) {
    const size_t pb = (size_t)v->cfg.pool_blocks;
    if ((rc = dev_alloc(&v->ix_bits, pb * kEdgeWords * 4)) || (rc = dev_alloc(&v->ix_pre, pb * kEdgeWords * 2)) ||
        (rc = dev_alloc(&v->ix_vcounts, pb * 4)) || (rc = dev_alloc(&v->ix_vchunks, (pb / kScanChunk + 1) * 4)) ||
        (rc = dev_alloc(&v->ix_vtotal, 4)))
      return rc;
  }
  a.ebits = v->ix_bits;
  a.epre = v->ix_pre;
  a.vcounts = v->ix_vcounts;
  a.vchunks = v->ix_vchunks;
  a.vtotal = v->ix_vtotal;
  uint32_t tot[2] = {0u, 0u};
  if (a.n_blocks) {
    VH_HIP(launch_mesh_index_count(a, v->cfg.storage, v->stream));
    VH_HIP(hipMemcpyAsync(&v->pin[0], v->mc_total, 4, hipMemcpyDeviceToHost, v->stream));
    VH_HIP(hipMemcpyAsync(&v->pin[1], v->ix_vtotal, 4, hipMemcpyDeviceToHost, v->stream));
  }
  VH_HIP(hipStreamSynchronize(v->stream));
  if (a.n_blocks) {
    tot[0] = v->pin[0];
    tot[1] = v->pin[1];
  }
  *n_faces = tot[0];
  *n_vertices = tot[1];
  if (tot[1] > 0x7fffffffu) return fail(VH_ERR_INVALID_ARG, "more than 2^31 vertices");
  return VH_OK;
}

// Second half: faces and vertices into buffers that hold them all.
int index_write(vh_volume* v, McArgs& a, int32_t* d_faces, int64_t n_faces, float* d_vertices, uint8_t* d_vrgb,
                int64_t n_vertices) {
  if (n_faces == 0) return VH_OK;
  a.faces = d_faces;
  a.capacity = (uint32_t)n_faces;
  a.vertices = d_vertices;
  a.vrgb = d_vrgb;
  a.vertex_cap = (uint32_t)n_vertices;
  VH_HIP(launch_mesh_index_write(a, v->cfg.storage, v->stream));
  VH_HIP(hipStreamSynchronize(v->stream));
  return VH_OK;
}

// ---------------------------------------------------------------- streaming internals
uint32_t words_per_block(const vh_volume* v) { return (uint32_t)(voxel_bytes(v->cfg.storage) * kB3 / 4); }

int ensure_stream_index(vh_volume* v, size_t n) {
  if (n <= v->st_cap) return VH_OK;
  if (n > 0xffffffffu / 2) return fail(VH_ERR_INVALID_ARG, "too many blocks to stream at once");
  if (v->st_idx) (void)hipFree(v->st_idx);
  if (v->st_keys) (void)hipFree(v->st_keys);
  v->st_idx = nullptr;
  v->st_keys = nullptr;
  v->st_cap = 0;
  int rc;
  if ((rc = dev_alloc(&v->st_idx, n * 2 * 4)) || (rc = dev_alloc(&v->st_keys, n * 8))) return rc;
  v->st_cap = (uint32_t)n;
  return VH_OK;
}

// A free host-store slot (new pinned slabs as needed).
int store_alloc_slot(vh_volume* v, uint32_t* slot) {
  if (!v->free_slots.empty()) {
    *slot = v->free_slots.back();
    v->free_slots.pop_back();
    return VH_OK;
  }
  const size_t s = v->slot_key.size();
  if (s >= ((size_t)1 << 31)) return fail(VH_ERR_OOM, "host store full");
  if ((s >> kSlabShift) >= v->slab_w.size()) {
    uint32_t *w = nullptr, *c = nullptr;
    VH_HIP(hipHostMalloc((void**)&w, (size_t)kSlabBlocks * words_per_block(v) * 4, hipHostMallocMapped));
    v->slab_w.push_back(w);
    if (v->colors) {
      VH_HIP(hipHostMalloc((void**)&c, (size_t)kSlabBlocks * kB3 * 4, hipHostMallocMapped));
      v->slab_c.push_back(c);
    }
    v->slab_table_dirty = true;
  }
  v->slot_key.push_back(kEmptyKey);
  *slot = (uint32_t)s;
  return VH_OK;
}

void store_free_slot(vh_volume* v, uint32_t slot) {
  v->store_index.erase(v->slot_key[slot]);
  v->slot_key[slot] = kEmptyKey;
  v->free_slots.push_back(slot);
}

uint32_t* store_words(vh_volume* v, uint32_t slot) {
  return v->slab_w[slot >> kSlabShift] + (size_t)(slot & (kSlabBlocks - 1u)) * words_per_block(v);
}
uint32_t* store_cols(vh_volume* v, uint32_t slot) {
  return v->slab_c[slot >> kSlabShift] + (size_t)(slot & (kSlabBlocks - 1u)) * kB3;
}

// The device address tables of the slabs, after slabs were added.
int upload_slab_table(vh_volume* v) {
  if (!v->slab_table_dirty) return VH_OK;
  const size_t n = v->slab_w.size();
  if (n > v->d_slab_cap) {
    if (v->d_slab_w) (void)hipFree(v->d_slab_w);
    if (v->d_slab_c) (void)hipFree(v->d_slab_c);
    v->d_slab_w = v->d_slab_c = nullptr;
    v->d_slab_cap = 0;
    const size_t cap = std::max<size_t>(64, 2 * n);
    int rc;
    if ((rc = dev_alloc(&v->d_slab_w, cap * sizeof(uint32_t*))) || (rc = dev_alloc(&v->d_slab_c, cap * sizeof(uint32_t*))))
      return rc;
    v->d_slab_cap = cap;
  }
  std::vector<uint32_t*> dw(n), dc(n, nullptr);
  for (size_t i = 0; i < n; ++i) {
    VH_HIP(hipHostGetDevicePointer((void**)&dw[i], v->slab_w[i], 0));
    if (v->colors) VH_HIP(hipHostGetDevicePointer((void**)&dc[i], v->slab_c[i], 0));
  }
  VH_HIP(hipMemcpy(v->d_slab_w, dw.data(), n * sizeof(uint32_t*), hipMemcpyHostToDevice));
  VH_HIP(hipMemcpy(v->d_slab_c, dc.data(), n * sizeof(uint32_t*), hipMemcpyHostToDevice));
  v->slab_table_dirty = false;
  return VH_OK;
}

StreamArgs stream_args(vh_volume* v, uint32_t birth) {
  StreamArgs a{};
  a.hash = HashView{v->keys, v->vals, v->cfg.log2_buckets, ((uint64_t)1 << v->cfg.log2_buckets) - 1};
  a.block_keys = v->block_keys;
  a.block_info = v->block_info;
  a.pool = static_cast<uint32_t*>(v->voxels);
  a.colors = v->colors;
  a.wpb = words_per_block(v);
  a.slab_words = v->d_slab_w;
  a.slab_cols = v->d_slab_c;
  a.ctl = v->ctl;
  a.pool_cap = (uint32_t)v->cfg.pool_blocks;
  a.birth = birth;
  a.cap = (float)v->cfg.weight_cap;
  return a;
}

// Squared distance (metres) from the centre of block key k (the mean of its 8^3 lattice
// points, voxel i at i * vox_size) to c.
double block_dist2(unsigned long long k, double vox, const double c[3]) {
  int b[3];
  unpack_key(k, b[0], b[1], b[2]);
  double d2 = 0.0;
  for (int i = 0; i < 3; ++i) {
    const double x = ((double)b[i] * kB + 0.5 * (kB - 1)) * vox - c[i];
    d2 += x * x;
  }
  return d2;
}

// Moves the resident blocks that `out` selects (by pool index, ascending) into the host
// store, compacts the pool and rebuilds the hash table. Queued work must have finished.
int stream_out_blocks(vh_volume* v, const std::vector<unsigned long long>& bk, const std::vector<uint32_t>& out,
                      uint32_t frame_id) {
  const uint32_t N = (uint32_t)bk.size(), n_out = (uint32_t)out.size(), K = N - n_out;
  // kept blocks above K take the vacated indices below K, both in ascending order
  std::vector<uint32_t> holes, srcs;
  for (uint32_t b : out)
    if (b < K) holes.push_back(b);
  {
    size_t j = 0;
    for (uint32_t b = K; b < N; ++b) {
      while (j < out.size() && out[j] < b) ++j;
      if (j < out.size() && out[j] == b) continue;
      srcs.push_back(b);
    }
  }
  if (holes.size() != srcs.size()) return fail(VH_ERR_DEVICE, "stream-out compaction mismatch");
  std::vector<uint32_t> slots(n_out);
  uint32_t got = 0;  // slots allocated so far (returned to the free list on failure)
  int rc = ensure_stream_index(v, std::max<size_t>(std::max<size_t>(n_out, holes.size()), 1));
  for (; !rc && got < n_out; ++got)
    if ((rc = store_alloc_slot(v, &slots[got]))) break;
  if (rc || (rc = upload_slab_table(v))) {
    for (uint32_t i = 0; i < got; ++i) v->free_slots.push_back(slots[i]);
    return rc;
  }
  StreamArgs a = stream_args(v, frame_id - 1u);
  // 1. the outgoing blocks' words and colours straight into their host slots
  VH_HIP(hipMemcpyAsync(v->st_idx, out.data(), (size_t)n_out * 4, hipMemcpyHostToDevice, v->stream));
  VH_HIP(hipMemcpyAsync(v->st_idx + v->st_cap, slots.data(), (size_t)n_out * 4, hipMemcpyHostToDevice, v->stream));
  a.n = n_out;
  a.src = v->st_idx;
  a.slots = v->st_idx + v->st_cap;
  VH_HIP(launch_stream_out(a, v->stream));
  VH_HIP(hipStreamSynchronize(v->stream));
  for (uint32_t i = 0; i < n_out; ++i) {
    v->slot_key[slots[i]] = bk[out[i]];
    v->store_index[bk[out[i]]] = slots[i];
  }
  // 2. compaction
  if (!holes.empty()) {
    VH_HIP(hipMemcpyAsync(v->st_idx, srcs.data(), holes.size() * 4, hipMemcpyHostToDevice, v->stream));
    VH_HIP(hipMemcpyAsync(v->st_idx + v->st_cap, holes.data(), holes.size() * 4, hipMemcpyHostToDevice, v->stream));
    a.n = (uint32_t)holes.size();
    a.src = v->st_idx;
    a.dst = v->st_idx + v->st_cap;
    VH_HIP(launch_stream_move(a, v->stream));
  }
  // 3. the hash table from the K remaining keys; the pool counter
  VH_HIP(hipMemsetAsync(v->keys, 0xff, v->n_slots * sizeof(unsigned long long), v->stream));
  a.n = K;
  VH_HIP(launch_stream_rehash(a, v->stream));
  VH_HIP(hipMemcpyAsync(&v->ctl->pool_count, &K, sizeof K, hipMemcpyHostToDevice, v->stream));
  v->streamed_out += n_out;
  return sync_and_check(v);
}

// Moves the host-store blocks at `slots` into the pool (inserting new keys, merging keys that
// are resident again) and frees their slots; n_new of them are not resident (each takes a pool
// block, the others merge in place). Queued work must have finished.
int stream_in_blocks(vh_volume* v, const std::vector<uint32_t>& slots, uint32_t frame_id, uint32_t pool_count,
                     int64_t* moved, uint32_t n_new) {
  const uint32_t n = (uint32_t)slots.size();
  *moved = 0;
  if (!n) return VH_OK;
  if ((uint64_t)pool_count + n_new > (uint64_t)v->cfg.pool_blocks)
    return fail(VH_ERR_POOL_FULL, "the pool cannot take the blocks to stream in (nothing moved)");
  int rc = ensure_stream_index(v, n);
  if (rc || (rc = upload_slab_table(v))) return rc;
  std::vector<unsigned long long> keys(n);
  for (uint32_t i = 0; i < n; ++i) keys[i] = v->slot_key[slots[i]];
  StreamArgs a = stream_args(v, frame_id - 1u);
  VH_HIP(hipMemcpyAsync(v->st_keys, keys.data(), (size_t)n * 8, hipMemcpyHostToDevice, v->stream));
  VH_HIP(hipMemcpyAsync(v->st_idx, slots.data(), (size_t)n * 4, hipMemcpyHostToDevice, v->stream));
  a.n = n;
  a.keys = v->st_keys;
  a.slots = v->st_idx;
  a.placed = v->st_idx + v->st_cap;
  VH_HIP(launch_stream_insert(a, v->cfg.storage, v->stream));
  std::vector<uint32_t> placed(n);
  VH_HIP(hipMemcpyAsync(placed.data(), a.placed, (size_t)n * 4, hipMemcpyDeviceToHost, v->stream));
  VH_HIP(hipStreamSynchronize(v->stream));
  for (uint32_t i = 0; i < n; ++i) {
    if (placed[i] == kNoBlock) continue;
    store_free_slot(v, slots[i]);
    ++*moved;
    if (placed[i] & kMergeBit) ++v->stream_merged;
  }
  v->streamed_in += *moved;
  return sync_and_check(v);
}

// Of the host-store slots `slots`, how many hold a key that is not resident (resident: bk).
uint32_t count_new_keys(const vh_volume* v, const std::vector<uint32_t>& slots, const std::vector<unsigned long long>& bk) {
  std::unordered_set<unsigned long long> res(bk.begin(), bk.end());
  uint32_t n = 0;
  for (uint32_t s : slots) n += res.count(v->slot_key[s]) ? 0u : 1u;
  return n;
}

// Merges into the resident blocks every host-store block whose key is resident again (the
// weighted mean of vh_stream_in), so that no key is in both places. Queued work must have
// finished; bk: the resident keys.
int merge_resident_duplicates(vh_volume* v, const std::vector<unsigned long long>& bk, uint32_t frame_id,
                              uint32_t pool_count) {
  if (v->store_index.empty()) return VH_OK;
  std::vector<uint32_t> dup;
  for (unsigned long long k : bk) {
    auto it = v->store_index.find(k);
    if (it != v->store_index.end()) dup.push_back(it->second);
  }
  if (dup.empty()) return VH_OK;
  std::sort(dup.begin(), dup.end());
  int64_t moved = 0;
  return stream_in_blocks(v, dup, frame_id, pool_count, &moved, 0);
}

int ensure_gn_scratch(vh_volume* v) {
  if (v->gn_partials) return VH_OK;
  int rc = dev_alloc(&v->gn_partials, (size_t)2 * kGnGrid * kGnAcc * sizeof(double));
  if (rc) return rc;
  rc = dev_alloc(&v->gn_pose, 16 * sizeof(double));
  if (rc) return rc;
  rc = dev_alloc(&v->bw_pose, 16 * sizeof(float));
  if (rc) return rc;
  rc = dev_alloc(&v->gn_state, 2 * sizeof(GnState));
  if (rc) return rc;
  return dev_alloc(&v->gn_stats, (size_t)kGnStats * 65 * sizeof(double));
}

// Device scratch of one pose refinement: per-workgroup partial sums, the candidate pose, the
// solver state and one stats row per evaluation.
struct GnScratch {
  double* partials;  // 2 x kGnGrid * kGnAcc
  double* cand;      // 16
  GnState* state;    // 2
  double* stats;     // kGnStats * (iters + 1)
};

// iters Levenberg-Marquardt steps over n device points starting from c2w, then a last
// evaluation; the accepted pose is written back to c2w (FP64, bottom row 0 0 0 1) and the
// per-evaluation figures to *stats (may be NULL). Synchronous; allocates nothing.
int run_gn_refine(vh_volume* v, const float* d_pts, int64_t n, double c2w[16], int32_t iters, vh_gn_stats* stats,
                  const GnScratch& s) {
  GnStart start{};
  for (int k = 0; k < 12; ++k) start.pose[k] = c2w[k];
  start.lambda = 1e-4;
  VH_HIP(launch_gn_start(start, s.state, s.cand, v->stream));
  GnArgs a{};
  a.q = query_args(v);
  a.mu = v->cfg.trunc;
  a.pts = d_pts;
  a.n = n;
  VH_HIP(enqueue_gn_evaluations(v, a, iters, s.partials, s.state, s.cand, s.stats));
  double host_stats[kGnStats * 64];
  VH_HIP(hipMemcpyAsync(c2w, s.cand, 12 * sizeof(double), hipMemcpyDeviceToHost, v->stream));
  VH_HIP(hipMemcpyAsync(host_stats, s.stats, (size_t)kGnStats * (iters + 1) * sizeof(double), hipMemcpyDeviceToHost,
                        v->stream));
  VH_HIP(hipStreamSynchronize(v->stream));
  c2w[12] = c2w[13] = c2w[14] = 0.0;
  c2w[15] = 1.0;
  if (stats) {
    std::memset(stats, 0, sizeof *stats);
    stats->iters = iters;
    for (int it = 0; it <= iters && it < 64; ++it) {
      stats->n_valid[it] = (int32_t)host_stats[kGnStats * it];
      stats->rms[it] = host_stats[kGnStats * it + 1];
      stats->step_norm[it] = host_stats[kGnStats * it + 2];
      for (int k = 0; k < 12; ++k) stats->pose[it][k] = host_stats[kGnStats * it + 5 + k];
    }
  }
  return VH_OK;
}

// Solver scratch of vh_gn_refine_device_noalloc: a file-scope device array, one copy per GPU
// (each GPU holds it, 264 KB, until the process ends), laid out as GnScratch with room for
// kNoAllocMaxIters + 1 stats rows. Host code reaches it only through the per-device address
// that hipGetSymbolAddress returns; the array's name is not a device address on the host.
constexpr int kNoAllocMaxIters = 63;
constexpr size_t kScrPartials = 0;
constexpr size_t kScrCand = kScrPartials + (size_t)2 * kGnGrid * kGnAcc;
constexpr size_t kScrState = kScrCand + 16;
constexpr size_t kScrStats = kScrState + 160;
constexpr size_t kScrDoubles = kScrStats + (size_t)kGnStats * (kNoAllocMaxIters + 1);
static_assert(2 * sizeof(GnState) <= 160 * sizeof(double), "GnState outgrew its scratch slots");
static __device__ __align__(256) double scratch[kScrDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

// Device address of `scratch` on the current device (resolved once per device).
int scratch_for_device(int device, double** out) {
  if (device < 0 || device >= kMaxDevices) return fail(VH_ERR_INVALID_ARG, "device ordinal beyond 63");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    VH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(VH_ERR_DEVICE, "hipGetSymbolAddress returned no address for the solver scratch");
    g_scratch[device] = static_cast<double*>(p);
  }
  *out = g_scratch[device];
  return VH_OK;
}

}  // namespace

extern "C" {

const char* vh_last_error(void) { return g_last_error.c_str(); }
int vh_abi_version(void) { return VH_ABI_VERSION; }

void vh_config_default(vh_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof *c);
  c->width = 640;
  c->height = 480;
  c->fx = 525.0f;
  c->fy = 525.0f;
  c->cx = 319.5f;
  c->cy = 239.5f;
  c->max_depth = 5.0f;
  c->min_depth = 0.1f;
  c->vox_size = 0.01f;
  c->trunc = 0.04f;
  c->log2_buckets = 18;
  c->pool_blocks = 1 << 18;
  c->storage = VH_STORAGE_F32;
  c->weight_cap = 0;
  c->device = 0;
}

int vh_create(const vh_config* cfg, vh_volume** out) {
  if (!cfg || !out) return fail(VH_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  int rc = validate(cfg);
  if (rc) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    return fail(VH_ERR_NO_DEVICE, "no HIP device available");
  }
  if (cfg->device < 0 || cfg->device >= ndev) return fail(VH_ERR_INVALID_ARG, "device ordinal out of range");
  vh_volume* v = new (std::nothrow) vh_volume();
  if (!v) return fail(VH_ERR_OOM, "host allocation failed");
  v->cfg = *cfg;
  v->batch = cfg->seq_batch ? cfg->seq_batch : kMaxBatch;
  v->device = cfg->device;
  if ((rc = set_device(v))) { delete v; return rc; }
  if (hipStreamCreateWithFlags(&v->stream, hipStreamNonBlocking) != hipSuccess) {
    delete v;
    return fail(VH_ERR_DEVICE, "hipStreamCreate failed");
  }
  if (hipHostMalloc((void**)&v->pin, kPinWords * 4, hipHostMallocDefault) != hipSuccess) {
    v->pin = nullptr;
    free_volume(v);
    return fail(VH_ERR_OOM, "hipHostMalloc failed");
  }
  v->n_slots = (uint64_t)kSlotsPerBucket << cfg->log2_buckets;
  // blocks listed per batch, per stripe (one stripe per image region, so shares are uneven):
  // room for four times an even share of the whole pool
  v->stripe_cap = (uint32_t)(cfg->pool_blocks / (kStripes / 4) + 4096);
  const size_t pool = (size_t)cfg->pool_blocks;
  const size_t frame_px = (size_t)cfg->width * cfg->height;
  if ((rc = dev_alloc(&v->keys, v->n_slots * 8)) || (rc = dev_alloc(&v->vals, v->n_slots * 8)) ||
      (rc = dev_alloc(&v->marks, 2 * v->n_slots * 4)) || (rc = dev_alloc(&v->block_keys, pool * 8)) ||
      (rc = dev_alloc(&v->block_info, pool * 8)) || (rc = dev_alloc(&v->dm, (size_t)2 * kMaxBatch * frame_px * 4)) ||
      (cfg->carve && (rc = dev_alloc(&v->tilemax, (size_t)2 * kMaxBatch * ((cfg->width + kTile - 1) / kTile) *
                                                     ((cfg->height + kTile - 1) / kTile) * sizeof(float)))) ||
      (rc = dev_alloc(&v->voxels, pool * kB3 * voxel_bytes(cfg->storage))) ||
      (rc = dev_alloc(&v->touched, (size_t)2 * kStripes * v->stripe_cap * 4)) ||
      (rc = dev_alloc(&v->ctl, sizeof(FrameCtl))) ||
      (rc = dev_alloc(&v->stage_depth, frame_px * 4)) || (rc = dev_alloc(&v->stage_pose, 16 * sizeof(float))) ||
      (cfg->color && ((rc = dev_alloc(&v->colors, pool * kB3 * 4)) || (rc = dev_alloc(&v->stage_rgb, frame_px * 3)) ||
                      (rc = dev_alloc(&v->rgba, 2 * kMaxBatch * frame_px * 4))))) {
    std::string msg = g_last_error;
    free_volume(v);
    return fail(rc, msg);
  }
  hipError_t e = hipMemsetAsync(v->marks, 0, 2 * v->n_slots * 4, v->stream);
  if (e == hipSuccess) e = hipMemsetAsync(v->ctl, 0, sizeof(FrameCtl), v->stream);
  if (e != hipSuccess) {
    free_volume(v);
    return fail(VH_ERR_DEVICE, std::string("hipMemsetAsync failed: ") + hipGetErrorString(e));
  }
  if ((rc = reset_tables(v))) {
    std::string msg = g_last_error;
    free_volume(v);
    return fail(rc, msg);
  }
  *out = v;
  return VH_OK;
}

int vh_destroy(vh_volume* v) {
  if (!v) return fail(VH_ERR_INVALID_ARG, "null volume");
  free_volume(v);
  return VH_OK;
}

int vh_reset(vh_volume* v) {
  if (!v) return fail(VH_ERR_INVALID_ARG, "null volume");
  std::lock_guard<std::mutex> lk(v->mu);
  int rc = set_device(v);
  if (rc) return rc;
  VH_HIP(hipStreamSynchronize(v->stream));
  return reset_tables(v);
}

int vh_integrate(vh_volume* v, const vh_frame* f) {
  if (!v || !f || !f->depth) return fail(VH_ERR_INVALID_ARG, "null argument");
  if (f->depth_format != VH_DEPTH_U16 && f->depth_format != VH_DEPTH_F32)
    return fail(VH_ERR_INVALID_ARG, "unknown depth format");
  if (f->depth_format == VH_DEPTH_U16 && !(f->depth_scale > 0.0f))
    return fail(VH_ERR_INVALID_ARG, "depth_scale must be > 0 for uint16 depth");
  for (int k = 0; k < 16; ++k)
    if (!std::isfinite(f->c2w[k])) return fail(VH_ERR_INVALID_ARG, "pose has a non-finite entry");
  std::lock_guard<std::mutex> lk(v->mu);
  int rc = set_device(v);
  if (rc) return rc;
  const size_t elem = f->depth_format == VH_DEPTH_U16 ? 2 : 4;
  const size_t row = (size_t)v->cfg.width * elem;
  const size_t pitch = f->pitch ? (size_t)f->pitch : row;
  if (pitch < row) return fail(VH_ERR_INVALID_ARG, "pitch smaller than a row");
  if (v->colors && !f->rgb) return fail(VH_ERR_INVALID_ARG, "a colour volume needs the frame's RGB image");
  VH_HIP(hipMemcpy2DAsync(v->stage_depth, row, f->depth, pitch, row, (size_t)v->cfg.height, hipMemcpyHostToDevice,
                          v->stream));
  VH_HIP(hipMemcpyAsync(v->stage_pose, f->c2w, 16 * sizeof(float), hipMemcpyHostToDevice, v->stream));
  if (v->colors)
    VH_HIP(hipMemcpyAsync(v->stage_rgb, f->rgb, (size_t)v->cfg.width * v->cfg.height * 3, hipMemcpyHostToDevice,
                          v->stream));
  const int which = f->depth_format == VH_DEPTH_U16 ? 0 : 1;
  // the staged uint16 graph bakes the depth scale; re-capture when it changes
  if (which == 0 && f->depth_scale != v->graph_scale) {
    drop_graphs(v, 1);
    v->graph_scale = f->depth_scale;
  }
  rc = launch_frame(v, which, f->depth_scale);
  if (rc) return rc;
  return sync_and_check(v);
}

int vh_integrate_async(vh_volume* v, const vh_frame* f) {
  if (!v || !f || !f->depth) return fail(VH_ERR_INVALID_ARG, "null argument");
  if (f->depth_format != VH_DEPTH_U16 && f->depth_format != VH_DEPTH_F32)
    return fail(VH_ERR_INVALID_ARG, "unknown depth format");
  if (f->depth_format == VH_DEPTH_U16 && !(f->depth_scale > 0.0f))
    return fail(VH_ERR_INVALID_ARG, "depth_scale must be > 0 for uint16 depth");
  for (int k = 0; k < 16; ++k)
    if (!std::isfinite(f->c2w[k])) return fail(VH_ERR_INVALID_ARG, "pose has a non-finite entry");
  std::lock_guard<std::mutex> lk(v->mu);
  int rc = set_device(v);
  if (rc) return rc;
  constexpr int R = vh_volume::kRing;
  const size_t elem = f->depth_format == VH_DEPTH_U16 ? 2 : 4;
  const size_t row = (size_t)v->cfg.width * elem;
  const size_t pitch = f->pitch ? (size_t)f->pitch : row;
  if (pitch < row) return fail(VH_ERR_INVALID_ARG, "pitch smaller than a row");
  if (v->colors && !f->rgb) return fail(VH_ERR_INVALID_ARG, "a colour volume needs the frame's RGB image");
  const size_t frame_px = (size_t)v->cfg.width * v->cfg.height;
  if (!v->copy_stream) {
    if ((rc = dev_alloc(&v->ring_depth, (size_t)R * frame_px * 4)) ||
        (rc = dev_alloc(&v->ring_pose, (size_t)R * 16 * sizeof(float))) ||
        (v->colors && (rc = dev_alloc(&v->ring_rgb, (size_t)R * frame_px * 3))))
      return rc;
    VH_HIP(hipHostMalloc((void**)&v->ring_pose_host, (size_t)R * 16 * sizeof(float), hipHostMallocDefault));
    for (int k = 0; k < R; ++k) {
      VH_HIP(hipEventCreateWithFlags(&v->ring_copied[k], hipEventDisableTiming));
      VH_HIP(hipEventCreateWithFlags(&v->ring_done[k], hipEventDisableTiming));
    }
    VH_HIP(hipStreamCreateWithFlags(&v->copy_stream, hipStreamNonBlocking));
  }
  const int which = f->depth_format == VH_DEPTH_U16 ? 6 : 7;
  if (which == 6 && f->depth_scale != v->ring_scale) {  // the ring uint16 graph bakes the scale
    VH_HIP(hipStreamSynchronize(v->stream));
    drop_graphs(v, 1 << 6);
    v->ring_scale = f->depth_scale;
  }
  // The slots are laid out with the element size of the frames in flight: before a frame of
  // the other format is copied, the queued frames finish (their slots would overlap).
  if (v->ring_format >= 0 && v->ring_format != f->depth_format) VH_HIP(hipStreamSynchronize(v->stream));
  v->ring_format = f->depth_format;
  const int slot = (int)(v->ring_next % R);
  if (v->ring_next >= (uint64_t)R) {
    VH_HIP(hipStreamWaitEvent(v->copy_stream, v->ring_done[slot], 0));  // the frame that read the slot
    VH_HIP(hipEventSynchronize(v->ring_copied[slot]));                    // the pose staging is free
  }
  std::memcpy(v->ring_pose_host + 16 * slot, f->c2w, 16 * sizeof(float));
  char* dst = static_cast<char*>(v->ring_depth) + (size_t)slot * v->cfg.width * v->cfg.height * elem;
  VH_HIP(hipMemcpy2DAsync(dst, row, f->depth, pitch, row, (size_t)v->cfg.height, hipMemcpyHostToDevice,
                          v->copy_stream));
  VH_HIP(hipMemcpyAsync(v->ring_pose + 16 * slot, v->ring_pose_host + 16 * slot, 16 * sizeof(float),
                        hipMemcpyHostToDevice, v->copy_stream));
  if (v->colors)
    VH_HIP(hipMemcpyAsync(v->ring_rgb + (size_t)slot * frame_px * 3, f->rgb, frame_px * 3, hipMemcpyHostToDevice,
                          v->copy_stream));
  VH_HIP(hipEventRecord(v->ring_copied[slot], v->copy_stream));
  VH_HIP(hipStreamWaitEvent(v->stream, v->ring_copied[slot], 0));
  if ((rc = launch_frame(v, which, f->depth_scale))) return rc;
  VH_HIP(hipEventRecord(v->ring_done[slot], v->stream));
  ++v->ring_next;
  return VH_OK;
}

int vh_bind_sequence(vh_volume* v, const void* d_depth, int32_t fmt, float scale, const float* d_c2w, int32_t n) {
  if (!v || !d_depth || !d_c2w || n <= 0) return fail(VH_ERR_INVALID_ARG, "bad sequence");
  if (fmt != VH_DEPTH_U16 && fmt != VH_DEPTH_F32) return fail(VH_ERR_INVALID_ARG, "unknown depth format");
  if (fmt == VH_DEPTH_U16 && !(scale > 0.0f)) return fail(VH_ERR_INVALID_ARG, "depth_scale must be > 0");
  std::lock_guard<std::mutex> lk(v->mu);
  int rc = set_device(v);
  if (rc) return rc;
  VH_HIP(hipStreamSynchronize(v->stream));
  const bool same = v->seq_depth == d_depth && v->seq_pose == d_c2w && v->seq_n == n && v->seq_format == fmt &&
                    v->seq_scale == scale;
  v->seq_depth = d_depth;
  v->seq_pose = d_c2w;
  v->seq_n = n;
  v->seq_format = fmt;
  v->seq_scale = scale;
  v->track = false;
  // the captured graphs hold the sequence pointers
  if (!same || v->seq_rgb) drop_graphs(v, 4 | 8 | 16 | 32);
  v->seq_rgb = nullptr;
  const uint32_t zero = 0;
  VH_HIP(hipMemcpyAsync(&v->ctl->cursor[1], &zero, sizeof zero, hipMemcpyHostToDevice, v->stream));
  VH_HIP(hipStreamSynchronize(v->stream));
  return VH_OK;
}

int vh_bind_tracking(vh_volume* v, const void* d_depth, int32_t fmt, float scale, int32_t n,
                     const vh_track_config* tc) {
  if (!v || !d_depth || !tc || !tc->d_prior_c2w || !tc->d_out_c2w || n <= 0)
    return fail(VH_ERR_INVALID_ARG, "bad sequence");
  if (tc->gn_iters < 0 || tc->gn_iters > 63) return fail(VH_ERR_INVALID_ARG, "gn_iters must be in 0..63");
  if (tc->point_stride < 1 || tc->point_stride > 64)
    return fail(VH_ERR_INVALID_ARG, "point_stride must be in 1..64");
  if (std::isnan(tc->max_point_depth)) return fail(VH_ERR_INVALID_ARG, "max_point_depth is NaN");
  if (tc->prior_mode != VH_PRIOR_ABSOLUTE && tc->prior_mode != VH_PRIOR_INCREMENT)
    return fail(VH_ERR_INVALID_ARG, "unknown prior_mode");
  for (float x : {tc->prior_sigma_t, tc->prior_sigma_rot, tc->depth_points})
    if (!(x >= 0.0f) || !std::isfinite(x))
      return fail(VH_ERR_INVALID_ARG, "prior_sigma_t, prior_sigma_rot and depth_points must be finite and >= 0");
  for (int32_t r : tc->reserved)
    if (r) return fail(VH_ERR_INVALID_ARG, "reserved fields must be zero");
  if (fmt != VH_DEPTH_U16 && fmt != VH_DEPTH_F32) return fail(VH_ERR_INVALID_ARG, "unknown depth format");
  if (fmt == VH_DEPTH_U16 && !(scale > 0.0f)) return fail(VH_ERR_INVALID_ARG, "depth_scale must be > 0");
  std::lock_guard<std::mutex> lk(v->mu);
  int rc = set_device(v);
  if (rc) return rc;
  VH_HIP(hipStreamSynchronize(v->stream));
  if ((rc = ensure_gn_scratch(v))) return rc;
  const int stride = tc->point_stride;
  const size_t npts = (size_t)((v->cfg.width + stride - 1) / stride) * ((v->cfg.height + stride - 1) / stride);
  if (npts > v->trk_pts_cap) {
    if (v->trk_pts) (void)hipFree(v->trk_pts);
    v->trk_pts = nullptr;
    v->trk_pts_cap = 0;
    drop_graphs(v, 8);
    if ((rc = dev_alloc(&v->trk_pts, npts * 3 * sizeof(float)))) return rc;
    v->trk_pts_cap = npts;
  }
  const float zmax = std::min(v->cfg.max_depth, tc->max_point_depth > 0.0f ? tc->max_point_depth : v->cfg.max_depth);
  const int incr = tc->prior_mode == VH_PRIOR_INCREMENT;
  const bool same = v->track && v->seq_depth == d_depth && v->trk_prior == tc->d_prior_c2w &&
                    v->trk_out == tc->d_out_c2w && v->trk_stats == tc->d_stats && v->seq_n == n &&
                    v->seq_format == fmt && v->seq_scale == scale && v->trk_iters == tc->gn_iters &&
                    v->trk_stride == stride && v->trk_max_depth == zmax && v->trk_increment == incr &&
                    v->trk_sigma_t == tc->prior_sigma_t && v->trk_sigma_r == tc->prior_sigma_rot &&
                    v->trk_points == tc->depth_points;
  v->seq_depth = d_depth;
  v->seq_pose = tc->d_prior_c2w;
  v->seq_n = n;
  v->seq_format = fmt;
  v->seq_scale = scale;
  v->track = true;
  v->trk_prior = tc->d_prior_c2w;
  v->trk_out = tc->d_out_c2w;
  v->trk_stats = tc->d_stats;
  v->trk_iters = tc->gn_iters;
  v->trk_stride = stride;
  v->trk_max_depth = zmax;
  v->trk_increment = incr;
  v->trk_sigma_t = tc->prior_sigma_t;
  v->trk_sigma_r = tc->prior_sigma_rot;
  v->trk_points = tc->depth_points;
  if (!same || v->seq_rgb) drop_graphs(v, 4 | 8 | 16 | 32);
  v->seq_rgb = nullptr;
  const uint32_t zero = 0;
  VH_HIP(hipMemcpyAsync(&v->ctl->cursor[1], &zero, sizeof zero, hipMemcpyHostToDevice, v->stream));
  VH_HIP(hipStreamSynchronize(v->stream));
  return VH_OK;
}

int vh_bind_sequence_rgb(vh_volume* v, const uint8_t* d_rgb) {
  if (!v || !d_rgb) return fail(VH_ERR_INVALID_ARG, "null argument");
  if (!v->colors) return fail(VH_ERR_INVALID_ARG, "the volume has no colour");
  std::lock_guard<std::mutex> lk(v->mu);
  if (!v->seq_depth) return fail(VH_ERR_INVALID_ARG, "no sequence bound");
  int rc = set_device(v);
  if (rc) return rc;
  if (v->seq_rgb != d_rgb) {
    VH_HIP(hipStreamSynchronize(v->stream));
    drop_graphs(v, 4 | 8 | 16 | 32);
    v->seq_rgb = d_rgb;
  }
  return VH_OK;
}

int vh_integrate_next(vh_volume* v, int32_t n) {
  if (!v || n < 0) return fail(VH_ERR_INVALID_ARG, "bad argument");
  if (!v->seq_depth) return fail(VH_ERR_INVALID_ARG, "no sequence bound");
  if (v->colors && !v->seq_rgb)
    return fail(VH_ERR_INVALID_ARG, "a colour volume needs the sequence's RGB images (vh_bind_sequence_rgb)");
  std::lock_guard<std::mutex> lk(v->mu);
  int rc = set_device(v);
  if (rc) return rc;
  if (v->track) {
    // per frame: tracking, allocation, integration (each pose depends on the volume so far)
    for (int k = 0; k < n; ++k) {
      rc = launch_frame(v, 3, v->seq_scale);
      if (rc) return rc;
    }
    return VH_OK;
  }
  // batches of up to `batch` frames; the last may be shorter
  const int K = v->batch;
  if (v->profiling == 1) {
    for (int f = 0; f < n; f += K)
      if ((rc = launch_frame(v, 2, v->seq_scale, std::min(K, n - f), std::min(K, n - f)))) return rc;
    return VH_OK;
  }
  // allocation of batch b+1 overlaps the integration of batch b; nothing stays allocated
  // but unintegrated when the call returns
  if (n == 0) return VH_OK;
  int prev = std::min(K, n);
  if ((rc = launch_frame(v, 4, v->seq_scale, 0, prev))) return rc;
  for (int f = prev; f < n; f += K) {
    const int cur = std::min(K, n - f);
    if ((rc = launch_frame(v, 2, v->seq_scale, prev, cur))) return rc;
    prev = cur;
  }
  return launch_frame(v, 5, v->seq_scale, prev, 0);
}

int vh_sync(vh_volume* v) {
  if (!v) return fail(VH_ERR_INVALID_ARG, "null volume");
  std::lock_guard<std::mutex> lk(v->mu);
  int rc = set_device(v);
  if (rc) return rc;
  return sync_and_check(v);
}

int vh_set_profiling(vh_volume* v, int32_t on) {
  if (!v) return fail(VH_ERR_INVALID_ARG, "null volume");
  std::lock_guard<std::mutex> lk(v->mu);
  int rc = set_device(v);
  if (rc) return rc;
  VH_HIP(hipStreamSynchronize(v->stream));
  rc = collect_profile(v);
  if (rc) return rc;
  if (on < 0 || on > 2) return fail(VH_ERR_INVALID_ARG, "profiling mode must be 0, 1 or 2");
  v->profiling = on;
  v->alloc_ms = v->integrate_ms = v->track_ms = v->carve_ms = v->fused_ms = 0.0;
  v->profiled_frames = v->profiled_launches = v->fused_launches = 0;
  return VH_OK;
}

int vh_get_stats(vh_volume* v, vh_stats* s) {
  if (!v || !s) return fail(VH_ERR_INVALID_ARG, "null argument");
  std::lock_guard<std::mutex> lk(v->mu);
  int rc = set_device(v);
  if (rc) return rc;
  VH_HIP(hipStreamSynchronize(v->stream));
  if ((rc = collect_profile(v))) return rc;
  FrameCtl c{};
  if ((rc = read_ctl(v, &c))) return rc;
  std::memset(s, 0, sizeof *s);
  s->frames = (int64_t)c.frames;
  s->blocks_allocated = std::min<int64_t>(c.pool_count, v->cfg.pool_blocks);
  s->pool_capacity = v->cfg.pool_blocks;
  s->slots_used = c.pool_count;
  s->touched_last = c.touched_last;
  s->touched_total = (int64_t)c.touched_total;
  s->new_total = s->blocks_allocated;
  s->storage = v->cfg.storage;
  s->flags = (int32_t)c.flags;
  s->host_blocks = (int64_t)v->store_index.size();
  s->streamed_out_total = v->streamed_out;
  s->streamed_in_total = v->streamed_in;
  s->stream_merged_total = v->stream_merged;
  s->integrate_ms = v->integrate_ms;
  s->alloc_ms = v->alloc_ms;
  s->profiled_frames = v->profiled_frames;
  s->profiled_launches = v->profiled_launches;
  s->blocks_integrated = (int64_t)c.integrated_total;
  s->carved_total = (int64_t)c.carved_total;
  s->carve_ms = v->carve_ms;
  s->fused_ms = v->fused_ms;
  s->fused_launches = v->fused_launches;
  s->track_ms = v->track_ms;
  return VH_OK;
}

void* vh_stream(vh_volume* v) { return v ? (void*)v->stream : nullptr; }

int vh_extract_mesh_device_rgb(vh_volume* v, float* d_xyz, int32_t* d_cubes, uint8_t* d_rgb, int64_t capacity,
                               int64_t* n_tris) {
  if (!v || !n_tris || capacity < 0 || (capacity > 0 && !d_xyz)) return fail(VH_ERR_INVALID_ARG, "bad argument");
  if (d_rgb && !v->colors) return fail(VH_ERR_INVALID_ARG, "the volume has no colour");
  std::lock_guard<std::mutex> lk(v->mu);
  int rc = set_device(v);
  if (rc) return rc;
  return run_mc(v, d_xyz, d_cubes, d_rgb, capacity, n_tris);
}

int vh_extract_mesh_device(vh_volume* v, float* d_xyz, int32_t* d_cubes, int64_t capacity, int64_t* n_tris) {
  return vh_extract_mesh_device_rgb(v, d_xyz, d_cubes, nullptr, capacity, n_tris);
}

int vh_extract_mesh(vh_volume* v, vh_mesh* out) {
  if (!v || !out) return fail(VH_ERR_INVALID_ARG, "null argument");
  std::memset(out, 0, sizeof *out);
  std::lock_guard<std::mutex> lk(v->mu);
  int rc = set_device(v);
  if (rc) return rc;
  int64_t n = 0;
  if ((rc = ensure_mc_buffer(v, std::max<int64_t>(v->mc_cap, 1 << 20)))) return rc;
  rc = run_mc(v, v->mc_tris, v->mc_cubes, v->mc_rgb, v->mc_cap, &n);
  if (rc == VH_ERR_TRUNCATED) {
    if ((rc = ensure_mc_buffer(v, n + (n >> 3) + 1024))) return rc;
    rc = run_mc(v, v->mc_tris, v->mc_cubes, v->mc_rgb, v->mc_cap, &n);
  }
  if (rc) return rc;
  out->n_tris = n;
  const size_t nt = (size_t)std::max<int64_t>(n, 1);
  out->xyz = static_cast<float*>(std::malloc(nt * 9 * sizeof(float)));
  out->cubes = static_cast<int32_t*>(std::malloc(nt * 4 * sizeof(int32_t)));
  if (v->colors) out->rgb = static_cast<uint8_t*>(std::malloc(nt * 9));
  if (!out->xyz || !out->cubes || (v->colors && !out->rgb)) {
    vh_mesh_free(out);
    return fail(VH_ERR_OOM, "host allocation failed");
  }
  if (n) {
    VH_HIP(hipMemcpy(out->xyz, v->mc_tris, (size_t)n * 9 * sizeof(float), hipMemcpyDeviceToHost));
    VH_HIP(hipMemcpy(out->cubes, v->mc_cubes, (size_t)n * 4 * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (out->rgb) VH_HIP(hipMemcpy(out->rgb, v->mc_rgb, (size_t)n * 9, hipMemcpyDeviceToHost));
  }
  return VH_OK;
}

int vh_extract_mesh_indexed_device(vh_volume* v, float* d_vertices, int64_t vertex_capacity, int32_t* d_faces,
                                   int64_t face_capacity, uint8_t* d_vrgb, int64_t* n_vertices, int64_t* n_faces) {
  if (!v || !n_vertices || !n_faces || vertex_capacity < 0 || face_capacity < 0 ||
      (vertex_capacity > 0 && !d_vertices) || (face_capacity > 0 && !d_faces))
    return fail(VH_ERR_INVALID_ARG, "bad argument");
  if (d_vrgb && !v->colors) return fail(VH_ERR_INVALID_ARG, "the volume has no colour");
  std::lock_guard<std::mutex> lk(v->mu);
  int rc = set_device(v);
  if (rc) return rc;
  McArgs a;
  int64_t nf = 0, nv = 0;
  if ((rc = index_count(v, a, &nf, &nv))) return rc;
  *n_faces = nf;
  *n_vertices = nv;
  if (nv > vertex_capacity || nf > face_capacity)
    return fail(VH_ERR_TRUNCATED, "indexed mesh larger than the output buffers");
  return index_write(v, a, d_faces, nf, d_vertices, d_vrgb, nv);
}

int vh_extract_mesh_indexed(vh_volume* v, vh_indexed_mesh* out) {
  if (!v || !out) return fail(VH_ERR_INVALID_ARG, "null argument");
  std::memset(out, 0, sizeof *out);
  std::lock_guard<std::mutex> lk(v->mu);
  int rc = set_device(v);
  if (rc) return rc;
  McArgs a;
  int64_t n = 0, nv = 0;
  if ((rc = index_count(v, a, &n, &nv))) return rc;
  // device output of the host variant, grown with 1/8 slack
  if (n > v->ix_fcap || nv > v->ix_vcap) {
    const int64_t fcap = std::max(n + (n >> 3), v->ix_fcap), vcap = std::max(nv + (nv >> 3), v->ix_vcap);
    void* old[] = {v->ix_faces, v->ix_verts, v->ix_vrgb};
    for (void* p : old)
      if (p) (void)hipFree(p);
    v->ix_faces = nullptr;
    v->ix_verts = nullptr;
    v->ix_vrgb = nullptr;
    v->ix_fcap = v->ix_vcap = 0;
    if ((rc = dev_alloc(&v->ix_faces, (size_t)fcap * 3 * 4)) || (rc = dev_alloc(&v->ix_verts, (size_t)vcap * 3 * 4)) ||
        (v->colors && (rc = dev_alloc(&v->ix_vrgb, (size_t)vcap * 3))))
      return rc;
    v->ix_fcap = fcap;
    v->ix_vcap = vcap;
  }
  if ((rc = index_write(v, a, v->ix_faces, n, v->ix_verts, v->ix_vrgb, nv))) return rc;
  out->n_vertices = nv;
  out->n_faces = n;
  const size_t cv = (size_t)std::max<int64_t>(nv, 1) * 3, cf = (size_t)std::max<int64_t>(n, 1) * 3;
  out->vertices = static_cast<float*>(std::malloc(cv * sizeof(float)));
  out->faces = static_cast<int32_t*>(std::malloc(cf * sizeof(int32_t)));
  if (v->colors) out->rgb = static_cast<uint8_t*>(std::malloc(cv));
  if (!out->vertices || !out->faces || (v->colors && !out->rgb)) {
    vh_indexed_mesh_free(out);
    return fail(VH_ERR_OOM, "host allocation failed");
  }
  if (nv) VH_HIP(hipMemcpy(out->vertices, v->ix_verts, (size_t)nv * 3 * sizeof(float), hipMemcpyDeviceToHost));
  if (nv && out->rgb) VH_HIP(hipMemcpy(out->rgb, v->ix_vrgb, (size_t)nv * 3, hipMemcpyDeviceToHost));
  if (n) VH_HIP(hipMemcpy(out->faces, v->ix_faces, (size_t)n * 3 * sizeof(int32_t), hipMemcpyDeviceToHost));
  return VH_OK;
}

int vh_indexed_mesh_free(vh_indexed_mesh* m) {
  if (!m) return fail(VH_ERR_INVALID_ARG, "null mesh");
  std::free(m->vertices);
  std::free(m->faces);
  std::free(m->rgb);
  std::memset(m, 0, sizeof *m);
  return VH_OK;
}

int vh_mesh_free(vh_mesh* m) {
  if (!m) return fail(VH_ERR_INVALID_ARG, "null mesh");
  std::free(m->xyz);
  std::free(m->cubes);
  std::free(m->rgb);
  std::memset(m, 0, sizeof *m);
  return VH_OK;
}

int vh_save_ply(vh_volume* v, const char* path) {
  if (!v || !path) return fail(VH_ERR_INVALID_ARG, "null argument");
  vh_indexed_mesh m{};
  int rc = vh_extract_mesh_indexed(v, &m);
  if (rc) return rc;
  FILE* f = std::fopen(path, "wb");
  if (!f) {
    vh_indexed_mesh_free(&m);
    return fail(VH_ERR_IO, std::string("cannot open ") + path);
  }
  std::fprintf(f,
               "ply\nformat binary_little_endian 1.0\nelement vertex %lld\nproperty float x\nproperty float y\n"
               "property float z\n%selement face %lld\nproperty list uchar int vertex_indices\nend_header\n",
               (long long)m.n_vertices, m.rgb ? "property uchar red\nproperty uchar green\nproperty uchar blue\n" : "",
               (long long)m.n_faces);
  bool ok = true;
  if (m.rgb) {
    // vertex records of 15 bytes: x, y, z, r, g, b
    std::vector<unsigned char> rec((size_t)m.n_vertices * 15);
    for (size_t i = 0; i < (size_t)m.n_vertices; ++i) {
      std::memcpy(&rec[15 * i], m.vertices + 3 * i, 12);
      std::memcpy(&rec[15 * i + 12], m.rgb + 3 * i, 3);
    }
    ok = std::fwrite(rec.data(), 1, rec.size(), f) == rec.size();
  } else {
    ok = std::fwrite(m.vertices, sizeof(float), (size_t)m.n_vertices * 3, f) == (size_t)m.n_vertices * 3;
  }
  // face records of 13 bytes: the count 3, then three int32 indices
  std::vector<unsigned char> frec((size_t)m.n_faces * 13);
  for (size_t t = 0; t < (size_t)m.n_faces; ++t) {
    frec[13 * t] = 3;
    std::memcpy(&frec[13 * t + 1], m.faces + 3 * t, 12);
  }
  ok = ok && std::fwrite(frec.data(), 1, frec.size(), f) == frec.size();
  ok = (std::fclose(f) == 0) && ok;
  vh_indexed_mesh_free(&m);
  return ok ? VH_OK : fail(VH_ERR_IO, "write failed");
}

int vh_export_blocks(vh_volume* v, int32_t* keys, float* tsdf, float* weight, int64_t capacity, int64_t* n) {
  if (!v || !n || capacity < 0) return fail(VH_ERR_INVALID_ARG, "bad argument");
  std::lock_guard<std::mutex> lk(v->mu);
  int rc = set_device(v);
  if (rc) return rc;
  FrameCtl c{};
  if ((rc = read_ctl(v, &c))) return rc;
  const int64_t nb = std::min<int64_t>(c.pool_count, v->cfg.pool_blocks);
  *n = nb;
  if (nb > capacity) return fail(VH_ERR_TRUNCATED, "export buffer too small");
  if (nb == 0) return VH_OK;
  std::vector<unsigned long long> bk((size_t)nb);
  VH_HIP(hipMemcpy(bk.data(), v->block_keys, (size_t)nb * 8, hipMemcpyDeviceToHost));
  if (keys)
    for (int64_t b = 0; b < nb; ++b) unpack_key(bk[(size_t)b], keys[3 * b], keys[3 * b + 1], keys[3 * b + 2]);
  if (tsdf || weight) {
    const size_t nv = (size_t)nb * kB3;
    if (v->cfg.storage == VH_STORAGE_F32) {
      std::vector<float> raw(nv * 2);
      VH_HIP(hipMemcpy(raw.data(), v->voxels, nv * 8, hipMemcpyDeviceToHost));
      for (size_t i = 0; i < nv; ++i) {
        if (tsdf) tsdf[i] = raw[2 * i];
        if (weight) weight[i] = raw[2 * i + 1];
      }
    } else {
      std::vector<uint32_t> raw(nv);
      VH_HIP(hipMemcpy(raw.data(), v->voxels, nv * 4, hipMemcpyDeviceToHost));
      const bool comp = v->cfg.storage == VH_STORAGE_F16R;
      for (size_t i = 0; i < nv; ++i) {
        const uint16_t h = (uint16_t)(raw[i] & 0xffffu);
        if (tsdf) tsdf[i] = comp ? f16r_decode(h, (int)(int8_t)(raw[i] >> 24)) : f16_bits_to_f32(h);
        if (weight) weight[i] = comp ? (float)((raw[i] >> 16) & 0xffu) : (float)(raw[i] >> 16);
      }
    }
  }
  return VH_OK;
}

int vh_export_block_colors(vh_volume* v, uint8_t* rgb, int64_t capacity, int64_t* n) {
  if (!v || !n || capacity < 0) return fail(VH_ERR_INVALID_ARG, "bad argument");
  if (!v->colors) return fail(VH_ERR_INVALID_ARG, "the volume has no colour");
  std::lock_guard<std::mutex> lk(v->mu);
  int rc = set_device(v);
  if (rc) return rc;
  FrameCtl c{};
  if ((rc = read_ctl(v, &c))) return rc;
  const int64_t nb = std::min<int64_t>(c.pool_count, v->cfg.pool_blocks);
  *n = nb;
  if (nb > capacity) return fail(VH_ERR_TRUNCATED, "export buffer too small");
  if (nb == 0 || !rgb) return VH_OK;
  const size_t nv = (size_t)nb * kB3;
  std::vector<uint32_t> raw(nv);
  VH_HIP(hipMemcpy(raw.data(), v->colors, nv * 4, hipMemcpyDeviceToHost));
  for (size_t i = 0; i < nv; ++i)
    for (int k = 0; k < 3; ++k) rgb[3 * i + k] = (uint8_t)(raw[i] >> (8 * k));
  return VH_OK;
}

int vh_query_sdf_device(vh_volume* v, const float* d_pts, int64_t n, float* d_sdf, uint8_t* d_valid) {
  if (!v || n < 0 || (n > 0 && !d_pts)) return fail(VH_ERR_INVALID_ARG, "bad argument");
  std::lock_guard<std::mutex> lk(v->mu);
  int rc = set_device(v);
  if (rc) return rc;
  if (n == 0) return VH_OK;
  VH_HIP(launch_query(query_args(v), v->cfg.storage, d_pts, n, d_sdf, d_valid, v->stream));
  return VH_OK;
}

int vh_query_sdf(vh_volume* v, const float* pts, int64_t n, float* sdf, uint8_t* valid) {
  if (!v || n < 0 || (n > 0 && (!pts || !sdf))) return fail(VH_ERR_INVALID_ARG, "bad argument");
  if (n == 0) return VH_OK;
  int rc = set_device(v);
  if (rc) return rc;
  float* d_pts = nullptr;
  float* d_sdf = nullptr;
  uint8_t* d_valid = nullptr;
  if ((rc = dev_alloc(&d_pts, (size_t)n * 12)) || (rc = dev_alloc(&d_sdf, (size_t)n * 4)) ||
      (rc = dev_alloc(&d_valid, (size_t)n))) {
    if (d_pts) (void)hipFree(d_pts);
    if (d_sdf) (void)hipFree(d_sdf);
    return rc;
  }
  struct Guard {
    void* p[3];
    ~Guard() { for (void* q : p) if (q) (void)hipFree(q); }
  } guard{{d_pts, d_sdf, d_valid}};
  VH_HIP(hipMemcpyAsync(d_pts, pts, (size_t)n * 12, hipMemcpyHostToDevice, v->stream));
  rc = vh_query_sdf_device(v, d_pts, n, d_sdf, d_valid);
  if (rc) return rc;
  VH_HIP(hipMemcpyAsync(sdf, d_sdf, (size_t)n * 4, hipMemcpyDeviceToHost, v->stream));
  if (valid) VH_HIP(hipMemcpyAsync(valid, d_valid, (size_t)n, hipMemcpyDeviceToHost, v->stream));
  VH_HIP(hipStreamSynchronize(v->stream));
  return VH_OK;
}

int vh_query_sdf_backward_device(vh_volume* v, const float* d_pts, int64_t n, const float* c2w, const float* d_grad_r,
                                 float* d_r, uint8_t* d_valid, float* d_J, double* d_pose_grad, float* d_voxel_grad) {
  if (!v || !c2w || n < 0 || (n > 0 && !d_pts)) return fail(VH_ERR_INVALID_ARG, "bad argument");
  if ((d_pose_grad || d_voxel_grad) && !d_grad_r) return fail(VH_ERR_INVALID_ARG, "gradients need d_grad_r");
  std::lock_guard<std::mutex> lk(v->mu);
  int rc = set_device(v);
  if (rc) return rc;
  if ((rc = ensure_gn_scratch(v))) return rc;
  float* d_T = v->bw_pose;
  VH_HIP(hipMemcpyAsync(d_T, c2w, 12 * sizeof(float), hipMemcpyHostToDevice, v->stream));
  VH_HIP(launch_backward(query_args(v), v->cfg.storage, d_pts, n, d_T, d_grad_r, d_r, d_valid, d_J,
                         d_pose_grad ? v->gn_partials : nullptr, kGnGrid, d_voxel_grad, v->stream));
  if (d_pose_grad) VH_HIP(launch_reduce6(v->gn_partials, kGnGrid, d_pose_grad, v->stream));
  // the scratch pose must outlive the kernels before the next call reuses it
  VH_HIP(hipStreamSynchronize(v->stream));
  return VH_OK;
}

int vh_gn_refine(vh_volume* v, const float* pts, int64_t n, double c2w[16], int32_t iters, vh_gn_stats* stats) {
  if (!v || !pts || !c2w || n <= 0 || iters < 0 || iters > 63) return fail(VH_ERR_INVALID_ARG, "bad argument");
  for (int k = 0; k < 16; ++k)
    if (!std::isfinite(c2w[k])) return fail(VH_ERR_INVALID_ARG, "pose has a non-finite entry");
  std::lock_guard<std::mutex> lk(v->mu);
  int rc = set_device(v);
  if (rc) return rc;
  if ((rc = ensure_gn_scratch(v))) return rc;
  float* d_pts = nullptr;
  if ((rc = dev_alloc(&d_pts, (size_t)n * 12))) return rc;
  struct Guard {
    void* p;
    ~Guard() { if (p) (void)hipFree(p); }
  } guard{d_pts};
  VH_HIP(hipMemcpyAsync(d_pts, pts, (size_t)n * 12, hipMemcpyHostToDevice, v->stream));
  return run_gn_refine(v, d_pts, n, c2w, iters, stats,
                       GnScratch{v->gn_partials, v->gn_pose, v->gn_state, v->gn_stats});
}

int vh_gn_refine_device_noalloc(vh_volume* v, const float* d_pts, int64_t n, double c2w[16], int32_t iters,
                                vh_gn_stats* stats) {
  if (!v || !d_pts || !c2w || n <= 0 || iters < 0) return fail(VH_ERR_INVALID_ARG, "bad argument");
  if (iters > kNoAllocMaxIters)
    return fail(VH_ERR_INVALID_ARG, "iters exceeds 63, the most the static solver scratch holds");
  for (int k = 0; k < 16; ++k)
    if (!std::isfinite(c2w[k])) return fail(VH_ERR_INVALID_ARG, "pose has a non-finite entry");
  std::lock_guard<std::mutex> lk(v->mu);
  int rc = set_device(v);
  if (rc) return rc;
  double* base = nullptr;
  if ((rc = scratch_for_device(v->device, &base))) return rc;
  return run_gn_refine(v, d_pts, n, c2w, iters, stats,
                       GnScratch{base + kScrPartials, base + kScrCand, reinterpret_cast<GnState*>(base + kScrState),
                                 base + kScrStats});
}

int vh_pool_view(vh_volume* v, void** d_voxels, void** d_block_keys, int64_t* n_blocks) {
  if (!v) return fail(VH_ERR_INVALID_ARG, "null volume");
  if (d_voxels) *d_voxels = v->voxels;
  if (d_block_keys) *d_block_keys = v->block_keys;
  if (n_blocks) {
    std::lock_guard<std::mutex> lk(v->mu);
    int rc = set_device(v);
    if (rc) return rc;
    FrameCtl c{};
    if ((rc = read_ctl(v, &c))) return rc;
    *n_blocks = std::min<int64_t>(c.pool_count, v->cfg.pool_blocks);
  }
  return VH_OK;
}

// ------------------------------------------------------------------ out-of-core streaming
namespace {
bool finite3(const float c[3]) { return std::isfinite(c[0]) && std::isfinite(c[1]) && std::isfinite(c[2]); }

// The control block and the resident blocks' keys in pool order; queued work has finished.
int resident_blocks(vh_volume* v, FrameCtl& c, std::vector<unsigned long long>& bk) {
  int rc = read_ctl(v, &c);
  if (rc) return rc;
  const uint32_t N = std::min<uint32_t>(c.pool_count, (uint32_t)v->cfg.pool_blocks);
  bk.resize(N);
  if (N) VH_HIP(hipMemcpy(bk.data(), v->block_keys, (size_t)N * 8, hipMemcpyDeviceToHost));
  return VH_OK;
}

struct CkptHeader {
  char magic[8];
  int32_t abi;
  int32_t words_per_block;
  vh_config cfg;
  uint32_t frame_id;
  uint32_t reserved;
  uint64_t frames;
  int64_t n_blocks;
};
constexpr char kCkptMagic[8] = {'V', 'H', 'V', 'O', 'L', 'U', 'M', 'E'};
}  // namespace

int vh_stream_out(vh_volume* v, const float center[3], float radius, int64_t* n_blocks) {
  if (!v || !center || !n_blocks || !finite3(center) || !(radius >= 0.0f))
    return fail(VH_ERR_INVALID_ARG, "bad argument");
  *n_blocks = 0;
  std::lock_guard<std::mutex> lk(v->mu);
  int rc = set_device(v);
  if (rc || (rc = sync_and_check(v))) return rc;
  FrameCtl c{};
  std::vector<unsigned long long> bk;
  if ((rc = resident_blocks(v, c, bk))) return rc;
  const double cc[3] = {center[0], center[1], center[2]}, r2 = (double)radius * (double)radius;
  std::vector<uint32_t> out;
  for (uint32_t b = 0; b < (uint32_t)bk.size(); ++b)
    if (block_dist2(bk[b], v->cfg.vox_size, cc) > r2) out.push_back(b);
  if (out.empty()) return VH_OK;
  // a key that left before and was observed again: its stored copy merges into the resident
  // block first, so the block leaves as one copy (the pool and its order do not change)
  if ((rc = merge_resident_duplicates(v, bk, c.frame_id, c.pool_count))) return rc;
  *n_blocks = (int64_t)out.size();
  return stream_out_blocks(v, bk, out, c.frame_id);
}

int vh_stream_in(vh_volume* v, const float center[3], float radius, int64_t* n_blocks) {
  if (!v || !center || !n_blocks || !finite3(center) || !(radius >= 0.0f))
    return fail(VH_ERR_INVALID_ARG, "bad argument");
  *n_blocks = 0;
  std::lock_guard<std::mutex> lk(v->mu);
  int rc = set_device(v);
  if (rc || (rc = sync_and_check(v))) return rc;
  const double cc[3] = {center[0], center[1], center[2]}, r2 = (double)radius * (double)radius;
  std::vector<uint32_t> slots;
  for (const auto& kv : v->store_index)
    if (block_dist2(kv.first, v->cfg.vox_size, cc) <= r2) slots.push_back(kv.second);
  std::sort(slots.begin(), slots.end());
  if (slots.empty()) return VH_OK;
  FrameCtl c{};
  std::vector<unsigned long long> bk;
  if ((rc = resident_blocks(v, c, bk))) return rc;
  return stream_in_blocks(v, slots, c.frame_id, c.pool_count, n_blocks, count_new_keys(v, slots, bk));
}

int vh_get_config(vh_volume* v, vh_config* out) {
  if (!v || !out) return fail(VH_ERR_INVALID_ARG, "null argument");
  *out = v->cfg;
  return VH_OK;
}

int vh_save_volume(vh_volume* v, const char* path) {
  if (!v || !path) return fail(VH_ERR_INVALID_ARG, "null argument");
  std::lock_guard<std::mutex> lk(v->mu);
  int rc = set_device(v);
  if (rc || (rc = sync_and_check(v))) return rc;
  FrameCtl c{};
  std::vector<unsigned long long> bk;
  if ((rc = resident_blocks(v, c, bk))) return rc;
  // each key once: stored copies of resident keys merge into the resident blocks first
  if ((rc = merge_resident_duplicates(v, bk, c.frame_id, c.pool_count))) return rc;
  const uint32_t wpb = words_per_block(v);
  std::vector<uint32_t> host;  // occupied store slots, ascending
  for (const auto& kv : v->store_index) host.push_back(kv.second);
  std::sort(host.begin(), host.end());
  const int64_t n_res = (int64_t)bk.size(), n_host = (int64_t)host.size();
  CkptHeader h{};
  std::memcpy(h.magic, kCkptMagic, 8);
  h.abi = VH_ABI_VERSION;
  h.words_per_block = (int32_t)wpb;
  h.cfg = v->cfg;
  h.frame_id = c.frame_id;
  h.frames = c.frames;
  h.n_blocks = n_res + n_host;
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(VH_ERR_IO, std::string("cannot open ") + path);
  bool ok = std::fwrite(&h, sizeof h, 1, f) == 1;
  ok = ok && std::fwrite(bk.data(), 8, (size_t)n_res, f) == (size_t)n_res;
  for (size_t i = 0; ok && i < host.size(); ++i) ok = std::fwrite(&v->slot_key[host[i]], 8, 1, f) == 1;
  // voxel words (then colours): resident blocks in rounds of 65,536, then the store's
  std::vector<uint32_t> buf;
  for (int pass = 0; pass < (v->colors ? 2 : 1) && ok; ++pass) {
    const size_t per = pass == 0 ? wpb : (size_t)kB3;
    const uint32_t* src = pass == 0 ? static_cast<const uint32_t*>(v->voxels) : v->colors;
    for (int64_t b0 = 0; b0 < n_res && ok; b0 += 65536) {
      const size_t n = (size_t)std::min<int64_t>(65536, n_res - b0);
      buf.resize(n * per);
      if (hipMemcpy(buf.data(), src + (size_t)b0 * per, n * per * 4, hipMemcpyDeviceToHost) != hipSuccess) {
        std::fclose(f);
        return fail(VH_ERR_DEVICE, "device read failed");
      }
      ok = std::fwrite(buf.data(), 4, n * per, f) == n * per;
    }
    for (size_t i = 0; ok && i < host.size(); ++i)
      ok = std::fwrite(pass == 0 ? store_words(v, host[i]) : store_cols(v, host[i]), 4, per, f) == per;
  }
  ok = (std::fclose(f) == 0) && ok;
  return ok ? VH_OK : fail(VH_ERR_IO, "write failed");
}

int vh_load_volume(const char* path, int32_t device, vh_volume** out) {
  if (!path || !out) return fail(VH_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(VH_ERR_IO, std::string("cannot open ") + path);
  CkptHeader h{};
  if (std::fread(&h, sizeof h, 1, f) != 1 || std::memcmp(h.magic, kCkptMagic, 8) != 0) {
    std::fclose(f);
    return fail(VH_ERR_IO, "not a volume checkpoint");
  }
  if (h.abi != VH_ABI_VERSION || h.n_blocks < 0 || h.n_blocks >= ((int64_t)1 << 31)) {
    std::fclose(f);
    return fail(VH_ERR_IO, "checkpoint of another library version");
  }
  vh_config cfg = h.cfg;
  if (device >= 0) cfg.device = device;
  vh_volume* v = nullptr;
  int rc = vh_create(&cfg, &v);
  if (rc) {
    std::fclose(f);
    return rc;
  }
  auto bail = [&](int code, const char* msg) {
    std::fclose(f);
    std::string m = msg ? msg : vh_last_error();
    vh_destroy(v);
    return fail(code, m);
  };
  const uint32_t wpb = words_per_block(v);
  if ((uint32_t)h.words_per_block != wpb) return bail(VH_ERR_IO, "checkpoint voxel format does not match its configuration");
  // the blocks go to the host store (slots 0 .. n-1, slab by slab), then into the pool
  const size_t n = (size_t)h.n_blocks;
  std::vector<unsigned long long> keys(n);
  if (std::fread(keys.data(), 8, n, f) != n) return bail(VH_ERR_IO, "checkpoint truncated");
  for (size_t i = 0; i < n; ++i) {
    uint32_t slot = 0;
    if ((rc = store_alloc_slot(v, &slot))) return bail(rc, nullptr);
    v->slot_key[slot] = keys[i];
    if (!v->store_index.emplace(keys[i], slot).second) return bail(VH_ERR_IO, "checkpoint holds a block twice");
  }
  for (int pass = 0; pass < (v->colors ? 2 : 1); ++pass) {
    const size_t per = pass == 0 ? wpb : (size_t)kB3;
    for (size_t s0 = 0; s0 < n; s0 += kSlabBlocks) {
      const size_t m = std::min<size_t>(kSlabBlocks, n - s0);
      uint32_t* dst = pass == 0 ? v->slab_w[s0 >> kSlabShift] : v->slab_c[s0 >> kSlabShift];
      if (std::fread(dst, 4, m * per, f) != m * per) return bail(VH_ERR_IO, "checkpoint truncated");
    }
  }
  std::fclose(f);
  // the frame counters continue from the checkpoint; every block goes to the pool if it fits
  FrameCtl c{};
  if ((rc = read_ctl(v, &c)) == VH_OK) {
    c.frame_id = h.frame_id ? h.frame_id : 1u;
    c.frames = h.frames;
    if (hipMemcpy(v->ctl, &c, sizeof c, hipMemcpyHostToDevice) != hipSuccess) rc = fail(VH_ERR_DEVICE, "device write failed");
  }
  if (rc == VH_OK && n <= (size_t)cfg.pool_blocks) {
    std::vector<uint32_t> slots(n);
    for (size_t i = 0; i < n; ++i) slots[i] = (uint32_t)i;
    int64_t moved = 0;
    std::lock_guard<std::mutex> lk(v->mu);
    rc = stream_in_blocks(v, slots, c.frame_id, 0u, &moved, (uint32_t)n);
  }
  if (rc) {
    std::string msg = vh_last_error();
    vh_destroy(v);
    return fail(rc, msg);
  }
  *out = v;
  return VH_OK;
}

}  // extern "C"
