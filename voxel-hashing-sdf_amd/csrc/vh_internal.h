// vh_internal.h - device data layout and kernel launchers of libvh (not part of the ABI).
//
// HBM layout of one volume:
//   keys[4 * 2^log2_buckets]        u64 packed block key per slot, ~0 = empty;
//                                   a bucket is 4 consecutive slots (32 B), probing walks
//                                   consecutive buckets;
//   vals[4 * 2^log2_buckets]        u64 per slot: pool index (low 32) | frame of birth (high 32);
//   stamps[4 * 2^log2_buckets]      u32 per slot: last frame that touched the block;
//   block_keys[pool_blocks]         u64 key of each pool block (pool order);
//   voxels[pool_blocks * 512]       FP32: float2 {tsdf, weight}; FP16: u32 {f16 tsdf | u16 w << 16};
//   touched[kStripes * stripe_cap]  u32 slot ids of the blocks crossed by this frame's bands,
//                                   in kStripes sub-lists with separate counters;
//   ctl                             counters of the frame pipeline (FrameCtl).
//
// Device-scope atomics on one address serialise (about 11 ns each on MI355X), so every
// counter that many workgroups update is striped: kStripes copies, one per 128-byte line.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vh_math.h"

namespace vh {

enum : uint32_t { kFlagPoolFull = 1u, kFlagHashFull = 2u, kFlagRange = 4u, kFlagTouchedFull = 8u };
constexpr int kMaxProbe = 64;  // buckets examined before a key is declared unplaceable
constexpr int kStripes = 64;   // striped counters / touched sub-lists
constexpr int kLine = 32;      // u32 per 128-byte line

struct FrameCtl {
  uint32_t frame_id;       // stamp of the frame being integrated (starts at 1)
  uint32_t pool_count;     // keys inserted (pool indices handed out, may pass capacity)
  uint32_t flags;          // kFlag* raised since the host last read them
  uint32_t touched_last;
  uint32_t cursor[2];      // next frame of the staged input (0) and bound sequence (1)
  uint32_t done_top;       // integration stripes finished this frame
  uint32_t pool_at_frame_start;
  unsigned long long touched_total;
  unsigned long long frames;
  unsigned long long new_total;
  uint32_t pad0[18];
  uint32_t tcount[kStripes * kLine];  // touched sub-list lengths (one line each)
  uint32_t done[kStripes * kLine];    // integration workgroups finished, per stripe
};

struct HashView {
  unsigned long long* keys;
  unsigned long long* vals;
  int log2_buckets;
  uint64_t bucket_mask;
};

struct FrameArgs {
  // input frames: frame f of the input is depth + f*W*H elements, pose poses + 16 f
  const void* depth;
  const float* poses;
  uint32_t n_seq;
  int32_t cursor_id;
  float depth_scale;
  // camera and volume parameters
  int32_t W, H;
  Intr K;
  float max_depth, min_depth, vox, bsz, mu, inv_mu, cap;
  // tables
  HashView hash;
  uint32_t* stamps;
  unsigned long long* block_keys;
  uint32_t pool_cap;
  uint32_t* touched;
  uint32_t stripe_cap;
  void* voxels;
  FrameCtl* ctl;
};

struct McArgs {
  HashView hash;
  const unsigned long long* block_keys;
  const void* voxels;
  uint32_t pool_cap;
  uint32_t n_blocks;   // pool blocks to extract
  float vox;
  uint32_t* counts;    // per pool block: triangles (count pass), then exclusive offsets
  uint32_t* chunk_sums;// per 1024 pool blocks
  uint32_t* total;     // total triangles
  float* tris;         // capacity * 9
  int32_t* cubes;      // capacity * 4, may be null
  uint32_t capacity;
};
constexpr int kScanChunk = 1024;

struct QueryArgs {
  HashView hash;
  const void* voxels;
  uint32_t pool_cap;
  float vox;
};

// Gauss-Newton / backward accumulators: 21 (J^T J upper) + 6 (J^T r) + 1 (r^2) + 1 (count).
constexpr int kGnAcc = 29;

// Levenberg-Marquardt state kept on the device between iterations.
struct GnState {
  double pose[12];  // accepted pose (row-major 3x4)
  double Ab[27];    // J^T J (upper, 21) and J^T r (6) at the accepted pose
  double cost;      // its cost
  double lambda;
  int has;          // an accepted pose exists
  int pad;
};
constexpr int kGnStats = 5;  // per-iteration stats: n_valid, rms, |delta|, lambda, accepted

struct GnArgs {
  QueryArgs q;
  const float* pts;   // camera-frame points
  int64_t n;
  const double* pose; // 12 doubles, row-major 3x4 camera-to-world
  double* partials;   // gridDim.x * kGnAcc
};

// launchers (vh_kernels.hip); all asynchronous on `st`
hipError_t launch_alloc(const FrameArgs& a, int depth_u16, hipStream_t st);
hipError_t launch_integrate(const FrameArgs& a, int storage, int depth_u16, int grid,
                            hipStream_t st);
// count pass, two-level exclusive scan, write pass (all on `st`)
hipError_t launch_mc(const McArgs& a, int storage, hipStream_t st);
hipError_t launch_query(const QueryArgs& q, int storage, const float* pts, int64_t n, float* sdf,
                        uint8_t* valid, hipStream_t st);
hipError_t launch_backward(const QueryArgs& q, int storage, const float* pts, int64_t n,
                           const float* T12, const float* grad_r, float* r, uint8_t* valid,
                           float* J, double* partials, int grid, float* voxel_grad,
                           hipStream_t st);
hipError_t launch_reduce6(const double* partials, int grid, double* out6, hipStream_t st);
hipError_t launch_gn_accum(const GnArgs& a, int storage, int grid, hipStream_t st);
hipError_t launch_gn_solve(const double* partials, int grid, GnState* state, double* cand,
                           double n_points, int final_step, double* stats_row, hipStream_t st);

}  // namespace vh
