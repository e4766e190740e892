// vh_internal.h - device data layout and kernel launchers of libvh (not part of the ABI).
//
// HBM layout of one volume:
//   keys[4 * 2^log2_buckets]        u64 packed block key per slot, ~0 = empty;
//                                   a bucket is 4 consecutive slots (32 B), probing walks
//                                   consecutive buckets;
//   vals[4 * 2^log2_buckets]        u64 per slot: pool index (low 32) | frame of birth (high 32);
//   marks[2][4 * 2^log2_buckets]    u32 per slot and batch parity: bit j set iff frame j of the
//                                   batch touched the block (cleared when the batch is integrated);
//   block_keys[pool_blocks]         u64 key of each pool block (pool order);
//   block_info[pool_blocks]         u64 per pool block: its hash slot (low 32) | birth frame (high 32);
//   voxels[pool_blocks * 512]       FP32: float2 {tsdf, weight}; FP16: u32 {f16 tsdf | u16 w << 16};
//   colors[pool_blocks * 512]       (colour volumes only) u32 {r | g << 8 | b << 16};
//   touched[2][kStripes * stripe_cap] u32 slot ids of the blocks crossed by a batch's bands
//                                   (each once), in kStripes sub-lists with separate
//                                   counters, one list per batch parity;
//   dm[2][8][W*H]                   per batch parity and frame of the batch, each pixel's depth / mu
//                                   (-inf where not a valid measurement), gathered by the integration;
//   tilemax[2][8][tiles]            carving volumes: per batch parity and frame of the batch, the
//                                   largest valid depth of each 16 x 16 pixel tile;
//   ctl                             counters of the frame pipeline (FrameCtl).
//
// Device-scope atomics on one address serialise (about 11 ns each on MI355X), so every
// counter that many workgroups update is striped: kStripes copies, one per 128-byte line.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vh.h"
#include "vh_math.h"

namespace vh {

// 1/z as the hardware reciprocal refined by one Newton step (3 instructions instead of the
// 11 of the IEEE division sequence). For every normal z whose reciprocal is normal
// (z < 2^126) this equals the correctly rounded quotient 1.0f / z bit for bit; the check
// over all such z is tests/device_checks/recip_exact.hip. Callers bound z to [2^-64, 2^64].
__device__ __forceinline__ float recip_rn(float z) {
  const float r0 = __builtin_amdgcn_rcpf(z);
  return __builtin_fmaf(__builtin_fmaf(-z, r0, 1.0f), r0, r0);
}

// a / b correctly rounded, for 1 <= b <= 2^25: y = RN(1/b) (recip_rn), q0 = RN(a*y), then
// one correction q0 + (a - b*q0)*y with the residual exact under FMA (Markstein's theorem:
// with y within half an ulp of 1/b and q0 within one ulp of a/b the corrected quotient is
// RN(a/b) unless the quotient is subnormal). 6 instructions instead of 11; for 0 < |a| <
// 2^-76 (where a/b could be subnormal) the IEEE division is taken instead. The check over
// all a for b = 1..256 is tests/device_checks/recip_exact.hip.
// The fallback branch is taken by the wave when any active lane needs it (a uniform branch
// on the ballot of the active lanes), so the common path carries no per-lane branch.
__device__ __forceinline__ float div_rn(float a, float b) {
  const float y = recip_rn(b);
  const float q0 = a * y;
  float q = __builtin_fmaf(__builtin_fmaf(-b, q0, a), y, q0);
  const bool tiny = __builtin_fabsf(a) < 0x1p-76f && a != 0.0f;
  if (__builtin_expect(__builtin_amdgcn_ballot_w64(tiny) != 0, 0)) q = tiny ? a / b : q;
  return q;
}

// div_rn for dividends that are exact non-negative integers (0 or >= 1; the colour update's
// c * w + rgb): the quotient by b <= 2^25 is then 0 or at least 2^-25, never subnormal, so
// the correction alone is exact and there is no fallback test.
__device__ __forceinline__ float div_rn_int(float a, float b) {
  const float y = recip_rn(b);
  const float q0 = a * y;
  return __builtin_fmaf(__builtin_fmaf(-b, q0, a), y, q0);
}

// div_rn for dividends that are rarely exactly +0 and never -0 (the TSDF mean's s*w + dist:
// dist is never -0, and x + y is -0 only when both are): a = +0 takes the fallback, whose
// quotient +0 is the one the correction gives, so the zero test can go.
__device__ __forceinline__ float div_rn_rare0(float a, float b) {
  const float y = recip_rn(b);
  const float q0 = a * y;
  float q = __builtin_fmaf(__builtin_fmaf(-b, q0, a), y, q0);
  const bool tiny = __builtin_fabsf(a) < 0x1p-76f;
  if (__builtin_expect(__builtin_amdgcn_ballot_w64(tiny) != 0, 0)) q = tiny ? a / b : q;
  return q;
}

enum : uint32_t { kFlagPoolFull = 1u, kFlagHashFull = 2u, kFlagRange = 4u, kFlagTouchedFull = 8u };
constexpr int kMaxProbe = 64;  // buckets examined before a key is declared unplaceable
constexpr int kStripes = 64;   // striped counters / touched sub-lists
constexpr int kLine = 32;      // u32 per 128-byte line
constexpr int kMaxBatch = 8;   // frames integrated per launch at most
constexpr int kLineCarve = 1 + kMaxBatch;  // word of a stripe line counting the mark bits the carve pass added
constexpr int kTile = 16;      // pixels per edge of a max-depth tile (carving)

struct FrameCtl {
  uint32_t frame_id;       // id of the first frame of the batch being integrated (starts at 1)
  uint32_t pool_count;     // keys inserted (pool indices handed out, may pass capacity)
  uint32_t flags;          // kFlag* raised since the host last read them
  uint32_t touched_last;
  uint32_t cursor[3];      // next frame of the staged input (0), the bound sequence (1) and the
                           // ring of asynchronously copied host frames (2)
  uint32_t done_top;       // stripes whose workgroups all finished this frame launch
  unsigned long long touched_total;
  unsigned long long frames;
  unsigned long long new_total;
  unsigned long long integrated_total;  // sum over batches of the blocks listed (read + written once)
  uint32_t batch_id;       // batches integrated; its parity selects the lists and marks
  uint32_t pad0;
  unsigned long long carved_total;  // block-frame updates added by carve passes (mark bits the bands did not set)
  uint32_t pad1[12];
  // per batch parity, one line per stripe: word 0 = listed blocks, word 1 + j = blocks frame
  // j of the batch touched first in the stripe, word kLineCarve = mark bits the carve pass added
  uint32_t tcount[2 * kStripes * kLine];
  uint32_t done[kStripes * kLine];        // workgroups of the frame launch finished, per stripe
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
  BandScale band;  // reciprocals for the band setup
  float max_depth, min_depth, vox, bsz, mu, inv_mu;
  float cap;  // weight cap, +inf when uncapped
  // tables
  HashView hash;
  uint32_t* marks;         // 2 x n_slots
  uint64_t n_slots;
  unsigned long long* block_keys;
  unsigned long long* block_info;  // per pool block: hash slot | birth frame << 32
  uint32_t pool_cap;
  uint32_t* touched;
  uint32_t stripe_cap;
  void* voxels;
  FrameCtl* ctl;
  // colour (both null when the volume has no colour): input frames of W*H*3 bytes (RGB,
  // frame f at rgb + f*W*H*3) and the colour pool, one u32 {r | g << 8 | b << 16} per voxel
  const uint8_t* rgb;
  uint32_t* colors;
  // colour volumes: the allocation pass packs each pixel of frame j of a batch into one word
  // {r | g << 8 | b << 16} at rgba + (parity * kMaxBatch + j) * W*H, which the integration
  // of the batch gathers
  uint32_t* rgba;
  // the allocation pass writes each pixel of frame j of a batch as scaled_depth (depth / mu,
  // -inf where not valid) to dm + (parity * kMaxBatch + j) * W*H; the integration gathers it
  float* dm;
  // carving volumes (null otherwise): the allocation pass writes the largest valid depth of each
  // kTile x kTile pixel tile of frame j of a batch to tilemax + (parity * kMaxBatch + j) *
  // tiles_x * tiles_y, which the carve pass reads
  float* tilemax;
  int32_t tiles_x, tiles_y;
};

struct McArgs {
  HashView hash;
  const unsigned long long* block_keys;
  const void* voxels;
  uint32_t pool_cap;
  uint32_t n_blocks;   // pool blocks to extract
  float vox;
  uint32_t* nbrs;      // per pool block: the 8 neighbourhood pool indices (k_mc_neighbours)
  uint32_t* counts;    // per pool block: triangles (count pass), then exclusive offsets
  uint32_t* chunk_sums;// per 1024 pool blocks
  uint32_t* total;     // total triangles
  float* tris;         // capacity * 9, may be null
  int32_t* cubes;      // capacity * 4, may be null
  uint32_t capacity;
  const uint32_t* colors;  // colour pool (null: no colour)
  uint8_t* rgb;            // capacity * 9 (r,g,b per vertex), may be null
  // indexed mesh (launch_mesh_index_*): lattice edges are owned by the block holding their
  // lower end point; bit axis * 512 + in-block index of a block's 48 words marks a used edge
  uint32_t* ebits;     // per pool block: 48 words of used-edge bits
  uint16_t* epre;      // per pool block: 48 exclusive popcount prefixes of ebits
  uint32_t* vcounts;   // per pool block: owned vertices, then exclusive offsets
  uint32_t* vchunks;   // per kScanChunk pool blocks
  uint32_t* vtotal;    // total vertices
  int32_t* faces;      // capacity * 3 vertex indices
  float* vertices;     // vertex_cap * 3 (metres)
  uint8_t* vrgb;       // vertex_cap * 3, may be null
  uint32_t vertex_cap;
};

// Read-only lookup (no concurrent inserts): pool index of `key`, or kNoBlock.
// Slots of a bucket fill in order, so the first empty slot ends the search.
__device__ __forceinline__ uint32_t find_block(const HashView& h, uint64_t key) {
  uint64_t b = bucket_of(key, h.log2_buckets);
  for (int p = 0; p < kMaxProbe; ++p) {
    const ulonglong2* bp = reinterpret_cast<const ulonglong2*>(h.keys + b * kSlotsPerBucket);
    const ulonglong2 k01 = bp[0], k23 = bp[1];
    const uint64_t ks[4] = {k01.x, k01.y, k23.x, k23.y};
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      if (ks[s] == key) return (uint32_t)h.vals[b * kSlotsPerBucket + s];
      if (ks[s] == kEmptyKey) return kNoBlock;
    }
    b = (b + 1) & h.bucket_mask;
  }
  return kNoBlock;
}

// Rounds to binary16 from a float that has itself been rounded: without the empty register
// constraint the compiler folds a preceding fused multiply-add and this conversion into one
// mixed-precision instruction (v_fma_mixlo_f16) that rounds once, straight to binary16, which
// differs from rounding to FP32 first (double rounding) in about one value in 10^4.
__device__ __forceinline__ uint16_t f32_to_f16_rn(float s) {
  __asm__ volatile("" : "+v"(s));
  return __builtin_bit_cast(uint16_t, (_Float16)s);
}

// Storage formats: unpack a voxel word to (tsdf, weight) in FP32, pack it back, the empty word.
template <int kStorage>
struct VoxelIO;

template <>
struct VoxelIO<VH_STORAGE_F32> {
  using Word = float2;
  static __device__ __forceinline__ void unpack(Word w, float& s, float& wt) { s = w.x; wt = w.y; }
  static __device__ __forceinline__ Word pack(float s, float wt) { return make_float2(s, wt); }
  static __device__ __forceinline__ Word zero() { return make_float2(0.0f, 0.0f); }
};
template <>
struct VoxelIO<VH_STORAGE_F16> {
  using Word = uint32_t;
  static __device__ __forceinline__ void unpack(Word w, float& s, float& wt) {
    s = (float)__builtin_bit_cast(_Float16, (uint16_t)(w & 0xffffu));
    wt = (float)(w >> 16);
  }
  static __device__ __forceinline__ Word pack(float s, float wt) {
    const uint16_t h = f32_to_f16_rn(s);
    return (uint32_t)h | ((uint32_t)wt << 16);
  }
  static __device__ __forceinline__ Word zero() { return 0u; }
};

template <>
struct VoxelIO<VH_STORAGE_F16R> {
  using Word = uint32_t;
  // value = h + r * ulp(h)/256: one fused multiply-add, exact (11 + 8 significant bits)
  static __device__ __forceinline__ void unpack(Word w, float& s, float& wt) {
    const float hf = (float)__builtin_bit_cast(_Float16, (uint16_t)(w & 0xffffu));
    s = __builtin_fmaf((float)(int)(int8_t)(w >> 24), f16r_scale_of(hf), hf);
    wt = (float)((w >> 16) & 0xffu);
  }
  static __device__ __forceinline__ Word pack(float s, float wt) {
    const uint16_t h = f32_to_f16_rn(s);
    const float hf = (float)__builtin_bit_cast(_Float16, h);
    const float r = fminf(127.0f, fmaxf(-128.0f, rintf((s - hf) * f16r_inv_scale_of(hf))));
    return (uint32_t)h | ((uint32_t)wt << 16) | ((uint32_t)(uint8_t)(int8_t)(int)r << 24);
  }
  static __device__ __forceinline__ Word zero() { return 0u; }
};

// Out-of-core streaming (launch_stream_*): blocks move between the pool and the host store,
// pinned host slabs of kSlabBlocks blocks that kernels read and write directly (wpb 32-bit
// words of voxels per block, plus 512 colour words); slot s lives in slab s >> kSlabShift.
constexpr uint32_t kSlabShift = 14, kSlabBlocks = 1u << kSlabShift;
struct StreamArgs {
  HashView hash;
  unsigned long long* block_keys;
  unsigned long long* block_info;  // per pool block: hash slot | birth frame << 32
  uint32_t* pool;              // voxel pool as 32-bit words (wpb per block)
  uint32_t* colors;            // colour pool (null: no colour)
  uint32_t wpb;                // 32-bit words per block: 1024 (FP32) or 512 (16-bit formats)
  uint32_t n;                  // blocks of this launch
  const uint32_t* src;         // out / move: pool blocks
  const uint32_t* dst;         // move: their new pool blocks
  const uint32_t* slots;       // out / insert: the blocks' host-store slots
  uint32_t* const* slab_words; // device addresses of the host slabs' voxel words
  uint32_t* const* slab_cols;  // and colour words (colour volumes)
  const unsigned long long* keys;  // insert: the blocks' keys
  uint32_t* placed;            // insert: pool block per block (kMergeBit: merged into a
                               // resident block), kNoBlock where the hash is full
  FrameCtl* ctl;
  uint32_t pool_cap;
  uint32_t birth;              // birth frame written with rebuilt and inserted keys
  float cap;                   // weight cap (0: none)
};
constexpr uint32_t kMergeBit = 0x80000000u;

constexpr int kScanChunk = 1024;
constexpr int kEdgeWords = 3 * kB3 / 32;  // used-edge words per block (indexed mesh)

struct QueryArgs {
  HashView hash;
  const void* voxels;
  uint32_t pool_cap;
  float vox;
};

// Gauss-Newton accumulators: 21 (J^T W J upper) + 6 (J^T W r) + 1 (r^2) + 1 (valid samples) +
// 1 (points considered: points with a non-finite coordinate are skipped) + 1 (robust cost).
constexpr int kGnAcc = 31;
// Residual model of the pose solver. A point measured at depth z has the axial noise of
// structured-light depth sensors, sigma(z) = 0.0012 + 0.0019 (z - 0.4)^2 metres (Nguyen, Izadi
// and Lovell 2012; the synthetic frames are drawn from it), so its normalised residual r (TSDF
// units, mu metres per unit) is scaled to e = r mu / sigma(z), and the cost is the Huber loss
// rho(e) = e^2/2 for |e| <= k, k (|e| - k/2) beyond, with k = kHuber.
constexpr double kHuber = 2.0;

// Levenberg-Marquardt state kept on the device between iterations.
struct GnState {
  double pose[12];  // accepted pose (row-major 3x4)
  double cand[12];  // the pose the next evaluation is made at
  double Ab[27];    // J^T J (upper, 21) and J^T r (6) at the accepted pose
  double cost;      // its cost
  double lambda;
  int has;          // an accepted pose exists
  int pad;
  // tracking only (zero otherwise): the motion prior 1/2 xi^T diag(lam_t I3, lam_r I3) xi with
  // xi = log(start^-1 T), and the depth term scaled by n_eff / (points considered)
  double start[12];
  double lam_t, lam_r, n_eff;
};
// per-evaluation stats: n_valid, rms, |delta|, lambda, accepted, accepted pose (12)
constexpr int kGnStats = 17;

// Starting pose (row-major 3x4) and damping of a refinement (k_gn_start).
struct GnStart {
  double pose[12];
  double lambda;
};

struct GnArgs {
  QueryArgs q;
  float mu;           // truncation (metres): residual scale of the noise model
  const float* pts;   // camera-frame points (a non-finite coordinate marks a point to skip)
  int64_t n;
  const double* pose; // without a previous evaluation: the pose to evaluate (12 doubles, 3x4)
  double* partials;   // gridDim.x * kGnAcc
  // the previous evaluation, stepped from at the start of the launch (prev_grid = 0: none)
  const double* prev_partials;
  int prev_grid;
  const GnState* st_in;  // solver state before that step
  GnState* st_out;       // after it (written by workgroup 0)
  double* cand_out;      // the new candidate (workgroup 0), or null
  double* stats_row;     // the previous evaluation's stats row (workgroup 0)
};

// Tracking over a bound sequence: frame f (the sequence cursor) starts Gauss-Newton from
// prior[f] (or, with `increment` and f > 0, from out[f-1] * prior[f]), back-projects its
// depth on a stride grid, and writes the accepted pose to out[f], which the frame's
// allocation and integration then use.
struct TrackArgs {
  const void* depth;      // n_seq frames of W*H
  int32_t depth_u16;
  float depth_scale;
  const float* prior;     // n_seq x 16 row-major camera-to-world (or motion from frame f-1)
  int32_t increment;      // prior[f] is a motion from the refined pose of frame f-1
  float* out;             // n_seq x 16
  double* stats;          // n_seq x (iters + 1) x kGnStats, or null
  const double* eval_stats; // the solver's rows for the current frame ((iters + 1) x kGnStats)
  uint32_t n_seq;
  FrameCtl* ctl;
  int32_t W, H, stride, pw, ph;  // image, stride, point grid (ceil(W/stride) x ceil(H/stride))
  Intr K;
  float max_depth;        // points need 0 < d <= max_depth
  float prior_sigma_t, prior_sigma_r, n_eff;  // motion prior and depth-term scale (vh_track_config)
  float* pts;             // pw * ph * 3
  double* pose;           // Gauss-Newton candidate (12)
  GnState* state;
  int32_t iters;
};

// launchers (vh_kernels.hip); all asynchronous on `st`
// Allocation of the next k frames (one batch, 1 <= k <= kMaxBatch), then their integration.
hipError_t launch_alloc(const FrameArgs& a, int depth_u16, int k, hipStream_t st);
hipError_t launch_integrate(const FrameArgs& a, int storage, int grid, int k, hipStream_t st);
// Carving: lists every resident block that frame j of the allocated batch (k frames) may update
// (conservative visibility and depth test per pool block) and sets its mark bit j.
hipError_t launch_carve(const FrameArgs& a, int k, hipStream_t st);
// Integration of the allocated batch (k_int frames) and allocation of the next one (k_alloc
// frames) in one launch (bound sequences; the next batch's allocation must not have run).
hipError_t launch_integrate_alloc(const FrameArgs& a, int storage, int depth_u16, int grid,
                                  int k_int, int k_alloc, hipStream_t st);
// count pass, two-level exclusive scan, write pass (all on `st`)
hipError_t launch_mc(const McArgs& a, int storage, hipStream_t st);
// Indexed mesh, first half: neighbour table, used-edge bits and face counts, vertex counts,
// and the exclusive scans of both (face total in *a.total, vertex total in *a.vtotal).
hipError_t launch_mesh_index_count(const McArgs& a, int storage, hipStream_t st);
// Second half (after the count on the same arguments): faces and vertices.
hipError_t launch_mesh_index_write(const McArgs& a, int storage, hipStream_t st);
// Streaming: host slot slots[i] = pool block src[i] (and colours), written over the link.
hipError_t launch_stream_out(const StreamArgs& a, hipStream_t st);
// pool[dst[i]] = pool[src[i]] with its key and colours (source and destination sets disjoint).
hipError_t launch_stream_move(const StreamArgs& a, hipStream_t st);
// Inserts block_keys[0 .. n) into an empty hash table, slot value = index | birth << 32.
hipError_t launch_stream_rehash(const StreamArgs& a, hipStream_t st);
// Inserts the host-store blocks at slots[]: new keys take pool blocks from ctl->pool_count
// and their words; keys already resident are merged into the resident block (placed[] says
// which).
hipError_t launch_stream_insert(const StreamArgs& a, int storage, hipStream_t st);
hipError_t launch_query(const QueryArgs& q, int storage, const float* pts, int64_t n, float* sdf,
                        uint8_t* valid, hipStream_t st);
hipError_t launch_backward(const QueryArgs& q, int storage, const float* pts, int64_t n,
                           const float* T12, const float* grad_r, float* r, uint8_t* valid,
                           float* J, double* partials, int grid, float* voxel_grad,
                           hipStream_t st);
hipError_t launch_reduce6(const double* partials, int grid, double* out6, hipStream_t st);
hipError_t launch_gn_accum(const GnArgs& a, int storage, int grid, hipStream_t st);
hipError_t launch_gn_solve(const double* partials, int grid, const GnState* st_in, GnState* st_out, double* cand,
                           int final_step, double* stats_row, hipStream_t st);
// cand = s.pose, solver state cleared with damping s.lambda.
hipError_t launch_gn_start(const GnStart& s, GnState* state, double* cand, hipStream_t st);
// Gauss-Newton state reset + prior pose of the current frame, and the depth back-projection,
// in one launch.
hipError_t launch_track_begin(const TrackArgs& t, hipStream_t st);
// Accepted pose (and the frame's solver stats) to the per-frame outputs.
hipError_t launch_track_end(const TrackArgs& t, hipStream_t st);

}  // namespace vh
