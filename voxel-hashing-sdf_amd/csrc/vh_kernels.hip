// vh_kernels.hip - gfx950 kernels of libvh.
//
// Per batch of k consecutive frames (one captured graph per launch): k_alloc (one lane per
// pixel of each frame: truncation-band DDA, hash insert, per-frame marks, compaction of the
// batch's block list) -> k_integrate (one wave per listed block, grid-stride: the block's
// voxels are read once, updated by each frame that touched it, written once; the last
// workgroup to finish closes the batch). Bound sequences run k_integrate_alloc: batch b is
// integrated while batch b+1 is allocated.
// Extraction: k_mc (one workgroup per pool block; 9^3 halo in LDS; wave prefix sums).
// Queries: k_query / k_backward / k_gn_accum + k_gn_solve (FP64 reductions).
//
// Kernel code follows the op order of vh_math.h exactly and is compiled with
// -ffp-contract=off, so FP32 results match a CPU twin of the same algorithm bit-for-bit.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "vh.h"
#include "vh_internal.h"
#include "vh_math.h"
#include "vh_mc_table.h"

namespace vh {

namespace {

__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
__device__ __forceinline__ uint32_t read_lane_u32(uint32_t v, int lane) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, lane);
}
__device__ __forceinline__ int first_lane(uint64_t m) { return __ffsll((unsigned long long)m) - 1; }
// A value every lane holds identically, moved to a scalar register.
__device__ __forceinline__ uint32_t uniform(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ unsigned long long uniform64(unsigned long long v) {
  return (unsigned long long)uniform((uint32_t)v) | (unsigned long long)uniform((uint32_t)(v >> 32)) << 32;
}

__device__ __forceinline__ void raise_flag(FrameCtl* ctl, uint32_t f) { atomicOr(&ctl->flags, f); }


// Lock-free find-or-insert. A key is written once into the first empty slot of its probe
// sequence by a compare-and-swap; a stale cached read can only show EMPTY where a key now
// lives, and the CAS then returns the real occupant, so no lane ever spins on another.
// Returns the slot id (or -1 when the probe limit is exhausted); `inserted` is set for the
// lane whose CAS created the key.
__device__ __forceinline__ int64_t find_or_insert(const HashView& h, uint64_t key, bool& inserted) {
  uint64_t b = bucket_of(key, h.log2_buckets);
  for (int p = 0; p < kMaxProbe; ++p) {
    unsigned long long* bk = h.keys + b * kSlotsPerBucket;
    const ulonglong2* bp = reinterpret_cast<const ulonglong2*>(bk);
    const ulonglong2 k01 = bp[0], k23 = bp[1];
    const uint64_t ks[4] = {k01.x, k01.y, k23.x, k23.y};
#pragma unroll
    for (int s = 0; s < 4; ++s)
      if (ks[s] == key) return (int64_t)(b * kSlotsPerBucket + s);
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      if (ks[s] != kEmptyKey) continue;
      const unsigned long long old = atomicCAS(bk + s, (unsigned long long)kEmptyKey, (unsigned long long)key);
      if (old == kEmptyKey) {
        inserted = true;
        return (int64_t)(b * kSlotsPerBucket + s);
      }
      if (old == key) return (int64_t)(b * kSlotsPerBucket + s);
    }
    b = (b + 1) & h.bucket_mask;
  }
  return -1;
}

template <bool kU16>
__device__ __forceinline__ float depth_at(const void* base, float scale, int pix) {
  if (kU16) return (float)reinterpret_cast<const uint16_t*>(base)[(uint32_t)pix] * scale;
  return reinterpret_cast<const float*>(base)[(uint32_t)pix];
}

// ============================================================================ allocation
// Workgroup = a tile of 16 x 64 pixels, 256 lanes, each lane one pixel in each of the tile's
// four 16 x 16 bands. Phase 1: every lane walks its pixels' truncation bands in turn and
// inserts the block keys into an LDS hash set, so each distinct block of the tile is kept
// once. Phase 2: the distinct keys go to the
// global hash (find or insert), take a pool index if new, and mark the block for this frame
// of the batch (an atomic OR of the frame's bit into the slot's mark); the lanes that found
// the mark empty append the slot to the workgroup's stripe of the batch's block list, and
// the lanes that found their frame's bit clear count the block for that frame.
constexpr int kSetSlots = 1024;  // LDS key set (8 KB)
constexpr int kUniqCap = 1024;   // distinct keys kept per tile (8 KB)
// 16-row bands per tile of a batched launch. One workgroup per four bands pays the LDS clear,
// the barriers and the chain of dependent global atomics of phase 2 once for all four, and
// merges the bands' keys: the fused batch launch ran 3.8 % faster than with one band
// (room_medium_5mm), 2, 6 and 8 bands gained less or lost. A single frame's launch keeps one
// band per workgroup, which fills the GPU with 4x as many workgroups.
constexpr int kAllocRows = 4;
__host__ __device__ constexpr uint32_t alloc_tiles(int W, int H, int rows) {
  return (uint32_t)(((W + 15) >> 4) * ((H + 16 * rows - 1) / (16 * rows)));
}

static_assert(kSetSlots == 1 << 10, "set_hash gives 10 bits");
__device__ __forceinline__ uint32_t set_hash(uint64_t key) {
  return (((uint32_t)key ^ (uint32_t)(key >> 32)) * 0x9e3779b1u) >> 22;
}

// The block list, its stripe lengths and the slot marks are double-buffered by batch
// parity, so the allocation of batch b+1 can run while batch b is integrated. Line s of a
// parity's counters: word 0 = listed blocks of stripe s, word 1 + j = blocks that frame j of
// the batch touched first in that stripe.
__device__ __forceinline__ uint32_t* list_count_of(FrameCtl* ctl, uint32_t parity) {
  return ctl->tcount + parity * (kStripes * kLine);
}
__device__ __forceinline__ uint32_t* list_of(const FrameArgs& a, uint32_t parity) {
  return a.touched + (size_t)parity * kStripes * a.stripe_cap;
}
__device__ __forceinline__ uint32_t* marks_of(const FrameArgs& a, uint32_t parity) {
  return a.marks + (size_t)parity * a.n_slots;
}
// Input slot of frame j of a batch whose first frame is input `cur` (consecutive, wrapping).
__device__ __forceinline__ uint32_t input_slot(uint32_t cur, uint32_t j, uint32_t n_seq) {
  uint32_t fs = cur % n_seq + j;
  while (fs >= n_seq) fs -= n_seq;
  return fs;
}

// Global part for the active lanes' keys: find-or-insert, pool index for new keys (one
// atomic per wave), the frame's mark, the list append and the frame's count (one atomic per
// wave each). `frame` is the frame's id (stored as the birth frame of new blocks), jb its
// index in the batch.
__device__ void claim_keys(const FrameArgs& a, uint64_t key, bool act, uint32_t frame, uint32_t jb,
                           uint32_t parity, int lane, uint32_t stripe) {
  FrameCtl* ctl = a.ctl;
  bool inserted = false;
  int64_t slot = -1;
  if (act) {
    slot = find_or_insert(a.hash, key, inserted);
    if (slot < 0) raise_flag(ctl, kFlagHashFull);
  }
  // The pool reservation and the mark depend only on the slot: both are in flight together,
  // and the slot-value stores go out just before the stripe atomics, so the chain is bucket
  // load -> (CAS) -> pool + mark -> stripe.
  const uint64_t im = __ballot(inserted);
  const int leader = im ? first_lane(im) : 0;
  uint32_t pbase = 0;
  if (im && lane == leader) pbase = atomicAdd(&ctl->pool_count, (uint32_t)__popcll(im));
  const uint32_t bit = 1u << jb;
  uint32_t old_mark = bit;
  if (slot >= 0) old_mark = atomicOr(&marks_of(a, parity)[slot], bit);
  if (im) {
    pbase = read_lane_u32(pbase, leader);
    if (inserted) {
      const uint32_t idx = pbase + lanes_below(im);
      uint32_t stored = kNoBlock;
      if (idx < a.pool_cap) {
        a.block_keys[idx] = key;
        a.block_info[idx] = (unsigned long long)slot | ((unsigned long long)frame << 32);
        stored = idx;
      } else {
        raise_flag(ctl, kFlagPoolFull);
      }
      a.hash.vals[slot] = (unsigned long long)stored | ((unsigned long long)frame << 32);
    }
  }
  uint32_t* line = list_count_of(ctl, parity) + stripe * kLine;
  const uint64_t cm = __ballot((old_mark & bit) == 0u);  // first touch of the block by this frame
  if (cm && lane == first_lane(cm)) atomicAdd(&line[1 + jb], (uint32_t)__popcll(cm));
  const bool first = old_mark == 0u;  // first touch by the batch: list the block
  const uint64_t fm = __ballot(first);
  if (fm) {
    const int fl = first_lane(fm);
    uint32_t tbase = 0;
    if (lane == fl) tbase = atomicAdd(&line[0], (uint32_t)__popcll(fm));
    tbase = read_lane_u32(tbase, fl);
    if (first) {
      const uint32_t pos = tbase + lanes_below(fm);
      if (pos < a.stripe_cap) {
        list_of(a, parity)[(size_t)stripe * a.stripe_cap + pos] = (uint32_t)slot;
      } else {
        // the block cannot be listed: its mark goes back to zero, so that a later lane of this
        // batch that marks it tries to list it again, and no mark outlives the batch unlisted
        atomicExch(&marks_of(a, parity)[slot], 0u);
        raise_flag(ctl, kFlagTouchedFull);
      }
    }
  }
}

// Allocation of one tile of frame `frame` (index jb of its batch, input slot fs): 16 pixel
// columns by 16 * kRows rows, 256 lanes, each lane one pixel of every 16-row band.
// The allocation's LDS: the key set, the tile's distinct keys and, for carving volumes, the
// bands' largest depths.
template <int kRows>
struct AllocLds {
  unsigned long long set[kSetSlots];
  unsigned long long uniq[kUniqCap];
  uint32_t nuniq;
  float tmax[4][kRows];
};

// kColor: the volume keeps colour (a.rgba is set), known at compile time so that no branch
// separates the tile's loads.
template <bool kU16, int kRows, bool kColor>
__device__ void alloc_tile(const FrameArgs& a, uint32_t tile, uint32_t frame, uint32_t jb, uint32_t parity,
                           uint32_t fs, AllocLds<kRows>& L) {
  unsigned long long* s_set = L.set;
  unsigned long long* s_uniq = L.uniq;
  uint32_t& s_nuniq = L.nuniq;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int tiles_x = (a.W + 15) >> 4;
  const int u = ((int)(tile % tiles_x) << 4) + ((wave & 1) << 3) + (lane & 7);
  const int v0 = (int)(tile / tiles_x) * (16 * kRows) + ((wave >> 1) << 3) + (lane >> 3);
  const uint32_t stripe = tile & (kStripes - 1);
  FrameCtl* ctl = a.ctl;
  const float* Tp = a.poses + 16 * (size_t)fs;
  float T[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) T[k] = Tp[k];
  const size_t elem = kU16 ? 2 : 4;
  const void* base = reinterpret_cast<const char*>(a.depth) + (size_t)fs * a.W * a.H * elem;

  // the lane's depths (and colours) of all bands: every load is issued before the first use
  // (a pixel outside the image loads a pixel inside it, whose value is not used), so the loads
  // are in flight together
  float dr[kRows];
  uint32_t px3[kRows][3];
  const int uc = min(u, a.W - 1);
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    const int vc = min(v0 + 16 * r, a.H - 1);
    dr[r] = depth_at<kU16>(base, a.depth_scale, vc * a.W + uc);
    if constexpr (kColor) {
      const uint8_t* px = a.rgb + ((size_t)fs * a.W * a.H + (size_t)vc * a.W + uc) * 3;
#pragma unroll
      for (int c = 0; c < 3; ++c) px3[r][c] = px[c];
    }
  }
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    const int v = v0 + 16 * r;
    if (u < a.W && v < a.H) {
      const size_t staged = ((size_t)parity * kMaxBatch + jb) * a.W * a.H + (size_t)v * a.W + u;
      // the pixel's colour, packed for the integration's one-word gathers
      if constexpr (kColor) a.rgba[staged] = px3[r][0] | px3[r][1] << 8 | px3[r][2] << 16;
      // the pre-scaled depth the integration of this frame gathers
      a.dm[staged] = scaled_depth(dr[r], a.max_depth, a.inv_mu);
    } else {
      dr[r] = 0.0f;
    }
  }

  // carving volumes: the largest valid depth of each 16 x 16 band (a max-depth tile)
  float (*s_tmax)[kRows] = L.tmax;
  if (a.tilemax) {
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      float m = depth_valid(dr[r], a.max_depth) ? dr[r] : 0.0f;
#pragma unroll
      for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
      if (lane == 0) s_tmax[wave][r] = m;
    }
  }
  for (int i = tid; i < kSetSlots; i += 256) s_set[i] = kEmptyKey;
  if (tid == 0) s_nuniq = 0;
  __syncthreads();
  if (a.tilemax && tid < kRows) {
    const int ty = (int)(tile / tiles_x) * kRows + tid;
    if (ty < a.tiles_y)
      a.tilemax[((size_t)parity * kMaxBatch + jb) * a.tiles_x * a.tiles_y + (size_t)ty * a.tiles_x + tile % tiles_x] =
          fmaxf(fmaxf(s_tmax[0][tid], s_tmax[1][tid]), fmaxf(s_tmax[2][tid], s_tmax[3][tid]));
  }

#pragma unroll 1
  for (int r = 0; r < kRows; ++r) {
    const int v = v0 + 16 * r;
    BandDDA s;
    bool active = false;
    if (u < a.W && v < a.H && depth_valid(dr[r], a.max_depth)) {
      active = band_setup(T, a.K, a.band, a.mu, a.min_depth, u, v, dr[r], s, [](float x) { return recip_rn(x); });
      if (!active) raise_flag(ctl, kFlagRange);
    }
    // phase 1: distinct keys of the tile into LDS; keys that do not fit are claimed directly
    while (__ballot(active)) {
      bool spill = false;
      uint64_t key = kEmptyKey;
      if (active) {
        key = pack_key(s.ix, s.iy, s.iz);
        uint32_t h = set_hash(key);
        spill = true;
        for (int p = 0; p < 32; ++p) {
          const unsigned long long old = atomicCAS(&s_set[h], (unsigned long long)kEmptyKey, (unsigned long long)key);
          if (old == key) { spill = false; break; }
          if (old == kEmptyKey) {
            const uint32_t n = atomicAdd(&s_nuniq, 1u);
            if (n < kUniqCap) { s_uniq[n] = key; spill = false; }
            break;
          }
          h = (h + 1) & (kSetSlots - 1);
        }
      }
      if (__ballot(spill)) claim_keys(a, key, spill, frame, jb, parity, lane, stripe);
      if (active) active = band_next(s);
    }
  }
  __syncthreads();
  // phase 2: the tile's distinct keys
  const uint32_t n = min(s_nuniq, (uint32_t)kUniqCap);
  for (uint32_t b0 = 0; b0 < n; b0 += 256) {
    const uint32_t i = b0 + tid;
    const bool act = i < n;
    const uint32_t wave_first = b0 + (uint32_t)(wave << 6);
    if (wave_first >= n) break;  // whole wave idle (uniform per wave)
    claim_keys(a, act ? s_uniq[i] : kEmptyKey, act, frame, jb, parity, lane, stripe);
  }
}

// Allocation of the batch about to be integrated: `tiles` workgroups per frame.
template <bool kU16, int kRows, bool kColor>
__global__ __launch_bounds__(256) void k_alloc(FrameArgs a, uint32_t tiles) {
  __shared__ AllocLds<kRows> L;
  const uint32_t jb = blockIdx.x / tiles;
  const FrameCtl* ctl = a.ctl;
  alloc_tile<kU16, kRows, kColor>(a, blockIdx.x - jb * tiles, ctl->frame_id + jb, jb, ctl->batch_id & 1u,
                          input_slot(ctl->cursor[a.cursor_id], jb, a.n_seq), L);
}

// ============================================================================ integration
// tsdf_update (vh_math.h) with RN(1/(w + 1)) by recip_rn (the correctly rounded reciprocal).
// cap_or_inf: the weight cap, +inf when uncapped (min is then the identity). kTab: RN(1/w)
// from the workgroup's table (rcp_table) instead of the reciprocal instructions; the launch
// takes it when every weight it can meet is below kRcpTab - 1 (the values are equal).
constexpr uint32_t kRcpTab = 1024;
template <bool kTab>
__device__ __forceinline__ void mean_update(float& s, float& w, float dist, float cap_or_inf, const float* rcp_tab) {
  const float w_new = w + 1.0f;
  const float r = kTab ? rcp_tab[(uint32_t)w_new] : recip_rn(w_new);
  s = __builtin_fmaf(dist - s, r, s);
  w = fminf(w_new, cap_or_inf);
}

// Colour update of the reference (src/tsdf.cu:743-745): per channel (c * w_old + rgb) / w_new
// in FP32 (the product and sum are exact integers below 2^24; the quotient correctly
// rounded by div_rn), truncated to 8 bits. w_old is the stored weight before this update;
// c and px are packed {r | g << 8 | b << 16}.
__device__ __forceinline__ uint32_t color_update(uint32_t c, float w_old, uint32_t px) {
  const float w_new = w_old + 1.0f;
  uint32_t out = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float num = __builtin_fmaf((float)((c >> (8 * k)) & 255u), w_old, (float)((px >> (8 * k)) & 255u));
    out |= (uint32_t)(uint8_t)div_rn_int(num, w_new) << (8 * k);
  }
  return out;
}

// One wave per listed block: lane l owns voxels l + 64k (x = k, y = l >> 3, z = l & 7), so
// each of its 8 loads is one coalesced 512-byte wave access and all 8 are in flight at once.
// A launch integrates a batch of k frames (ids f0 .. f0+k-1, inputs cur .. cur+k-1): the list
// holds every block that any of them touched, once, with the mask of those frames. The
// block's voxels are read once, updated by each marked frame in frame order (each update
// unpacks and packs the storage word, exactly as that frame's own load and store would), and
// written once; a voxel gets the same sequence of updates as frame-by-frame integration.
// Worker w of `workers` (one wave each) takes an equal share of the list, its 64 stripes
// seen as one index space; workers synchronise only within their own wave.
// kColor: the colour pool is updated from each frame's packed RGB at the same pixels.
// Where a block's voxels fall in a frame, which selects frame_update's variant:
constexpr int kPlaceEdge = 0;    // some voxel may project outside the image: per-voxel checks
constexpr int kPlaceInside = 1;  // every voxel inside: no checks, depth gathered from the image
constexpr int kPlaceWindow = 2;  // every voxel inside, its pixels in a window staged in LDS
// LDS window per wave (pixels): the pre-scaled depth of the pixel rectangle that a block's
// voxels project into, rows of a power-of-two pitch <= 64. A 5 mm block at 2 m covers about
// 12 x 12 pixels; at 0.5 m about 50 x 50, which does not fit and is gathered from the image.
constexpr uint32_t kWin = 512;

// A block's window in a frame: the image rectangle of columns ulo .. ulo + cols - 1 and rows
// vlo .. vlo + rows - 1 (both inside the image), in LDS at pitch 2^shift.
struct DepthWindow {
  uint32_t first;  // byte offset 4 (vlo W + ulo) of the rectangle's first pixel in the image
  uint32_t shift;  // log2 of the LDS row pitch in pixels, <= 6
  uint32_t rows, cols;
  uint32_t bias;   // LDS byte offset of pixel (u, v) = (bits(M + v) << (shift + 2)) + (bits(M + u) << 2) - bias
};

// Smallest and largest of x over lanes 0, 7, 56 and 63, the lanes whose columns end at the
// block's 8 corners: the half-row mirror (DPP) pairs lane 0 with 7 and 56 with 63, then two
// reads of lanes 0 and 56.
__device__ __forceinline__ uint32_t half_mirror(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x141, 0xf, 0xf, false);
}
__device__ __forceinline__ int32_t corner_min(int32_t x) {
  x = min(x, (int32_t)half_mirror((uint32_t)x));
  return min(__builtin_amdgcn_readlane(x, 0), __builtin_amdgcn_readlane(x, 56));
}
__device__ __forceinline__ int32_t corner_max(int32_t x) {
  x = max(x, (int32_t)half_mirror((uint32_t)x));
  return max(__builtin_amdgcn_readlane(x, 0), __builtin_amdgcn_readlane(x, 56));
}

// Placement of a block in a frame (wave-uniform). The lane's column runs from voxel 0 to
// voxel 7 of cl + q*ax (homogeneous coordinates); the block's lattice box is the convex hull
// of its 8 corners, the column end points of lanes 0, 7, 56 and 63. For depths > 0 the
// perspective image of a convex combination is a convex combination of the images, and a
// voxel's own projection differs from the exact one, as the corners' do, by far less than
// 1/2 pixel. So if every end point has z in [2^-60, 2^60], every voxel's z lies in
// [2^-64, 2^64], and with umin, umax the least and largest u = X/z of the corners a voxel's
// pixel round(u) lies in [round(umin - 1/2), round(umax + 1/2)] (for errors below 1/2 these
// roundings cannot cross the voxel's); likewise v. That rectangle is the block's window.
// kPlaceEdge unless the window lies inside the image; then kPlaceWindow when it fits kWin at a
// pitch of at most 64 pixels, else kPlaceInside. The bounds are the bits of (u -+ 1/2) + M as
// signed integers: for positive sums below 2^24 they are M's plus the rounded bound, larger
// sums exceed every image bound and negative ones lie below them.
__device__ __forceinline__ int block_place(const FrameArgs& a, const float ax[3], const float cl[3], DepthWindow& w) {
  bool zin = true;
  float u[2], v[2];
#pragma unroll
  for (int e = 0; e < 2; ++e) {  // voxel 0 of the column is cl itself
    const float z = e ? __builtin_fmaf(7.0f, ax[2], cl[2]) : cl[2];
    const float iz = recip_rn(z);
    u[e] = (e ? __builtin_fmaf(7.0f, ax[0], cl[0]) : cl[0]) * iz;
    v[e] = (e ? __builtin_fmaf(7.0f, ax[1], cl[1]) : cl[1]) * iz;
    zin = zin & (z >= 0x1p-60f) & (z <= 0x1p60f);
  }
  // computed whatever the placement, so that the window stays in scalar registers
  const int32_t lu = corner_min((int32_t)__float_as_uint((fminf(u[0], u[1]) - 0.5f) + kPixMagic));
  const int32_t hu = corner_max((int32_t)__float_as_uint((fmaxf(u[0], u[1]) + 0.5f) + kPixMagic));
  const int32_t lv = corner_min((int32_t)__float_as_uint((fminf(v[0], v[1]) - 0.5f) + kPixMagic));
  const int32_t hv = corner_max((int32_t)__float_as_uint((fmaxf(v[0], v[1]) + 0.5f) + kPixMagic));
  const uint32_t cols = (uint32_t)(hu - lu) + 1u, rows = (uint32_t)(hv - lv) + 1u;
  // pitch: the power of two >= cols (2 for one column), at most 64
  const uint32_t shift = 32u - (uint32_t)__builtin_clz(min(cols - 1u, 63u) | 1u);
  w.first = 4u * (((uint32_t)lv - kPixMagicBits) * (uint32_t)a.W + ((uint32_t)lu - kPixMagicBits));
  w.shift = shift;
  w.rows = rows;
  w.cols = cols;
  w.bias = ((uint32_t)lv << (shift + 2u)) + ((uint32_t)lu << 2);
  const int32_t mb = (int32_t)kPixMagicBits;
  const bool inside = __builtin_amdgcn_ballot_w64(!zin) == 0 && lu >= mb && hu <= mb + a.W - 1 && lv >= mb &&
                      hv <= mb + a.H - 1;
  const bool fits = cols <= 64u && rows <= (kWin >> shift);
  return inside ? (fits ? kPlaceWindow : kPlaceInside) : kPlaceEdge;
}

// LDS addresses as integers: the byte address of an LDS location, and a read from one.
typedef __attribute__((address_space(3))) float LdsFloat;
__device__ __forceinline__ uint32_t lds_address(const float* p) { return (uint32_t)(size_t)(const LdsFloat*)p; }
__device__ __forceinline__ float lds_read(uint32_t byte_address) { return *(const LdsFloat*)(size_t)byte_address; }

// The window's pixels of pre-scaled depth image dmf into the wave's LDS window at byte address
// win (row r at win + 4 (r << shift)), by loads that write LDS directly: load t covers rows
// t * 64 >> shift onwards, lane l column l mod 2^shift of row l >> shift, and writes word
// 64 t + l. The loads use no registers, so the caller computes its pixels while they are in
// flight and waits for them (window_ready) before its first read.
__device__ __forceinline__ void stage_window(const FrameArgs& a, const float* dmf, const DepthWindow& w, uint32_t win) {
  const uint32_t lane = threadIdx.x & 63u, w4 = 4u * (uint32_t)a.W;
  const uint32_t row = lane >> w.shift, col = lane & ((1u << w.shift) - 1u), step = 64u >> w.shift;
  const char* src = reinterpret_cast<const char*>(dmf) + w.first;
  const uint32_t off = row * w4 + 4u * col;
  const int left = col < w.cols && row < w.rows ? (int)(w.rows - row) : 0;  // rows the lane loads, at most
  const uint32_t n_it = (w.rows + step - 1u) >> (6u - w.shift);             // loads of the widest lane
  for (uint32_t t = 0; t < n_it; ++t) {
    const char* rows_t = src + t * step * w4;  // uniform
    if ((int)(t * step) < left)
      __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(rows_t + off),
                                       (__attribute__((address_space(3))) void*)(size_t)(win + 256u * t), 4, 0, 0);
  }
}
// Wait for the window's loads (vmcnt 0; the compiler does not order a read through an integer
// LDS address after them by itself).
__device__ __forceinline__ void window_ready() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// One frame's update of the lane's 8 voxels of a block (homogeneous column cl, step ax): per
// half, all projections, then all pre-scaled depth (and colour) fetches, then the updates, so
// the fetches are in flight together. Branch-free: every lane computes, the observation flag
// selects. dmf: the frame's pre-scaled depth image (scaled_depth of each pixel).
// kPlace (block_place's result): kPlaceInside and kPlaceWindow leave out the range checks,
// which are known to pass (the values computed are the same); kPlaceWindow reads the depth
// from the window that stage_window is loading (pitch 2^wshift; wbase: the LDS byte address of
// pixel (u, v) is (bits(M + v) << (wshift + 2)) + (bits(M + u) << 2) + wbase modulo 2^32, the
// window's address minus DepthWindow::bias). (Issuing the next frame's gathers before
// this frame's updates, with two register sets, was slower: 92 against 86 us per FP32 batch
// launch.)
template <int kStorage, bool kColor, int kPlace, bool kTab>
__device__ __forceinline__ void frame_update(const FrameArgs& a, const float ax[3], const float cl[3],
                                             const float* dmf, const uint32_t* rgba, uint32_t wshift,
                                             uint32_t wbase, typename VoxelIO<kStorage>::Word old[8],
                                             uint32_t cold[8], uint32_t& upd, const float* rcp_tab) {
  using IO = VoxelIO<kStorage>;
  using Word = typename IO::Word;
  constexpr bool kInside = kPlace != kPlaceEdge, kWindow = kPlace == kPlaceWindow;
  const uint32_t w4 = (4u * (uint32_t)a.W) & 0xffffffu, bias = pix_bias(a.W);
  // two halves of 4 voxels: fewer values live at once (occupancy) for 4 fetches in flight
#pragma unroll
  for (int h = 0; h < 8; h += 4) {
    float zc[4];
    int pix[4];
    uint32_t off[4];   // kInside: byte offset of the pixel in the depth (and RGB) image, off a scalar base
    uint32_t loff[4];  // kWindow: byte offset of the pixel in the LDS window
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float fq = (float)(h + r);  // voxel 0 is the column itself (fma(0, a, c) = c but for -0)
      const float X = h + r ? __builtin_fmaf(fq, ax[0], cl[0]) : cl[0];
      const float Y = h + r ? __builtin_fmaf(fq, ax[1], cl[1]) : cl[1];
      zc[r] = h + r ? __builtin_fmaf(fq, ax[2], cl[2]) : cl[2];
      const float iz = recip_rn(zc[r]);
      if constexpr (kInside) {
        const uint32_t ub = __float_as_uint(__builtin_fmaf(X, iz, kPixMagic));
        const uint32_t vb = __float_as_uint(__builtin_fmaf(Y, iz, kPixMagic));
        if constexpr (kWindow) {
          loff[r] = (((vb << wshift) + ub) << 2) + wbase;
          if constexpr (kColor) off[r] = (vb & 0xffffffu) * w4 + ((ub << 2) - bias);
        } else {
          off[r] = (vb & 0xffffffu) * w4 + ((ub << 2) - bias);
        }
      } else {
        pix[r] = pixel_of(X, Y, zc[r], iz, a.W, a.H);
      }
    }
    float dv[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if constexpr (kWindow) {
        if (h == 0 && r == 0) window_ready();  // the pixels are computed: now wait for the window
        dv[r] = lds_read(loff[r]);
      }
      else if constexpr (kInside) dv[r] = *reinterpret_cast<const float*>(reinterpret_cast<const char*>(dmf) + off[r]);
      else dv[r] = dmf[max(pix[r], 0)];
    }
    uint32_t px[4];
    if constexpr (kColor) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if constexpr (kInside) px[r] = *reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(rgba) + off[r]);
        else px[r] = rgba[max(pix[r], 0)];
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int q = h + r;
      const float t = __builtin_fmaf(-zc[r], a.inv_mu, dv[r]);
      const bool ok = (kInside || pix[r] >= 0) & (t > -1.0f);
      const float dist = fminf(1.0f, t);
      float sv, wv;
      IO::unpack(old[q], sv, wv);
      if constexpr (kColor) {
        const uint32_t c = color_update(cold[q], wv, px[r]);
        cold[q] = ok ? c : cold[q];
      }
      mean_update<kTab>(sv, wv, dist, a.cap, rcp_tab);
      const Word nw = IO::pack(sv, wv);
      old[q] = ok ? nw : old[q];
      if constexpr (kColor) upd |= (uint32_t)ok << q;
    }
    __builtin_amdgcn_sched_barrier(0);  // keep the halves apart: the scheduler would merge them
  }
}

// Per frame of a batch, staged in LDS by stage_frames: the pose (row-major 3x4) and the
// frame's homogeneous voxel steps (block_frame_h).
struct FrameLds {
  float T[12];
  BlockFrame h;
};

// s_frm: the batch's k frames; rcp_tab: RN(1/w) for w < kRcpTab (read when kTab); s_win: the
// LDS byte address of the wave's depth window (kWin pixels). kShared: the block origin of
// each frame is computed once per block (launches of several frames), else in the frame loop.
template <int kStorage, bool kColor, bool kTab, bool kShared>
__device__ void integrate_worker(const FrameArgs& a, uint32_t worker, uint32_t workers, uint32_t f0, uint32_t k,
                                 uint32_t parity, unsigned long long* s_val, unsigned long long* s_key,
                                 uint32_t* s_mark, const FrameLds* s_frm, const float* rcp_tab, uint32_t s_win) {
  using IO = VoxelIO<kStorage>;
  using Word = typename IO::Word;
  FrameCtl* ctl = a.ctl;
  const int lane = threadIdx.x & 63;
  const uint32_t* lcount = list_count_of(ctl, parity);
  const uint32_t* list = list_of(a, parity);
  uint32_t* marks = marks_of(a, parity);
  // lane s holds stripe s's length and exclusive offset
  const uint32_t my_len = min(lcount[lane * kLine], a.stripe_cap);
  uint32_t incl = my_len;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)incl, off);
    if (lane >= off) incl += y;
  }
  const uint32_t my_off = incl - my_len;
  const uint32_t n = read_lane_u32(incl, 63);
  const size_t frame_px = (size_t)a.W * a.H;

  // Lane j fetches the list entry, slot value, key and mark of this worker's j-th block
  // (clearing the mark for the next batch of this parity) and parks them in LDS, so the
  // per-block chain is a single round trip (voxels + depth). The block loop reads them back
  // from LDS: a use of the loaded registers inside the loop would make the compiler wait for
  // every outstanding vector memory operation, the previous block's stores included (vmcnt
  // counts stores on this architecture).
  for (uint32_t t0 = worker; t0 < n; t0 += 64u * workers) {
    const uint32_t tj = t0 + (uint32_t)lane * workers;
    const uint32_t tq = min(tj, n - 1u);
    int lo = 0, hi = 63;  // stripe of entry tq: the last stripe whose offset is <= tq
#pragma unroll
    for (int it = 0; it < 6; ++it) {  // all lanes active: the shuffles read defined values
      const int mid = (lo + hi + 1) >> 1;
      const uint32_t om = (uint32_t)__shfl((int)my_off, mid);
      if (om <= tq) lo = mid;
      else hi = mid - 1;
    }
    const uint32_t stripe_off = (uint32_t)__shfl((int)my_off, lo);
    unsigned long long jval = ~0ull, jkey = 0ull;
    uint32_t jmark = 0u;
    if (tj < n) {
      const uint32_t sl = list[(size_t)lo * a.stripe_cap + (tj - stripe_off)];
      jval = a.hash.vals[sl];
      jkey = a.hash.keys[sl];
      jmark = marks[sl];
      marks[sl] = 0u;
    }
    s_val[lane] = jval;
    s_key[lane] = jkey;
    s_mark[lane] = jmark;
    __builtin_amdgcn_wave_barrier();  // one wave owns this slice: it only needs its own writes
    const uint32_t nj = min(64u, (n - t0 + workers - 1) / workers);
    for (uint32_t j = 0; j < nj; ++j) {
      const unsigned long long val = s_val[j];
      const unsigned long long keyj = s_key[j];
      const uint32_t mark = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_mark[j]) & ((1u << k) - 1u);
      const uint32_t idx = (uint32_t)val;
      if (idx >= a.pool_cap) {
        if (lane == 0) raise_flag(ctl, kFlagPoolFull);
        continue;
      }
      const bool fresh = (uint32_t)(val >> 32) - f0 < k;  // born in this batch: starts at zero
      int bx, by, bz;
      unpack_key(keyj, bx, by, bz);
      Word* vp = reinterpret_cast<Word*>(a.voxels) + (size_t)idx * kB3 + lane;
      Word old[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) old[q] = fresh ? IO::zero() : vp[64 * q];
      uint32_t* cp = kColor ? a.colors + (size_t)idx * kB3 + lane : nullptr;
      uint32_t cold[8];
      if constexpr (kColor) {
#pragma unroll
        for (int q = 0; q < 8; ++q) cold[q] = fresh ? 0u : cp[64 * q];
      }
      uint32_t upd = 0u;  // colour volumes: bit q set when voxel q was observed in the batch
      // the block origin in homogeneous pixel coordinates per frame of the batch: with kShared
      // lane j < k computes frame j's and the frame loop takes it with readlane, so these
      // wave-uniform values are computed once per block rather than by every lane in every
      // frame (a single-frame launch would only add the readlanes, and computes it inline)
      float hbl[3] = {0.0f, 0.0f, 0.0f};
      if (kShared && (uint32_t)lane < k) {
        float T[12], cb[3];
#pragma unroll
        for (int q = 0; q < 12; ++q) T[q] = s_frm[lane].T[q];
        block_origin_cam(T, a.vox, bx, by, bz, cb);
        to_pixel_h(a.K, cb, hbl);
      }
      for (uint32_t m = mark; m != 0u; m &= m - 1u) {
        const uint32_t jf = (uint32_t)__builtin_ctz(m);
        const FrameLds& F = s_frm[jf];
        const BlockFrame hf = F.h;
        float hb[3], cl[3];
        if constexpr (kShared) {
#pragma unroll
          for (int r = 0; r < 3; ++r)
            hb[r] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(hbl[r]), (int)jf));
        } else {
          float T[12], cb[3];
#pragma unroll
          for (int q = 0; q < 12; ++q) T[q] = F.T[q];
          block_origin_cam(T, a.vox, bx, by, bz, cb);
          to_pixel_h(a.K, cb, hb);
        }
        column_h(hf, hb, lane >> 3, lane & 7, cl);
        const size_t fo = ((size_t)parity * kMaxBatch + jf) * frame_px;
        const uint32_t* rgba = kColor ? a.rgba + fo : nullptr;
        const float* dmf = a.dm + fo;
        DepthWindow win;
        const int place = block_place(a, hf.ax, cl, win);
        if (place == kPlaceWindow) {
          stage_window(a, dmf, win, s_win);
          frame_update<kStorage, kColor, kPlaceWindow, kTab>(a, hf.ax, cl, dmf, rgba, win.shift, s_win - win.bias,
                                                             old, cold, upd, rcp_tab);
        } else if (place == kPlaceInside) {
          frame_update<kStorage, kColor, kPlaceInside, kTab>(a, hf.ax, cl, dmf, rgba, 0u, 0u, old, cold, upd,
                                                             rcp_tab);
        } else {
          frame_update<kStorage, kColor, kPlaceEdge, kTab>(a, hf.ax, cl, dmf, rgba, 0u, 0u, old, cold, upd, rcp_tab);
        }
      }
      // Depth-only volumes write the block whole (unobserved voxels keep the words read,
      // fresh blocks zero there; the wave owns the block in this launch), which saves the
      // per-voxel mask; colour volumes (12 B per voxel) write fresh blocks whole and other
      // blocks where observed.
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        if constexpr (kColor) {
          if (fresh || ((upd >> q) & 1u)) {
            vp[64 * q] = old[q];
            cp[64 * q] = cold[q];
          }
        } else {
          vp[64 * q] = old[q];
        }
      }
    }
    __builtin_amdgcn_wave_barrier();  // the next batch of entries overwrites s_val / s_key / s_mark
  }
}

// End of a workgroup of a batch launch (all 256 lanes call it): the last of `groups`
// workgroups to finish (two-level striped counter) closes the batch (first frame f0, k
// frames, list parity): its per-frame counts are summed into the statistics and cleared with
// the list lengths; the frame id, input cursor and batch id advance. No fence before the
// counters: the closing workgroup reads no voxel data, and the voxel writes are published by
// the kernel boundary (a device-scope fence per workgroup would write back the L2 thousands
// of times per batch).
__device__ void batch_done(const FrameArgs& a, uint32_t groups, uint32_t f0, uint32_t k, uint32_t parity) {
  __shared__ uint32_t s_closes;
  FrameCtl* ctl = a.ctl;
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t g = blockIdx.x & (kStripes - 1);
    const uint32_t in_group = groups / kStripes + (g < groups % kStripes ? 1u : 0u);
    const uint32_t live = min(groups, (uint32_t)kStripes);  // stripes holding a workgroup
    uint32_t closes = 0;
    if (atomicAdd(&ctl->done[g * kLine], 1u) == in_group - 1) closes = atomicAdd(&ctl->done_top, 1u) == live - 1;
    s_closes = closes;
  }
  __syncthreads();
  if (threadIdx.x >= 64 || !s_closes) return;
  // The counts were final before this launch started (the batch's allocation ran earlier):
  // each lane reads and clears one stripe's line, one round trip in all.
  const int lane = threadIdx.x;
  uint32_t* line = list_count_of(ctl, parity) + lane * kLine;
  uint32_t listed = min(line[0], a.stripe_cap);
  uint32_t carved = line[kLineCarve];
  uint32_t per[kMaxBatch];
#pragma unroll
  for (int j = 0; j < kMaxBatch; ++j) per[j] = (uint32_t)j < k ? line[1 + j] : 0u;
#pragma unroll
  for (int j = 0; j <= kMaxBatch; ++j)
    if ((uint32_t)j <= k) line[j] = 0u;
  line[kLineCarve] = 0u;
  ctl->done[lane * kLine] = 0u;
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    listed += (uint32_t)__shfl_xor((int)listed, m);
    carved += (uint32_t)__shfl_xor((int)carved, m);
#pragma unroll
    for (int j = 0; j < kMaxBatch; ++j) per[j] += (uint32_t)__shfl_xor((int)per[j], m);
  }
  if (lane == 0) {
    unsigned long long sum = 0;
#pragma unroll
    for (int j = 0; j < kMaxBatch; ++j) sum += per[j];
    uint32_t last = 0;
#pragma unroll
    for (int j = 0; j < kMaxBatch; ++j)
      if ((uint32_t)j + 1u == k) last = per[j];
    ctl->touched_last = last;
    ctl->touched_total += sum;
    ctl->integrated_total += listed;
    ctl->carved_total += carved;
    ctl->frames += k;
    ctl->cursor[a.cursor_id] += k;
    ctl->frame_id = f0 + k;
    ctl->batch_id += 1u;
    ctl->done_top = 0u;
  }
}

constexpr int kIntegrateWaves = 4;  // independent workers per workgroup
// Register budget of the integration kernels (measured on room_medium_5mm): depth only, 5
// waves per SIMD (96 VGPRs; uncapped the two frame_update variants take 106 and leave 4
// waves; capped, at most 7 values spill and every voxel format runs faster); colour, 4 waves
// (128 VGPRs: at 96 about 50 values spill into the frame loop, 13 % slower).
#define VH_INTEGRATE_OCCUPANCY __attribute__((amdgpu_waves_per_eu(kColor ? 4 : 5)))

// The k poses of a batch (inputs cur .. cur+k-1) into LDS, read by every block of the
// workgroup's workers (a global load per block and frame would put a memory round trip in
// front of each frame's projections). All 256 lanes call it.
__device__ __forceinline__ void stage_poses(const FrameArgs& a, uint32_t cur, uint32_t k, float (*s_pose)[12]) {
  const uint32_t t = threadIdx.x;
  if (t < 12u * k) {
    const uint32_t j = t / 12u, q = t - 12u * j;
    s_pose[j][q] = a.poses[16 * (size_t)input_slot(cur, j, a.n_seq) + q];
  }
  __syncthreads();
}
// The poses and the homogeneous steps of a batch's frames into LDS. All 256 lanes call it.
__device__ __forceinline__ void stage_frames(const FrameArgs& a, uint32_t cur, uint32_t k, FrameLds* s_frm,
                                             float* rcp_tab, bool tab) {
  const uint32_t t = threadIdx.x;
  if (t < 12u * k) {
    const uint32_t j = t / 12u, q = t - 12u * j;
    s_frm[j].T[q] = a.poses[16 * (size_t)input_slot(cur, j, a.n_seq) + q];
  }
  if (tab)
    for (uint32_t i = t; i < kRcpTab; i += blockDim.x) rcp_tab[i] = recip_rn((float)i);
  __syncthreads();
  if (t < k) {
    float T[12];
#pragma unroll
    for (int q = 0; q < 12; ++q) T[q] = s_frm[t].T[q];
    block_frame_h(T, a.vox, a.K, s_frm[t].h);
  }
  __syncthreads();
}

// True when every weight the batch can meet is below kRcpTab - 1: capped below it, or fewer
// frames integrated (with this batch) than that (a voxel gains at most 1 per frame; a merge
// of streamed blocks adds the weights of disjoint frames).
__device__ __forceinline__ bool rcp_table_covers(const FrameArgs& a, const FrameCtl* ctl, uint32_t k) {
  return a.cap < (float)(kRcpTab - 1) || uniform64(ctl->frames) + k < (unsigned long long)(kRcpTab - 1);
}

// The integration's LDS: per wave the staged list entries, slot values, keys and marks and
// the depth window; per workgroup the batch's frames and the reciprocal table.
struct IntegrateLds {
  FrameLds frm[kMaxBatch];  // first: a frame's words are then within one read's offset range
  unsigned long long val[kIntegrateWaves][64], key[kIntegrateWaves][64];
  uint32_t mark[kIntegrateWaves][64];
  float rcp[kRcpTab];
  float win[kIntegrateWaves][kWin];
};

// The workgroup's share of the batch of k frames (first frame f0, list parity, first input
// cur). All 256 lanes call it.
template <int kStorage, bool kColor, bool kShared>
__device__ __forceinline__ void integrate_group(const FrameArgs& a, uint32_t group, uint32_t groups, uint32_t f0,
                                                uint32_t k, uint32_t parity, uint32_t cur, IntegrateLds& L) {
  const int wave = threadIdx.x >> 6;
  const bool tab = rcp_table_covers(a, a.ctl, k);
  stage_frames(a, cur, k, L.frm, L.rcp, tab);
  const uint32_t worker = group * kIntegrateWaves + wave, workers = groups * kIntegrateWaves;
  const uint32_t win = uniform(lds_address(L.win[wave]));
  if (tab)
    integrate_worker<kStorage, kColor, true, kShared>(a, worker, workers, f0, k, parity, L.val[wave], L.key[wave],
                                                      L.mark[wave], L.frm, L.rcp, win);
  else
    integrate_worker<kStorage, kColor, false, kShared>(a, worker, workers, f0, k, parity, L.val[wave], L.key[wave],
                                                       L.mark[wave], L.frm, L.rcp, win);
}

// Integration of the batch of k frames that the last allocation launch prepared (kShared: k > 1).
template <int kStorage, bool kColor, bool kShared>
__global__ __launch_bounds__(64 * kIntegrateWaves) VH_INTEGRATE_OCCUPANCY void k_integrate(FrameArgs a, uint32_t k) {
  __shared__ IntegrateLds L;
  const FrameCtl* ctl = a.ctl;
  // control values are uniform: keep them (and what derives from them) in scalar registers
  const uint32_t f0 = uniform(ctl->frame_id), parity = uniform(ctl->batch_id) & 1u,
                 cur = uniform(ctl->cursor[a.cursor_id]);
  integrate_group<kStorage, kColor, kShared>(a, blockIdx.x, gridDim.x, f0, k, parity, cur, L);
  batch_done(a, gridDim.x, f0, k, parity);
}

// Integration of batch b (k_int frames) and allocation of batch b+1 (k_alloc frames) in one
// launch (bound sequences): the integration workgroups come first in the grid, so they take
// every CU slot at the start, and the allocation workgroups (short, latency-bound) fill the
// slots that the integration's tail frees. Every workgroup counts towards the close of batch
// b, so the frame id, cursor and batch id the allocation workgroups read at their start
// cannot advance under them.
template <int kStorage, bool kU16, bool kColor>
__global__ __launch_bounds__(64 * kIntegrateWaves) VH_INTEGRATE_OCCUPANCY void k_integrate_alloc(FrameArgs a, uint32_t int_groups,
                                                                         uint32_t alloc_groups, uint32_t tiles,
                                                                         uint32_t k_int) {
  // a workgroup either allocates or integrates: the two share the LDS
  __shared__ union {
    AllocLds<kAllocRows> alloc;
    IntegrateLds integ;
  } L;
  const uint32_t b = blockIdx.x;
  const bool alloc = b >= int_groups;
  const uint32_t idx = alloc ? b - int_groups : b;
  const FrameCtl* ctl = a.ctl;
  const uint32_t f0 = uniform(ctl->frame_id), batch = uniform(ctl->batch_id), cur = uniform(ctl->cursor[a.cursor_id]);
  if (alloc) {
    const uint32_t jb = idx / tiles;
    alloc_tile<kU16, kAllocRows, kColor>(a, idx - jb * tiles, f0 + k_int + jb, jb, (batch + 1u) & 1u,
                                 input_slot(cur, k_int + jb, a.n_seq), L.alloc);
  } else {  // uniform branch: the whole workgroup integrates
    integrate_group<kStorage, kColor, true>(a, idx, int_groups, f0, k_int, batch & 1u, cur, L.integ);
  }
  batch_done(a, gridDim.x, f0, k_int, batch & 1u);
}

// ============================================================================ carving
// Conservative form of the observation test for a whole block: false only if no voxel of block
// (bx, by, bz) can pass frame T's test (z >= 2^-64, its pixel inside the image, 0 < d <=
// max_depth and d - z > -mu there). The voxels lie in the box spanned by the block's lattice
// points 0..7 along each axis. The box's 8 corners in camera coordinates bound every voxel's z
// (the camera map is affine), and, when the whole box lies in front of the camera, every
// voxel's projection (the perspective image of a convex set in z > 0 is the convex hull of its
// corners' images). Margins cover the rounding of the voxel arithmetic of the integration
// (frame_update) against these bounds: every camera coordinate it forms is within
// 4 * 2^-23 * mag of the exact one (mag bounds the magnitudes summed), so dz bounds the depth
// error and eps the pixel error ((fx + |u - cx|) * error / z for u = fx x / z + cx).
// tmax: the frame's max-depth tiles; a voxel is only updated if its z is below its pixel's
// depth + mu, so a block whose nearest corner lies beyond the largest valid depth of the tiles
// its pixels fall in, plus mu, is not.
__device__ bool block_may_update(const FrameArgs& a, const float* T, int bx, int by, int bz, const float* tmax) {
  BlockFrame bf;
  block_frame(T, a.vox, bf);
  float cb[3];
  block_origin_cam(T, a.vox, bx, by, bz, cb);
  const float mag = fabsf(cb[0]) + fabsf(cb[1]) + fabsf(cb[2]) + 24.0f * a.vox;
  const float dz = 1e-4f + 4e-6f * mag;
  const float zfar = a.max_depth + a.mu + dz;
  // bounding sphere of the box in z first (centre at lattice offset 3.5, radius 3.5 sqrt(3) vox)
  const float zc = cb[2] + 3.5f * ((bf.ax[2] + bf.ay[2]) + bf.az[2]);
  const float rad = 6.07f * a.vox;
  if (zc + rad < -dz || zc - rad > zfar) return false;
  float P[8][3];
  float zmin = INFINITY, zmax = -INFINITY;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const float ex = (c & 1) ? 7.0f : 0.0f, ey = (c & 2) ? 7.0f : 0.0f, ez = (c & 4) ? 7.0f : 0.0f;
#pragma unroll
    for (int r = 0; r < 3; ++r) P[c][r] = cb[r] + ((ex * bf.ax[r] + ey * bf.ay[r]) + ez * bf.az[r]);
    zmin = fminf(zmin, P[c][2]);
    zmax = fmaxf(zmax, P[c][2]);
  }
  if (zmax < -dz || zmin > zfar) return false;
  if (!(zmin > dz)) return true;  // the box reaches the camera plane: no projection bound
  float umin = INFINITY, umax = -INFINITY, vmin = INFINITY, vmax = -INFINITY;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const float iz = __builtin_amdgcn_rcpf(P[c][2]);
    const float u = a.K.fx * (P[c][0] * iz) + a.K.cx, v = a.K.fy * (P[c][1] * iz) + a.K.cy;
    umin = fminf(umin, u);
    umax = fmaxf(umax, u);
    vmin = fminf(vmin, v);
    vmax = fmaxf(vmax, v);
  }
  const float rel = 2e-6f * mag / zmin;
  const float eu = 1.0f + rel * (a.K.fx + fmaxf(fabsf(umin - a.K.cx), fabsf(umax - a.K.cx)));
  const float ev = 1.0f + rel * (a.K.fy + fmaxf(fabsf(vmin - a.K.cy), fabsf(vmax - a.K.cy)));
  if (umax < -0.5f - eu || umin > (float)a.W - 0.5f + eu || vmax < -0.5f - ev || vmin > (float)a.H - 0.5f + ev)
    return false;
  // pixels the voxels can round to (fmaxf / fminf also map NaN bounds to the whole image)
  const int x0 = (int)fmaxf(0.0f, floorf(umin - eu)), x1 = (int)fminf((float)(a.W - 1), ceilf(umax + eu));
  const int y0 = (int)fmaxf(0.0f, floorf(vmin - ev)), y1 = (int)fminf((float)(a.H - 1), ceilf(vmax + ev));
  if (!tmax) return true;
  const int tx0 = x0 / kTile, tx1 = x1 / kTile, ty0 = y0 / kTile, ty1 = y1 / kTile;
  if ((tx1 - tx0 + 1) * (ty1 - ty0 + 1) > 64) return true;  // a near block: not worth the reads
  float D = 0.0f;
  for (int ty = ty0; ty <= ty1; ++ty)
    for (int tx = tx0; tx <= tx1; ++tx) D = fmaxf(D, tmax[ty * a.tiles_x + tx]);
  return zmin <= D + a.mu + dz;
}

// Carve pass of the allocated batch (k frames, first frame id f0, list parity): one lane per
// resident pool block (the lanes of a fixed grid stride over the pool). Frame j may update a
// block born at or before frame f0 + j (a block born later in the batch did not exist when
// frame j was integrated); for each such frame that may update it (block_may_update) the
// block's mark gets bit j, and a block whose mark was empty (no band of the batch crossed it)
// is appended to the batch's list, in the workgroup's stripe; the carve word of that stripe's
// line counts the mark bits the pass added. Runs after the batch's allocation and before its
// integration.
constexpr uint32_t kCarveGroups = 2048;
__global__ __launch_bounds__(256) void k_carve(FrameArgs a, uint32_t k) {
  __shared__ float s_pose[kMaxBatch][12];
  FrameCtl* ctl = a.ctl;
  const uint32_t f0 = uniform(ctl->frame_id), parity = uniform(ctl->batch_id) & 1u,
                 cur = uniform(ctl->cursor[a.cursor_id]);
  const uint32_t n = min(uniform(ctl->pool_count), a.pool_cap);
  stage_poses(a, cur, k, s_pose);
  const int lane = threadIdx.x & 63;
  const uint32_t stripe = blockIdx.x & (kStripes - 1);
  uint32_t* line = list_count_of(ctl, parity) + stripe * kLine;
  uint32_t* marks = marks_of(a, parity);
  const size_t ntile = (size_t)a.tiles_x * a.tiles_y;
  for (uint32_t b0 = blockIdx.x * 256u; b0 < n; b0 += gridDim.x * 256u) {
    const uint32_t b = b0 + threadIdx.x;
    uint32_t mask = 0u, slot = 0u;
    if (b < n) {
      const unsigned long long info = a.block_info[b];
      slot = (uint32_t)info;
      const int32_t born = (int32_t)((uint32_t)(info >> 32) - f0);  // birth frame relative to f0
      int bx, by, bz;
      unpack_key(a.block_keys[b], bx, by, bz);
      for (uint32_t j = 0; j < k; ++j)
        if (born <= (int32_t)j &&
            block_may_update(a, s_pose[j], bx, by, bz, a.tilemax + ((size_t)parity * kMaxBatch + j) * ntile))
          mask |= 1u << j;
    }
    uint32_t old = 0u;
    if (mask) old = atomicOr(&marks[slot], mask);
    // block-frame updates the carve pass adds (frames whose bands did not cross the block)
    uint32_t added = (uint32_t)__popc(mask & ~old);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) added += (uint32_t)__shfl_xor((int)added, o);
    if (lane == 0 && added) atomicAdd(&line[kLineCarve], added);
    const bool first = mask != 0u && old == 0u;  // no band of the batch listed the block
    const uint64_t fm = __ballot(first);
    if (fm) {
      const int fl = first_lane(fm);
      uint32_t tbase = 0;
      if (lane == fl) tbase = atomicAdd(&line[0], (uint32_t)__popcll(fm));
      tbase = read_lane_u32(tbase, fl);
      if (first) {
        const uint32_t pos = tbase + lanes_below(fm);
        if (pos < a.stripe_cap) {
          list_of(a, parity)[(size_t)stripe * a.stripe_cap + pos] = slot;
        } else {
          atomicExch(&marks[slot], 0u);
          raise_flag(ctl, kFlagTouchedFull);
        }
      }
    }
  }
}

// ============================================================================ marching cubes
constexpr int kHalo = 9;
constexpr int kHalo3 = kHalo * kHalo * kHalo;
constexpr int kItemCap = 512;  // (cube, triangle) items listed per round (a block averages ~80)
// Extraction workgroups: kMcWaves independent waves, one block each, each with its own LDS
// (McShared). A wave only exchanges data with itself through LDS, and executes its LDS
// instructions in order, so wave_sync (a barrier to the compiler's reordering) suffices
// between a write and other lanes' reads. One wave per workgroup: four (24 instead of 16
// resident waves per CU) made the room_medium_5mm extraction slower, 0.87 against 0.69 ms.
constexpr int kMcWaves = 1;
__device__ __forceinline__ int mc_lane() { return (int)(threadIdx.x & 63u); }
__device__ __forceinline__ int mc_wave() { return (int)(threadIdx.x >> 6); }
__device__ __forceinline__ void wave_sync() { __builtin_amdgcn_wave_barrier(); }

// Cases in corner-plane order: bit 4 * dx + 2 * dy + dz of a permuted case is the sign bit of
// the cube corner at offset (dx, dy, dz), so that the low nibble holds the x = cx face and the
// high nibble the x = cx + 1 face, each in slot order (dy << 1) | dz. kPermOfCase maps a case
// (bit c: reference corner c) to its permuted form, four cases per word.
struct McPerm {
  uint32_t w[64];
};
constexpr McPerm mc_perm_table() {
  McPerm t{};
  for (int ci = 0; ci < 256; ++ci) {
    uint32_t p = 0;
    for (int c = 0; c < 8; ++c)
      p |= (uint32_t)((ci >> c) & 1) << (4 * corner_dx(c) + 2 * corner_dy(c) + corner_dz(c));
    t.w[ci >> 2] |= p << (8 * (ci & 3));
  }
  return t;
}
constexpr McPerm kPermHost = mc_perm_table();  // the compile-time copy that mc_perm_ok checks
constexpr bool mc_perm_ok() {  // a bijection fixing 0 and 255, and corner 0 of the x = cx face at bit 0
  bool seen[256] = {};
  for (int ci = 0; ci < 256; ++ci) {
    const uint32_t p = (kPermHost.w[ci >> 2] >> (8 * (ci & 3))) & 255u;
    if (seen[p]) return false;
    seen[p] = true;
  }
  return (kPermHost.w[0] & 255u) == 0u && (kPermHost.w[63] >> 24) == 255u && ((kPermHost.w[0] >> 8) & 255u) == 1u;
}
static_assert(mc_perm_ok(), "permuted case table is not a bijection");
__device__ __constant__ const McPerm kPermOfCase = mc_perm_table();

// LDS of one extraction wave (about 4.9 KB).
struct McShared {
  float val[kHalo3];   // 9^3 tile (x*81 + y*9 + z): tsdf, NaN where unobserved
  uint64_t col[81];    // column words (mc_col_bits) of the 9 x 9 (y, z) columns, index y*9 + z
  uint32_t nb[8];      // pool blocks of the neighbourhood, index (dx << 2) | (dy << 1) | dz
  uint16_t pc[256];    // by permuted case: case | triangles << 8 (case 255 stored as 0)
  uint8_t cnt[256];    // triangles of each case (table count)
  uint8_t cases[512];  // case per cube (x*64 + y*8 + z); 0 = no triangles
};
static_assert(offsetof(McShared, cnt) % 4 == 0, "mc_put_counts writes S.cnt four entries at a time");
// Item lists of the write passes (k_mc_write, k_ix_write).
struct McItems {
  uint32_t lane_off[65];     // exclusive prefix of table triangles per lane
  uint16_t items[kItemCap];  // (cube << 3) | t of the current round
};

// The table's triangle counts into S.cnt and S.pc in two steps, so that the loads can be
// issued first and the LDS writes made after the block's own loads are out (one lane: cases
// 4l .. 4l + 3).
struct McCounts {
  uint64_t t[4];
  uint32_t p;
};
__device__ __forceinline__ McCounts mc_fetch_counts() {
  const int lane = mc_lane();
  McCounts c;
#pragma unroll
  for (int j = 0; j < 4; ++j) c.t[j] = kTriTable[4 * lane + j];
  c.p = kPermOfCase.w[lane];
  return c;
}
__device__ __forceinline__ void mc_put_counts(McShared& S, const McCounts& c) {
  const int lane = mc_lane();
  uint32_t packed = 0u;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint32_t ci = 4u * (uint32_t)lane + (uint32_t)j, n = (uint32_t)(c.t[j] >> 60);
    packed |= n << (8 * j);
    S.pc[(c.p >> (8 * j)) & 255u] = (uint16_t)((ci == 255u ? 0u : ci) | n << 8);
  }
  *reinterpret_cast<uint32_t*>(&S.cnt[4 * lane]) = packed;
}

// Column word of one (y, z) column of the tile, built from its 9 values (x = 0..8) by
// mc_col_bits: bit 4x (x < 8) or bit 32 (x = 8) set iff the value is negative, bit 33 + x
// iff it is unobserved (NaN), bit 42 + x iff |value| < tiny. The 4x spacing interleaves the
// four columns of a cube's corners into one nibble per x-plane.
__device__ __forceinline__ void mc_col_bits(float v, int x, float tiny, uint32_t& q, uint32_t& m) {
  const uint32_t neg = v < 0.0f ? 1u : 0u;
  if (x < 8) q |= neg << (4 * x);
  else m |= neg;
  m |= (__builtin_isnan(v) ? 2u : 0u) << x;
  m |= (__builtin_fabsf(v) < tiny ? 1024u : 0u) << x;
}
__device__ __forceinline__ uint64_t mc_col_word(uint32_t q, uint32_t m) { return (uint64_t)q | (uint64_t)m << 32; }

// Threshold of mc_count_exact for the block at lattice origin (gx0, gy0, gz0): above the
// snap bound 1e-5 and above ulp(c) for every lattice coordinate c of the block and its
// +1 layer (ulp(c) <= |c| * 2^-23).
__device__ __forceinline__ float mc_tiny(int gx0, int gy0, int gz0) {
  const int m = max(max(max(abs(gx0), abs(gx0 + 8)), max(abs(gy0), abs(gy0 + 8))), max(abs(gz0), abs(gz0 + 8)));
  return 2e-5f + (float)(m + 1) * 0x1p-22f;
}

// One wave per pool block. The block's 8^3 voxels and the +x/+y/+z faces, edges and corner
// of its 7 neighbours (one hash lookup each) are staged as a 9^3 tile in LDS: eight
// coalesced rows of the block itself, one load per face, one for the three edges and the
// corner, so no lane divides an index. A corner is observed iff its block exists and its
// weight is > 0; unobserved corners hold NaN. kColor: the colour pool words of the same
// voxels go to col (9^3, same order) with the same accesses.
// Returns true iff the staged values (observed ones) include both a negative value and a
// non-negative one: otherwise every cube's case is 0 or 255 (or it is unobserved), and the
// block has no triangles. The column words of the 81 columns go to S.col, with kTiny the
// |value| < mc_tiny bits (otherwise none).
// The neighbourhood word of lane l of block b (k_mc_neighbours): lanes 0..7 hold the 8 entries.
__device__ __forceinline__ uint32_t mc_nbr(const McArgs& a, uint32_t b) {
  return a.nbrs[(size_t)b * 8 + (mc_lane() & 7)];
}
// nbr: mc_nbr(a, b), which the caller may have loaded early.
template <int kStorage, bool kColor, bool kTiny = false>
__device__ __forceinline__ bool mc_stage(const McArgs& a, uint32_t b, uint32_t nbr, McShared& S, uint32_t* col,
                                         int& gx0, int& gy0, int& gz0) {
  using IO = VoxelIO<kStorage>;
  using Word = typename IO::Word;
  const int lane = mc_lane();
  int bx, by, bz;
  unpack_key(a.block_keys[b], bx, by, bz);
  // the neighbourhood in lanes 0..7, read by the wave from its registers; loaded before the
  // block's own voxels, so that waiting for it does not wait for them
  const uint32_t nbv = lane < 8 ? nbr : kNoBlock;
  auto nb_of = [&](int o) { return (uint32_t)__builtin_amdgcn_readlane((int)nbv, o); };
  const Word* vox = reinterpret_cast<const Word*>(a.voxels);
  const float nan = __builtin_nanf("");
  auto value = [&](Word w) {
    float sv, wv;
    IO::unpack(w, sv, wv);
    return wv > 0.0f ? sv : nan;
  };
  const int hi = lane >> 3, lo = lane & 7;
  // this block: voxel lane + 64k = (k, hi, lo)
  const Word* own = vox + (size_t)b * kB3 + lane;
  Word w[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) w[k] = own[64 * k];
  if (lane < 8) S.nb[lane] = nbv;
  // faces x = 8 (voxel (0, hi, lo) of +x), y = 8 ((hi, 0, lo) of +y), z = 8 ((hi, lo, 0) of +z);
  // an absent neighbour's loads read this block instead (their values are not used), so that
  // every load is issued at once, with no wait between them
  const uint32_t nfx = nb_of(4), nfy = nb_of(2), nfz = nb_of(1);
  const auto safe = [&](uint32_t n) { return (size_t)(n != kNoBlock ? n : b) * kB3; };
  const Word wfx = vox[safe(nfx) + lane];
  const Word wfy = vox[safe(nfy) + hi * 64 + lo];
  const Word wfz = vox[safe(nfz) + hi * 64 + lo * 8];
  // edges (8, 8, z) of +x+y, (8, y, 8) of +x+z, (x, 8, 8) of +y+z (8 lanes each), corner
  // (8, 8, 8) of +x+y+z (lane 24)
  int edst = 0, ecol = 0, ex = 8;  // tile index, column and x of the lane's edge value
  uint32_t enb = kNoBlock;
  int esrc = 0;
  const uint32_t ne6 = nb_of(6), ne5 = nb_of(5), ne3 = nb_of(3), ne7 = nb_of(7);
  if (lane < 8) { enb = ne6; esrc = lane; edst = 8 * 81 + 8 * 9 + lane; ecol = 72 + lane; }
  else if (lane < 16) { enb = ne5; esrc = (lane - 8) * 8; edst = 8 * 81 + (lane - 8) * 9 + 8; ecol = (lane - 8) * 9 + 8; }
  else if (lane < 24) { enb = ne3; esrc = (lane - 16) * 64; edst = (lane - 16) * 81 + 8 * 9 + 8; ecol = 80; ex = lane - 16; }
  else if (lane == 24) { enb = ne7; esrc = 0; edst = kHalo3 - 1; ecol = 80; }
  const Word we = vox[safe(enb) + esrc];
  if constexpr (kColor) {
    const uint32_t* cp = a.colors;
    uint32_t c[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) c[k] = cp[(size_t)b * kB3 + lane + 64 * k];
    const uint32_t rfx = cp[safe(nfx) + lane], rfy = cp[safe(nfy) + hi * 64 + lo], rfz = cp[safe(nfz) + hi * 64 + lo * 8],
                   re = cp[safe(enb) + esrc];
    const uint32_t cfx = nfx != kNoBlock ? rfx : 0u, cfy = nfy != kNoBlock ? rfy : 0u, cfz = nfz != kNoBlock ? rfz : 0u,
                   ce = enb != kNoBlock ? re : 0u;
#pragma unroll
    for (int k = 0; k < 8; ++k) col[k * 81 + hi * 9 + lo] = c[k];
    col[8 * 81 + hi * 9 + lo] = cfx;
    col[hi * 81 + 8 * 9 + lo] = cfy;
    col[hi * 81 + lo * 9 + 8] = cfz;
    if (lane < 25) col[edst] = ce;
  }
  // lanes holding a negative / a non-negative value (NaN, unobserved, is neither), gathered as
  // ballots so that the OR is one scalar instruction per value
  uint64_t neg = 0u, nonneg = 0u;
  auto put = [&](int i, float v) {
    S.val[i] = v;
    neg |= __builtin_amdgcn_ballot_w64(v < 0.0f);
    nonneg |= __builtin_amdgcn_ballot_w64(v >= 0.0f);
  };
  gx0 = bx * kB;
  gy0 = by * kB;
  gz0 = bz * kB;
  const float tiny = kTiny ? mc_tiny(gx0, gy0, gz0) : 0.0f;
  // the border columns (y, 8), (8, z) and (8, 8) collect their values from other lanes
  if (lane < 17) S.col[lane < 8 ? lane * 9 + 8 : 64 + lane] = 0u;
  // own column (hi, lo): x = 0..7 from the block, x = 8 from the +x face
  uint32_t cq = 0u, cm = 0u;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const float v = value(w[k]);
    put(k * 81 + hi * 9 + lo, v);
    mc_col_bits(v, k, tiny, cq, cm);
  }
  const float vfx = nfx != kNoBlock ? value(wfx) : nan, vfy = nfy != kNoBlock ? value(wfy) : nan,
              vfz = nfz != kNoBlock ? value(wfz) : nan, ve = enb != kNoBlock ? value(we) : nan;
  put(8 * 81 + hi * 9 + lo, vfx);
  mc_col_bits(vfx, 8, tiny, cq, cm);
  S.col[hi * 9 + lo] = mc_col_word(cq, cm);
  put(hi * 81 + 8 * 9 + lo, vfy);
  put(hi * 81 + lo * 9 + 8, vfz);
  // the edge values: lanes 25.. hold NaN, which adds no bit, so the ballots are taken by the
  // whole wave (a ballot inside the branch would leave neg and nonneg differing across lanes)
  if (lane < 25) S.val[edst] = ve;
  neg |= __builtin_amdgcn_ballot_w64(ve < 0.0f);
  nonneg |= __builtin_amdgcn_ballot_w64(ve >= 0.0f);
  wave_sync();  // the border words are clear
  uint32_t yq = 0u, ym = 0u, zq = 0u, zm = 0u, eq = 0u, em = 0u;
  mc_col_bits(vfy, hi, tiny, yq, ym);  // (hi, 8, lo): column (8, lo) at x = hi
  mc_col_bits(vfz, hi, tiny, zq, zm);  // (hi, lo, 8): column (lo, 8) at x = hi
  mc_col_bits(ve, ex, tiny, eq, em);
  atomicOr(reinterpret_cast<unsigned long long*>(&S.col[72 + lo]), (unsigned long long)mc_col_word(yq, ym));
  atomicOr(reinterpret_cast<unsigned long long*>(&S.col[lo * 9 + 8]), (unsigned long long)mc_col_word(zq, zm));
  if (lane < 25) atomicOr(reinterpret_cast<unsigned long long*>(&S.col[ecol]), (unsigned long long)mc_col_word(eq, em));
  wave_sync();
  return neg != 0u && nonneg != 0u;
}

// Lane l classifies the 8 cubes of its (y, z) column (x = 0..7) from the column words of
// its 4 corner columns (slot k = (dy << 1) | dz: column (y + dy, z + dz)): interleaved, each
// x-plane is one nibble, and a cube's permuted case is the byte of its two faces. A cube has
// triangles iff its 8 corners are observed and its case is not 0 or 255 (bit c set iff
// corner c < 0, reference corner order); its case goes to S.cases. Returns the lane's table
// triangle count. near: bit x set iff cube x of the column has triangles and a corner with
// |value| < tiny (the kTiny bits of mc_stage; see mc_count_exact).
__device__ __forceinline__ uint32_t mc_classify(McShared& S, uint32_t& near) {
  const int lane = mc_lane(), i = (lane >> 3) * 9 + (lane & 7);
  const uint64_t w0 = S.col[i], w1 = S.col[i + 1], w2 = S.col[i + 9], w3 = S.col[i + 10];
  const uint32_t q = (uint32_t)w0 | (uint32_t)w1 << 1 | (uint32_t)w2 << 2 | (uint32_t)w3 << 3;
  const uint32_t m0 = (uint32_t)(w0 >> 32), m1 = (uint32_t)(w1 >> 32), m2 = (uint32_t)(w2 >> 32),
                 m3 = (uint32_t)(w3 >> 32);
  const uint32_t q8 = (m0 & 1u) | (m1 & 1u) << 1 | (m2 & 1u) << 2 | (m3 & 1u) << 3;  // x = 8 nibble
  const uint32_t um = m0 | m1 | m2 | m3;  // unobserved / small bits of any corner column
  uint32_t ntri = 0;
  near = 0u;
#pragma unroll
  for (int cx = 0; cx < 8; ++cx) {
    const uint32_t p = cx < 7 ? (q >> (4 * cx)) & 255u : (q >> 28) | (q8 << 4);
    const uint32_t e = S.pc[p];
    const bool obs = ((um >> (1 + cx)) & 3u) == 0u;
    const uint32_t ci = obs ? e & 255u : 0u, n = obs ? e >> 8 : 0u;
    if (n != 0u && ((um >> (10 + cx)) & 3u) != 0u) near |= 1u << cx;
    S.cases[cx * 64 + lane] = (uint8_t)ci;
    ntri += n;
  }
  return ntri;
}

// Exclusive prefix of the lanes' counts n into I.lane_off, I.lane_off[64] the total, which it
// returns.
__device__ __forceinline__ uint32_t mc_lane_prefix(McItems& I, uint32_t n) {
  const int lane = mc_lane();
  uint32_t x = n;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)x, off);
    if (lane >= off) x += y;
  }
  I.lane_off[lane] = x - n;
  if (lane == 63) I.lane_off[64] = x;
  wave_sync();
  return I.lane_off[64];
}

__device__ __forceinline__ uint32_t mc_wave_sum(uint32_t n) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) n += (uint32_t)__shfl_xor((int)n, m);
  return n;
}

// Per cube edge e, 5 bits at 5e: the offset of its lower end point in the cube (dx | dy << 1
// | dz << 2) and its axis << 3 (0 x, 1 y, 2 z).
constexpr uint64_t edge_lo_table() {
  uint64_t t = 0;
  for (int e = 0; e < 12; ++e) {
    const int a = edge_a(e), b = edge_b(e);
    const int dx = corner_dx(a) < corner_dx(b) ? corner_dx(a) : corner_dx(b);
    const int dy = corner_dy(a) < corner_dy(b) ? corner_dy(a) : corner_dy(b);
    const int dz = corner_dz(a) < corner_dz(b) ? corner_dz(a) : corner_dz(b);
    const int axis = corner_dx(a) != corner_dx(b) ? 0 : (corner_dy(a) != corner_dy(b) ? 1 : 2);
    t |= (uint64_t)(dx | dy << 1 | dz << 2 | axis << 3) << (5 * e);
  }
  return t;
}
constexpr uint64_t kEdgeLo = edge_lo_table();
// kEdgeUp bit e: the reference's first end point of cube edge e (edge_a) is its upper end.
constexpr uint32_t edge_up_table() {
  uint32_t m = 0;
  for (int e = 0; e < 12; ++e) {
    const int a = edge_a(e), b = edge_b(e);
    if (corner_dx(a) + corner_dy(a) + corner_dz(a) > corner_dx(b) + corner_dy(b) + corner_dz(b)) m |= 1u << e;
  }
  return m;
}
constexpr uint32_t kEdgeUp = edge_up_table();

// vertex_interp (the reference's VertexInterp, src/tsdf.cu:1640-1660) from tile corner ia to
// tile corner ib of one lattice edge along axis, with kColor vertex_interp_rgb's colour. The
// end points differ along that axis only, so the other two coordinates are the lower end
// point's (fx, fy, fz) exactly (mu * 0 adds zero), and p2 - p1 along the axis is exactly +-1
// (g: the lower end point's lattice coordinate on the axis; up: ia is the upper end). The
// quotient is formed on every lane and dropped where a snap applies, so that the wave takes no
// branch.
template <bool kColor>
__device__ __forceinline__ void mc_interp(const McShared& S, const uint32_t* col, int ia, int ib, uint32_t axis, int g,
                                          bool up, float fx, float fy, float fz, float& x, float& y, float& z,
                                          uint32_t& c) {
  const float va = S.val[ia], vb = S.val[ib];
  const float pl = (float)g, pu = (float)(g + 1);
  const float pa = up ? pu : pl, pb = up ? pl : pu;
  const bool sa = near_iso(va), sb = !sa && near_iso(vb), sd = !sa && !sb && near_iso(va - vb);
  const float mu = (0.0f - va) / (vb - va);
  const float t = sa || sd ? pa : (sb ? pb : pa + mu * (pb - pa));
  x = axis == 0u ? t : fx;
  y = axis == 1u ? t : fy;
  z = axis == 2u ? t : fz;
  if constexpr (kColor) {
    const uint32_t c1 = col[ia], c2 = col[ib];
    uint32_t cl = 0u;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int a = (int)((c1 >> (8 * k)) & 255u), b = (int)((c2 >> (8 * k)) & 255u);
      const float v = (float)a + mu * (float)(b - a);
      cl |= (uint32_t)(uint8_t)v << (8 * k);
    }
    c = sa || sd ? c1 : (sb ? c2 : cl);
  }
}

// Vertex (voxel units) of cube edge e of cube (cx, cy, cz): mc_interp from corner edge_a(e) to
// edge_b(e), the edge decoded from the packed tables.
template <bool kColor>
__device__ __forceinline__ void mc_vertex(const McShared& S, const uint32_t* col, int e, int cx, int cy, int cz,
                                          int gx0, int gy0, int gz0, float& x, float& y, float& z, uint32_t& c) {
  const uint32_t info = (uint32_t)(kEdgeLo >> (5 * e)) & 31u;
  const int lx = cx + (int)(info & 1u), ly = cy + (int)((info >> 1) & 1u), lz = cz + (int)((info >> 2) & 1u);
  const uint32_t axis = info >> 3;
  const int il = lx * 81 + ly * 9 + lz, iu = il + (axis == 0u ? 81 : (axis == 1u ? 9 : 1));
  const bool up = ((kEdgeUp >> e) & 1u) != 0u;
  const int g = axis == 0u ? gx0 + lx : (axis == 1u ? gy0 + ly : gz0 + lz);
  mc_interp<kColor>(S, col, up ? iu : il, up ? il : iu, axis, g, up, (float)(gx0 + lx), (float)(gy0 + ly),
                    (float)(gz0 + lz), x, y, z, c);
}

// Vertices (voxel units) of table triangle t of cube (cx, cy, cz) with case ci, and with
// kColor their colours; false for a degenerate triangle (two equal vertices; the
// reference's test at src/tsdf.cu:1055-1058, with its typo corrected).
template <bool kColor>
__device__ __forceinline__ bool mc_triangle(const McShared& S, const uint32_t* col, int ci, int t, int cx, int cy,
                                            int cz, int gx0, int gy0, int gz0, float* P, uint32_t* C) {
  const uint64_t pk = kTriTable[ci];
#pragma unroll
  for (int j = 0; j < 3; ++j)
    mc_vertex<kColor>(S, col, (int)((pk >> (4 * (3 * t + j))) & 15u), cx, cy, cz, gx0, gy0, gz0, P[3 * j],
                      P[3 * j + 1], P[3 * j + 2], C[j]);
  const bool d01 = P[0] == P[3] && P[1] == P[4] && P[2] == P[5];
  const bool d12 = P[3] == P[6] && P[4] == P[7] && P[5] == P[8];
  const bool d20 = P[6] == P[0] && P[7] == P[1] && P[8] == P[2];
  return !(d01 || d12 || d20);
}

// Triangles that k_mc_write will emit for the lane's column, without computing vertices
// where none can coincide. Vertices of one triangle lie on three distinct cube edges, in
// the closed edge segments (the other two coordinates are exact lattice values, the
// interpolation parameter lies in [0, 1]); two such points can only be equal at a corner
// the two edges share, and a vertex lands exactly on corner a of its edge only by a snap
// (|v_a| <= 1e-5, or |v_a - v_b| <= 1e-5 with opposite signs) or by rounding a + mu*(b - a)
// to a, which needs |mu| <= ulp(a)/2 and so |v_a| <= ulp(a) (|v_b - v_a| <= 2); likewise for
// b. A cube whose corners all have |value| >= tiny (tiny above both bounds) therefore emits
// exactly its table count; the others ("near" bits) are counted triangle by triangle.

// Edges of case ci whose two corners lie on opposite sides (bit e): corners 0-3 and 4-7 form
// the z = 0 and z = 1 rings (edges 0-3 and 4-7 join ring neighbours i, i + 1 mod 4), edges
// 8-11 join corner i to corner i + 4.
__host__ __device__ constexpr uint32_t crossed_edges(int ci) {
  const uint32_t lo = (uint32_t)ci & 15u, hi = ((uint32_t)ci >> 4) & 15u;
  const uint32_t rlo = ((lo >> 1) | (lo << 3)) & 15u, rhi = ((hi >> 1) | (hi << 3)) & 15u;
  return (lo ^ rlo) | ((hi ^ rhi) << 4) | ((lo ^ hi) << 8);
}
constexpr bool crossed_edges_ok() {
  for (int ci = 0; ci < 256; ++ci)
    for (int e = 0; e < 12; ++e)
      if ((((ci >> edge_a(e)) ^ (ci >> edge_b(e))) & 1) != (int)((crossed_edges(ci) >> e) & 1u)) return false;
  return true;
}
static_assert(crossed_edges_ok(), "crossed_edges disagrees with the edge table");

// Non-degenerate table triangles of a near cube (case ci at (cx, cy, cz)): vertex_end tells
// on which corner, if any, each vertex lands, so no positions are formed; a triangle is
// degenerate when two of its vertices land on the same corner (the only way two vertices
// on distinct edges coincide), which is what mc_triangle's position test finds. Corners are
// named by their tile index; edges are decoded as in mc_vertex.
__device__ __forceinline__ uint32_t mc_count_near(const McShared& S, int ci, int cx, int cy, int cz, int gx0,
                                                  int gy0, int gz0) {
  const uint64_t pk = kTriTable[ci];
  const int nt = S.cnt[ci];
  uint32_t n = 0;
  for (int t = 0; t < nt; ++t) {
    int cid[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int e = (int)((pk >> (4 * (3 * t + j))) & 15u);
      const uint32_t info = (uint32_t)(kEdgeLo >> (5 * e)) & 31u;
      const int lx = cx + (int)(info & 1u), ly = cy + (int)((info >> 1) & 1u), lz = cz + (int)((info >> 2) & 1u);
      const uint32_t axis = info >> 3;
      const int il = lx * 81 + ly * 9 + lz, iu = il + (axis == 0u ? 81 : (axis == 1u ? 9 : 1));
      const bool up = ((kEdgeUp >> e) & 1u) != 0u;
      const int ia = up ? iu : il, ib = up ? il : iu;
      const int g = axis == 0u ? gx0 + lx : (axis == 1u ? gy0 + ly : gz0 + lz);
      const float pl = (float)g, pu = (float)(g + 1);
      const int end = vertex_end(up ? pu : pl, up ? pl : pu, S.val[ia], S.val[ib]);
      cid[j] = end == 0 ? ia : (end == 1 ? ib : -1 - j);
    }
    n += (cid[0] != cid[1] && cid[1] != cid[2] && cid[2] != cid[0]) ? 1u : 0u;
  }
  return n;
}

// The lane's emitted triangles from its table count n: near cubes are counted triangle by
// triangle.
__device__ __forceinline__ uint32_t mc_count_exact(const McShared& S, uint32_t near, uint32_t n, int gx0, int gy0,
                                                   int gz0) {
  const int lane = mc_lane(), cy = lane >> 3, cz = lane & 7;
  for (uint32_t m = near; m; m &= m - 1u) {
    const int cx = __builtin_ctz(m);
    const int ci = S.cases[cx * 64 + lane];
    n += mc_count_near(S, ci, cx, cy, cz, gx0, gy0, gz0) - (uint32_t)S.cnt[ci];
  }
  return n;
}

// The block's (cube, table triangle) items in order lane, x, t, in rounds of kItemCap: each
// lane lists its own items of the round in LDS, then the wave takes them one per lane.
// Emit receives (valid, cube x/y/z, case, t, P[9], C[3]) per item slot of a 64-item step;
// With kPos, valid is false for empty slots and for degenerate triangles (two equal
// vertices; the reference's test at src/tsdf.cu:1055-1058, with its typo corrected), and
// kColor makes C hold the vertex colours interpolated from the colour tile col (otherwise
// zeros). Without kPos, P and C are not computed and every item is valid.
template <bool kColor, bool kPos, class Emit>
__device__ __forceinline__ void mc_items(McShared& S, McItems& I, const uint32_t* col, uint32_t total, int gx0, int gy0,
                                         int gz0, Emit emit) {
  const int lane = mc_lane();
  const uint32_t my0 = I.lane_off[lane], my1 = I.lane_off[lane + 1];
  for (uint32_t r0 = 0; r0 < total; r0 += kItemCap) {
    const uint32_t r1 = min(total, r0 + (uint32_t)kItemCap);
    if (my0 < r1 && my1 > r0) {
      uint32_t idx = my0;
      for (int cx = 0; cx < 8 && idx < r1; ++cx) {
        const int cube = cx * 64 + lane;
        const int n = S.cnt[S.cases[cube]];
        for (int t = 0; t < n; ++t, ++idx)
          if (idx >= r0 && idx < r1) I.items[idx - r0] = (uint16_t)((cube << 3) | t);
      }
    }
    wave_sync();
    for (uint32_t i0 = r0; i0 < r1; i0 += 64) {
      const uint32_t item = i0 + lane;
      bool valid = false;
      int ci = 0, t = 0, cx = 0, cy = 0, cz = 0;
      float P[9];
      uint32_t C[3] = {0u, 0u, 0u};
      if (item < r1) {
        const int it = I.items[item - r0];
        const int cube = it >> 3;
        t = it & 7;
        cx = cube >> 6;
        cy = (cube >> 3) & 7;
        cz = cube & 7;
        ci = S.cases[cube];
        if constexpr (kPos) valid = mc_triangle<kColor>(S, col, ci, t, cx, cy, cz, gx0, gy0, gz0, P, C);
        else valid = true;
      }
      emit(valid, gx0 + cx, gy0 + cy, gz0 + cz, ci, t, P, C);
    }
    wave_sync();  // the next round overwrites the items
  }
}


// The pool indices of each block's 8 neighbourhood blocks (+x/+y/+z and their combinations,
// index (dx << 2) | (dy << 1) | dz; itself at 0; kNoBlock where absent), one lane per entry,
// so that the extraction passes read one word instead of walking the hash per neighbour.
// With ebits (indexed mesh), each lane also clears 6 of its block's 48 used-edge words.
__global__ __launch_bounds__(256) void k_mc_neighbours(McArgs a) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (size_t)a.n_blocks * 8) return;
  const uint32_t b = (uint32_t)(i >> 3), d = (uint32_t)(i & 7);
  if (a.ebits) {
#pragma unroll
    for (int k = 0; k < 6; ++k) a.ebits[(size_t)b * kEdgeWords + d * 6 + k] = 0u;
  }
  uint32_t nb = b;
  if (d > 0) {
    int bx, by, bz;
    unpack_key(a.block_keys[b], bx, by, bz);
    const int nx = bx + (int)(d >> 2), ny = by + (int)((d >> 1) & 1), nz = bz + (int)(d & 1);
    nb = key_in_range(nx, ny, nz) ? find_block(a.hash, pack_key(nx, ny, nz)) : kNoBlock;
    if (nb >= a.pool_cap) nb = kNoBlock;
  }
  a.nbrs[i] = nb;
}

template <int kStorage>
__global__ __launch_bounds__(64 * kMcWaves) void k_mc_count(McArgs a) {
  __shared__ McShared SS[kMcWaves];
  McShared& S = SS[mc_wave()];
  const int lane = mc_lane();
  const uint32_t b = blockIdx.x * kMcWaves + (uint32_t)mc_wave();
  if (b >= a.n_blocks) return;
  const McCounts counts = mc_fetch_counts();
  int gx0, gy0, gz0;
  if (!mc_stage<kStorage, false, true>(a, b, mc_nbr(a, b), S, nullptr, gx0, gy0, gz0)) {  // no sign change: no triangles
    if (lane == 0) a.counts[b] = 0u;
    return;
  }
  mc_put_counts(S, counts);
  wave_sync();
  uint32_t near;
  uint32_t cnt = mc_classify(S, near);
  if (__ballot(near != 0u)) cnt = mc_count_exact(S, near, cnt, gx0, gy0, gz0);
  cnt = mc_wave_sum(cnt);
  if (lane == 0) a.counts[b] = cnt;
}

// Exclusive scan of counts within chunks of kScanChunk blocks (one workgroup of 256 each,
// 4 counts per lane); chunk totals go to chunk_sums.
__global__ __launch_bounds__(256) void k_scan_chunks(McArgs a) {
  __shared__ uint32_t s_w[4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t c0 = blockIdx.x * kScanChunk;
  uint32_t v[4], sum = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t i = c0 + 4 * tid + k;
    v[k] = i < a.n_blocks ? a.counts[i] : 0u;
    sum += v[k];
  }
  uint32_t x = sum;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)x, off);
    if (lane >= off) x += y;
  }
  if (lane == 63) s_w[wave] = x;
  __syncthreads();
  uint32_t pre = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    if (w < wave) pre += s_w[w];
    tot += s_w[w];
  }
  uint32_t run = pre + x - sum;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t i = c0 + 4 * tid + k;
    if (i < a.n_blocks) a.counts[i] = run;
    run += v[k];
  }
  if (tid == 0) a.chunk_sums[blockIdx.x] = tot;
}

// Exclusive scan of the chunk totals (one workgroup, sequential over groups of 256).
__global__ __launch_bounds__(256) void k_scan_top(McArgs a, uint32_t n_chunks) {
  __shared__ uint32_t s_w[4];
  __shared__ uint32_t s_carry;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid == 0) s_carry = 0;
  __syncthreads();
  for (uint32_t b0 = 0; b0 < n_chunks; b0 += 256) {
    const uint32_t i = b0 + tid;
    const uint32_t v = i < n_chunks ? a.chunk_sums[i] : 0u;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t y = (uint32_t)__shfl_up((int)x, off);
      if (lane >= off) x += y;
    }
    if (lane == 63) s_w[wave] = x;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      if (w < wave) pre += s_w[w];
      tot += s_w[w];
    }
    const uint32_t carry = s_carry;
    if (i < n_chunks) a.chunk_sums[i] = carry + pre + x - v;
    __syncthreads();
    if (tid == 0) s_carry = carry + tot;
    __syncthreads();
  }
  if (tid == 0) *a.total = s_carry;
}

// Exclusive offset of block b after k_scan_chunks / k_scan_top (the total for b == n).
__device__ __forceinline__ uint32_t block_offset(const uint32_t* counts, const uint32_t* chunk_sums, uint32_t b) {
  return chunk_sums[b / kScanChunk] + counts[b];
}
__device__ __forceinline__ uint32_t block_offset(const uint32_t* counts, const uint32_t* chunk_sums, uint32_t b,
                                                 uint32_t n, const uint32_t* total) {
  return b < n ? block_offset(counts, chunk_sums, b) : *total;
}

template <int kStorage, bool kColor>
__global__ __launch_bounds__(64 * kMcWaves) void k_mc_write(McArgs a) {
  __shared__ McShared SS[kMcWaves];
  __shared__ McItems IS[kMcWaves];
  __shared__ uint32_t s_cols[kMcWaves][kColor ? kHalo3 : 1];
  McShared& S = SS[mc_wave()];
  McItems& I = IS[mc_wave()];
  uint32_t* s_col = s_cols[mc_wave()];
  const uint32_t b = blockIdx.x * kMcWaves + (uint32_t)mc_wave();
  if (b >= a.n_blocks) return;
  {
    const uint32_t nbr = mc_nbr(a, b);  // with the offsets: the staging then needs one round trip
    const McCounts counts = mc_fetch_counts();
    uint32_t pos = block_offset(a.counts, a.chunk_sums, b);
    if (block_offset(a.counts, a.chunk_sums, b + 1, a.n_blocks, a.total) == pos) return;  // no triangles
    int gx0, gy0, gz0;
    mc_stage<kStorage, kColor>(a, b, nbr, S, s_col, gx0, gy0, gz0);
    mc_put_counts(S, counts);
    wave_sync();
    uint32_t near;
    const uint32_t total = mc_lane_prefix(I, mc_classify(S, near));
    mc_items<kColor, true>(S, I, s_col, total, gx0, gy0, gz0,
                           [&](bool valid, int x, int y, int z, int ci, int t, const float* P, const uint32_t* C) {
      const uint64_t m = __ballot(valid);
      const uint32_t p = pos + lanes_below(m);
      if (valid && p < a.capacity) {
        float* o = a.tris + (size_t)p * 9;
#pragma unroll
        for (int k = 0; k < 9; ++k) o[k] = P[k] * a.vox;
        if (a.cubes) {
          int32_t* q = a.cubes + (size_t)p * 4;
          q[0] = x; q[1] = y; q[2] = z; q[3] = (ci << 8) | t;
        }
        if (kColor && a.rgb) {
          uint8_t* o8 = a.rgb + (size_t)p * 9;
#pragma unroll
          for (int k = 0; k < 9; ++k) o8[k] = (uint8_t)(C[k / 3] >> (8 * (k % 3)));
        }
      }
      pos += (uint32_t)__popcll(m);
    });
  }
}

// ---------------------------------------------------------------- indexed mesh
// One vertex per used lattice edge (an edge some face lies on), owned by the block that holds
// the edge's lower end point; its id is the owner's vertex offset plus the edge's rank among
// the owner's used edges (bit order axis * 512 + in-block index). Faces are every table
// triangle of every cube with 8 observed corners (zero-area ones included: dropping them
// would open the mesh where the surface passes exactly through a lattice point), in the
// order of k_mc_write; the table's triangles of a case use exactly the edges whose corners
// differ in sign, so those are the edges marked. Vertices are interpolated from the lower
// end point to the upper one (VertexInterp, src/tsdf.cu:1640-1660, with that orientation):
// every face that uses an edge gets the same vertex, and the mesh has no cracks where
// neighbouring cubes would interpolate an edge in opposite directions (the reference welds
// float-identical coordinates instead, src/tsdf.cu:1772-1850).


// Owner slot (index into S.nb) of edge e of cube (cx, cy, cz) and the edge's bit in its
// owner's used-edge words: the lower end point (block-local, 0..8 per axis) decides both.
__device__ __forceinline__ void edge_owner(int e, int cx, int cy, int cz, int& o, int& bit) {
  const uint32_t info = (uint32_t)(kEdgeLo >> (5 * e)) & 31u;
  const int lx = cx + (int)(info & 1u), ly = cy + (int)((info >> 1) & 1u), lz = cz + (int)((info >> 2) & 1u);
  o = ((lx >> 3) << 2) | ((ly >> 3) << 1) | (lz >> 3);
  bit = (int)(info >> 3) * kB3 + (((lx & 7) << 6) | ((ly & 7) << 3) | (lz & 7));
}

// Marks the used edges and counts the faces (table counts) of each block. Edges the block
// owns are gathered in LDS and merged with one atomic OR per word; edges owned by a +x/+y/+z
// neighbour go to its words directly.
template <int kStorage>
__global__ __launch_bounds__(64 * kMcWaves) void k_ix_mark(McArgs a) {
  __shared__ McShared SS[kMcWaves];
  __shared__ uint32_t s_bitss[kMcWaves][kEdgeWords];
  McShared& S = SS[mc_wave()];
  uint32_t* s_bits = s_bitss[mc_wave()];
  const int lane = mc_lane();
  const uint32_t b = blockIdx.x * kMcWaves + (uint32_t)mc_wave();
  if (b >= a.n_blocks) return;
  const McCounts counts = mc_fetch_counts();
  {
    if (lane < kEdgeWords) s_bits[lane] = 0u;
    int gx0, gy0, gz0;
    if (!mc_stage<kStorage, false>(a, b, mc_nbr(a, b), S, nullptr, gx0, gy0, gz0)) {  // no sign change: no faces, no edges
      if (lane == 0) a.counts[b] = 0u;
      return;
    }
    mc_put_counts(S, counts);
    wave_sync();
    uint32_t near;
    const uint32_t total = mc_wave_sum(mc_classify(S, near));
    const int cy = lane >> 3, cz = lane & 7;
    for (int cx = 0; cx < 8; ++cx) {
      const int ci = S.cases[cx * 64 + lane];
      for (uint32_t m = ci ? crossed_edges(ci) : 0u; m; m &= m - 1u) {
        const int e = __builtin_ctz(m);
        int o, bit;
        edge_owner(e, cx, cy, cz, o, bit);
        if (o == 0) atomicOr(&s_bits[bit >> 5], 1u << (bit & 31));
        else atomicOr(&a.ebits[(size_t)S.nb[o] * kEdgeWords + (bit >> 5)], 1u << (bit & 31));
      }
    }
    if (lane == 0) a.counts[b] = total;
    wave_sync();
    if (lane < kEdgeWords && s_bits[lane]) atomicOr(&a.ebits[(size_t)b * kEdgeWords + lane], s_bits[lane]);
  }
}

// Per block: popcount prefixes of the 48 used-edge words and the block's vertex count (one
// wave per block, lanes 0..47 one word each).
__global__ __launch_bounds__(256) void k_ix_count(McArgs a) {
  const int lane = threadIdx.x & 63;
  for (uint32_t b = blockIdx.x * 4u + (threadIdx.x >> 6); b < a.n_blocks; b += gridDim.x * 4u) {
    const uint32_t w = lane < kEdgeWords ? a.ebits[(size_t)b * kEdgeWords + lane] : 0u;
    const uint32_t c = (uint32_t)__popc(w);
    uint32_t x = c;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t y = (uint32_t)__shfl_up((int)x, off);
      if (lane >= off) x += y;
    }
    if (lane < kEdgeWords) a.epre[(size_t)b * kEdgeWords + lane] = (uint16_t)(x - c);
    if (lane == kEdgeWords - 1) a.vcounts[b] = x;
  }
}

// Faces (vertex ids of each table triangle's corners) and the vertices each block owns.
template <int kStorage, bool kColor>
__global__ __launch_bounds__(64 * kMcWaves) void k_ix_write(McArgs a) {
  __shared__ McShared SS[kMcWaves];
  __shared__ McItems IS[kMcWaves];
  __shared__ uint32_t s_cols[kMcWaves][kColor ? kHalo3 : 1];
  __shared__ uint32_t s_vbases[kMcWaves][8];
  McShared& S = SS[mc_wave()];
  McItems& I = IS[mc_wave()];
  uint32_t* s_col = s_cols[mc_wave()];
  uint32_t* s_vbase = s_vbases[mc_wave()];
  const int lane = mc_lane();
  const uint32_t b = blockIdx.x * kMcWaves + (uint32_t)mc_wave();
  if (b >= a.n_blocks) return;
  {
    uint32_t pos = block_offset(a.counts, a.chunk_sums, b);
    if (block_offset(a.counts, a.chunk_sums, b + 1, a.n_blocks, a.total) == pos &&
        block_offset(a.vcounts, a.vchunks, b + 1, a.n_blocks, a.vtotal) == block_offset(a.vcounts, a.vchunks, b))
      return;  // no faces and no vertices of its own
    const McCounts counts = mc_fetch_counts();
    int gx0, gy0, gz0;
    mc_stage<kStorage, kColor>(a, b, mc_nbr(a, b), S, s_col, gx0, gy0, gz0);
    mc_put_counts(S, counts);
    wave_sync();
    if (lane < 8) {
      const uint32_t nb = S.nb[lane];
      s_vbase[lane] = nb != kNoBlock ? a.vchunks[nb / kScanChunk] + a.vcounts[nb] : 0u;
    }
    uint32_t near;
    const uint32_t total = mc_lane_prefix(I, mc_classify(S, near));  // ends with a barrier
    mc_items<false, false>(S, I, nullptr, total, gx0, gy0, gz0,
                           [&](bool valid, int x, int y, int z, int ci, int t, const float*, const uint32_t*) {
      const uint64_t m = __ballot(valid);
      const uint32_t p = pos + lanes_below(m);
      if (valid && p < a.capacity) {
        const uint64_t pk = kTriTable[ci];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          int o, bit;
          edge_owner((int)((pk >> (4 * (3 * t + j))) & 15u), x - gx0, y - gy0, z - gz0, o, bit);
          const size_t wi = (size_t)S.nb[o] * kEdgeWords + (bit >> 5);
          const uint32_t below = a.ebits[wi] & ((1u << (bit & 31)) - 1u);
          a.faces[(size_t)p * 3 + j] = (int32_t)(s_vbase[o] + a.epre[wi] + (uint32_t)__popc(below));
        }
      }
      pos += (uint32_t)__popcll(m);
    });
    // the block's own vertices: lane l < 48 walks the set bits of word l
    if (lane < kEdgeWords) {
      uint32_t w = a.ebits[(size_t)b * kEdgeWords + lane];
      uint32_t id = s_vbase[0] + a.epre[(size_t)b * kEdgeWords + lane];
      while (w) {
        const int bit = lane * 32 + __builtin_ctz(w);
        w &= w - 1u;
        const int axis = bit / kB3, vi = bit % kB3;
        const int lx = vi >> 6, ly = (vi >> 3) & 7, lz = vi & 7;
        const int ux = lx + (axis == 0), uy = ly + (axis == 1), uz = lz + (axis == 2);
        const int i0 = lx * 81 + ly * 9 + lz, i1 = ux * 81 + uy * 9 + uz;
        float px, py, pz;
        uint32_t c = 0u;
        const int g = axis == 0 ? gx0 + lx : (axis == 1 ? gy0 + ly : gz0 + lz);
        mc_interp<kColor>(S, s_col, i0, i1, (uint32_t)axis, g, false, (float)(gx0 + lx), (float)(gy0 + ly),
                          (float)(gz0 + lz), px, py, pz, c);
        if (id < a.vertex_cap) {
          float* q = a.vertices + (size_t)id * 3;
          q[0] = px * a.vox;
          q[1] = py * a.vox;
          q[2] = pz * a.vox;
          if (kColor) {
            uint8_t* q8 = a.vrgb + (size_t)id * 3;
            q8[0] = (uint8_t)c;
            q8[1] = (uint8_t)(c >> 8);
            q8[2] = (uint8_t)(c >> 16);
          }
        }
        ++id;
      }
    }
  }
}

// ============================================================================ streaming
// Out-of-core streaming (the reference's streamOutGPU2CPU / streamInCPU2GPU, src/tsdf.cu:
// 277-596, re-designed: whole 8^3 blocks move, not 5^3-block chunks, the pool stays dense,
// and the kernels read and write the pinned host store directly, so each block crosses the
// link once with no staging copy). One wave per block in the copy kernels, 64 lanes
// striding over its words.

__device__ __forceinline__ uint32_t* slab_words(const StreamArgs& a, uint32_t slot) {
  return a.slab_words[slot >> kSlabShift] + (size_t)(slot & (kSlabBlocks - 1u)) * a.wpb;
}
__device__ __forceinline__ uint32_t* slab_cols(const StreamArgs& a, uint32_t slot) {
  return a.slab_cols[slot >> kSlabShift] + (size_t)(slot & (kSlabBlocks - 1u)) * kB3;
}

__global__ __launch_bounds__(256) void k_st_out(StreamArgs a) {
  const int lane = threadIdx.x & 63;
  for (uint32_t i = blockIdx.x * 4u + (threadIdx.x >> 6); i < a.n; i += gridDim.x * 4u) {
    const size_t b = a.src[i];
    uint32_t* w = slab_words(a, a.slots[i]);
    for (uint32_t k = lane; k < a.wpb; k += 64) w[k] = a.pool[b * a.wpb + k];
    if (a.colors) {
      uint32_t* c = slab_cols(a, a.slots[i]);
      for (uint32_t k = lane; k < kB3; k += 64) c[k] = a.colors[b * kB3 + k];
    }
  }
}

__global__ __launch_bounds__(256) void k_st_move(StreamArgs a) {
  const int lane = threadIdx.x & 63;
  for (uint32_t i = blockIdx.x * 4u + (threadIdx.x >> 6); i < a.n; i += gridDim.x * 4u) {
    const size_t s = a.src[i], d = a.dst[i];
    for (uint32_t k = lane; k < a.wpb; k += 64) a.pool[d * a.wpb + k] = a.pool[s * a.wpb + k];
    if (a.colors)
      for (uint32_t k = lane; k < kB3; k += 64) a.colors[d * kB3 + k] = a.colors[s * kB3 + k];
    if (lane == 0) a.block_keys[d] = a.block_keys[s];
  }
}

__global__ __launch_bounds__(256) void k_st_rehash(StreamArgs a) {
  const uint32_t b = blockIdx.x * 256u + threadIdx.x;
  if (b >= a.n) return;
  bool inserted = false;
  const int64_t slot = find_or_insert(a.hash, a.block_keys[b], inserted);
  if (slot < 0) {
    raise_flag(a.ctl, kFlagHashFull);
  } else {
    a.hash.vals[slot] = (unsigned long long)b | ((unsigned long long)a.birth << 32);
    a.block_info[b] = (unsigned long long)slot | ((unsigned long long)a.birth << 32);
  }
}

// One lane per staged block: find or insert its key; new keys take pool blocks (one counter
// atomic per wave), resident keys are marked for merging.
__global__ __launch_bounds__(256) void k_st_insert(StreamArgs a) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool act = i < a.n;
  bool inserted = false;
  int64_t slot = -1;
  if (act) {
    slot = find_or_insert(a.hash, a.keys[i], inserted);
    if (slot < 0) raise_flag(a.ctl, kFlagHashFull);
  }
  const uint64_t im = __ballot(inserted);
  const int leader = im ? first_lane(im) : 0;
  uint32_t pbase = 0;
  if (im && lane == leader) pbase = atomicAdd(&a.ctl->pool_count, (uint32_t)__popcll(im));
  if (im) pbase = read_lane_u32(pbase, leader);
  if (!act) return;
  uint32_t placed = kNoBlock;
  if (inserted) {
    const uint32_t idx = pbase + lanes_below(im);
    uint32_t stored = kNoBlock;
    if (idx < a.pool_cap) {
      a.block_keys[idx] = a.keys[i];
      a.block_info[idx] = (unsigned long long)slot | ((unsigned long long)a.birth << 32);
      stored = placed = idx;
    } else {
      raise_flag(a.ctl, kFlagPoolFull);
    }
    a.hash.vals[slot] = (unsigned long long)stored | ((unsigned long long)a.birth << 32);
  } else if (slot >= 0) {
    const uint32_t v = (uint32_t)a.hash.vals[slot];
    if (v < a.pool_cap) placed = v | kMergeBit;
  }
  a.placed[i] = placed;
}

// Host-store blocks into the pool: copied, or merged voxel by voxel into a resident block
// of the same key (the observation histories combine): weights add (capped), the TSDF value
// is the weighted mean (s_r w_r + s_s w_s) / (w_r + w_s), colours per channel
// (c_r w_r + c_s w_s) / (w_r + w_s) truncated to 8 bits; a side with weight 0 contributes
// nothing (the other side is kept bit for bit).
template <int kStorage>
__global__ __launch_bounds__(256) void k_st_place(StreamArgs a) {
  using IO = VoxelIO<kStorage>;
  using Word = typename IO::Word;
  const int lane = threadIdx.x & 63;
  for (uint32_t i = blockIdx.x * 4u + (threadIdx.x >> 6); i < a.n; i += gridDim.x * 4u) {
    const uint32_t p = a.placed[i];
    if (p == kNoBlock) continue;
    const size_t d = p & ~kMergeBit;
    const uint32_t* hw = slab_words(a, a.slots[i]);
    const uint32_t* hc = a.colors ? slab_cols(a, a.slots[i]) : nullptr;
    if (!(p & kMergeBit)) {
      for (uint32_t k = lane; k < a.wpb; k += 64) a.pool[d * a.wpb + k] = hw[k];
      if (a.colors)
        for (uint32_t k = lane; k < kB3; k += 64) a.colors[d * kB3 + k] = hc[k];
      continue;
    }
    Word* rw = reinterpret_cast<Word*>(a.pool) + d * kB3;
    const Word* sw = reinterpret_cast<const Word*>(hw);
    for (uint32_t k = lane; k < kB3; k += 64) {
      float sr, wr, ss, ws;
      const Word r = rw[k], q = sw[k];
      IO::unpack(r, sr, wr);
      IO::unpack(q, ss, ws);
      if (!(ws > 0.0f)) continue;
      uint32_t* col = a.colors ? a.colors + d * kB3 + k : nullptr;
      const uint32_t cs = a.colors ? hc[k] : 0u;
      if (!(wr > 0.0f)) {
        rw[k] = q;
        if (col) *col = cs;
        continue;
      }
      const float wsum = wr + ws;
      const float pr = sr * wr, ps = ss * ws;
      const float sv = div_rn(pr + ps, wsum);
      const float wv = (a.cap > 0.0f && wsum > a.cap) ? a.cap : wsum;
      rw[k] = IO::pack(sv, wv);
      if (col) {
        const uint32_t cr = *col;
        uint32_t out = 0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const float xr = (float)((cr >> (8 * c)) & 255u) * wr, xs = (float)((cs >> (8 * c)) & 255u) * ws;
          out |= (uint32_t)(uint8_t)div_rn(xr + xs, wsum) << (8 * c);
        }
        *col = out;
      }
    }
  }
}

// ============================================================================ queries
// Trilinear sample of the normalised TSDF at world point q; false unless all 8 lattice
// corners are allocated and observed. The lattice position q/vox and its cell fraction are
// formed in FP64 (an FP32 q/vox at |q| ~ 3 m is off by ~2e-5 of a voxel), the blend runs in
// FP32. vid/w receive the pool voxel ids and weights; g the gradient w.r.t. q.
template <int kStorage>
__device__ __forceinline__ bool sample_sdf(const QueryArgs& q, double px, double py, double pz,
                                           float& sdf, float g[3], uint32_t vid[8], float w[8]) {
  using IO = VoxelIO<kStorage>;
  using Word = typename IO::Word;
  const double vox = (double)q.vox;
  const double gxd = px / vox, gyd = py / vox, gzd = pz / vox;
  if (!(fabs(gxd) < 8.0e6 && fabs(gyd) < 8.0e6 && fabs(gzd) < 8.0e6)) return false;
  const double flx = floor(gxd), fly = floor(gyd), flz = floor(gzd);
  const int ix = (int)flx, iy = (int)fly, iz = (int)flz;
  const float fx = (float)(gxd - flx), fy = (float)(gyd - fly), fz = (float)(gzd - flz);
  const Word* voxw = reinterpret_cast<const Word*>(q.voxels);
  // the corners' blocks first (one lookup per distinct block), then all 8 voxel loads at
  // once, then the weight tests: a test after each load would make every load wait for the
  // one before it
  uint64_t last_key = kEmptyKey;
  uint32_t blk = kNoBlock;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int cx = ix + (k >> 2), cy = iy + ((k >> 1) & 1), cz = iz + (k & 1);
    const uint64_t key = pack_key(cx >> 3, cy >> 3, cz >> 3);
    if (key != last_key) {
      last_key = key;
      blk = find_block(q.hash, key);
      if (blk >= q.pool_cap) return false;
    }
    vid[k] = blk * kB3 + (uint32_t)((((cx & 7) * 8) + (cy & 7)) * 8 + (cz & 7));
  }
  Word wd[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) wd[k] = voxw[vid[k]];
  float s[8];
  bool seen = true;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    float wv;
    IO::unpack(wd[k], s[k], wv);
    seen = seen & (wv > 0.0f);
  }
  if (!seen) return false;
  trilinear_weights(fx, fy, fz, w);
  float acc = 0.0f;
#pragma unroll
  for (int k = 0; k < 8; ++k) acc = acc + w[k] * s[k];
  sdf = acc;
  trilinear_dfrac(fx, fy, fz, s, g);
  g[0] = g[0] / q.vox;
  g[1] = g[1] / q.vox;
  g[2] = g[2] / q.vox;
  return true;
}

template <int kStorage>
__global__ __launch_bounds__(256) void k_query(QueryArgs q, const float* pts, int64_t n, float* sdf,
                                               uint8_t* valid) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    float r = 0.0f, g[3], w[8];
    uint32_t vid[8];
    const bool ok = sample_sdf<kStorage>(q, (double)pts[3 * i], (double)pts[3 * i + 1], (double)pts[3 * i + 2], r,
                                         g, vid, w);
    if (sdf) sdf[i] = ok ? r : 0.0f;
    if (valid) valid[i] = ok ? 1 : 0;
  }
}

// q = R p + t in FP64 from a row-major 3x4 pose (float or double entries).
template <class P>
__device__ __forceinline__ void pose_apply_d(const P* T, float x, float y, float z, double& qx, double& qy, double& qz) {
  const double X = x, Y = y, Z = z;
  qx = ((X * (double)T[0] + Y * (double)T[1]) + Z * (double)T[2]) + (double)T[3];
  qy = ((X * (double)T[4] + Y * (double)T[5]) + Z * (double)T[6]) + (double)T[7];
  qz = ((X * (double)T[8] + Y * (double)T[9]) + Z * (double)T[10]) + (double)T[11];
}

__device__ __forceinline__ double shfl_xor_d(double v, int m) {
  const long long b = __double_as_longlong(v);
  const int lo = __shfl_xor((int)(b & 0xffffffffll), m);
  const int hi = __shfl_xor((int)(b >> 32), m);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sums acc[0..N) over the workgroup (256 lanes) into out[0..N); fixed tree, deterministic.
template <int N>
__device__ __forceinline__ void wg_reduce_store(double (&acc)[N], double* out) {
  __shared__ double s_part[4][N];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    double v = acc[k];
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += shfl_xor_d(v, m);
    acc[k] = v;
  }
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < N; ++k) s_part[wave][k] = acc[k];
  }
  __syncthreads();
  if (threadIdx.x < N) {
    const int k = threadIdx.x;
    out[k] = ((s_part[0][k] + s_part[1][k]) + s_part[2][k]) + s_part[3][k];
  }
  __syncthreads();
}

// Sums acc[0..N) (N <= 32) over the workgroup (256 lanes) into out[0..N) by recursive halving
// within each wave: at lane distance m = 32, 16, 8, 4, 2 every lane sends the half of its
// values its partner (lane ^ m) keeps and adds the partner's half of its own, so the values
// per lane halve each step (31 double exchanges instead of 6 per value); a last exchange at
// distance 1 completes the wave's sums, which lane 2 i + 1 ... hold for value i; the four waves
// are added in order. Fixed tree, deterministic.
template <int N>
__device__ __forceinline__ void wg_reduce_store_rs(const double (&acc)[N], double* out) {
  static_assert(N <= 32, "at most 32 values");
  __shared__ double s_part[4][32];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double v[32];
#pragma unroll
  for (int k = 0; k < 32; ++k) v[k] = k < N ? acc[k] : 0.0;
#pragma unroll
  for (int h = 16; h >= 1; h >>= 1) {  // h values per lane after this step; distance 2h
    const bool upper = (lane & (2 * h)) != 0;  // this lane keeps the upper half
#pragma unroll
    for (int i = 0; i < h; ++i) {
      const double send = upper ? v[i] : v[i + h];
      const double keep = upper ? v[i + h] : v[i];
      v[i] = keep + shfl_xor_d(send, 2 * h);
    }
  }
  const double total = v[0] + shfl_xor_d(v[0], 1);
  // value index of this lane: the halves it kept, from bit 5 (most significant) to bit 1
  const int idx = ((lane >> 5) & 1) * 16 + ((lane >> 4) & 1) * 8 + ((lane >> 3) & 1) * 4 + ((lane >> 2) & 1) * 2 +
                  ((lane >> 1) & 1);
  if ((lane & 1) == 0) s_part[wave][idx] = total;
  __syncthreads();
  if (threadIdx.x < N) {
    const int k = threadIdx.x;
    out[k] = ((s_part[0][k] + s_part[1][k]) + s_part[2][k]) + s_part[3][k];
  }
  __syncthreads();
}

// Residuals r_i = sdf(R p_i + t) and their derivatives with respect to a left se(3)
// perturbation (rho, phi): dr/drho = grad, dr/dphi = q x grad.
template <int kStorage>
__global__ __launch_bounds__(256) void k_backward(QueryArgs q, const float* pts, int64_t n, const float* T12,
                                                  const float* grad_r, float* r_out, uint8_t* valid, float* J,
                                                  double* partials, float* voxel_grad) {
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double qx, qy, qz;
    pose_apply_d(T12, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], qx, qy, qz);
    float r = 0.0f, g[3] = {0, 0, 0}, w[8];
    uint32_t vid[8];
    const bool ok = sample_sdf<kStorage>(q, qx, qy, qz, r, g, vid, w);
    float j[6] = {0, 0, 0, 0, 0, 0};
    if (ok) {
      const double gx = g[0], gy = g[1], gz = g[2];
      j[0] = g[0]; j[1] = g[1]; j[2] = g[2];
      j[3] = (float)(qy * gz - qz * gy);
      j[4] = (float)(qz * gx - qx * gz);
      j[5] = (float)(qx * gy - qy * gx);
    }
    if (r_out) r_out[i] = ok ? r : 0.0f;
    if (valid) valid[i] = ok ? 1 : 0;
    if (J) {
#pragma unroll
      for (int k = 0; k < 6; ++k) J[6 * i + k] = j[k];
    }
    const float up = grad_r ? grad_r[i] : 0.0f;
    if (ok && grad_r) {
#pragma unroll
      for (int k = 0; k < 6; ++k) acc[k] += (double)up * (double)j[k];
      if (voxel_grad) {
#pragma unroll
        for (int k = 0; k < 8; ++k) atomicAdd(voxel_grad + vid[k], up * w[k]);
      }
    }
  }
  if (partials) wg_reduce_store<6>(acc, partials + 6 * (size_t)blockIdx.x);
}

__global__ __launch_bounds__(64) void k_reduce6(const double* partials, int grid, double* out6) {
  if (threadIdx.x < 6) {
    double s = 0.0;
    for (int g = 0; g < grid; ++g) s += partials[6 * (size_t)g + threadIdx.x];
    out6[threadIdx.x] = s;
  }
}

// ============================================================================ Gauss-Newton
// Huber loss of a scaled residual e and its iteratively reweighted least-squares weight
// psi(e)/e (kHuber, vh_internal.h).
__device__ __forceinline__ double huber_rho(double e) {
  const double ae = fabs(e);
  return ae <= kHuber ? 0.5 * e * e : kHuber * (ae - 0.5 * kHuber);
}
__device__ __forceinline__ double huber_weight(double e) {
  const double ae = fabs(e);
  return ae <= kHuber ? 1.0 : kHuber / ae;
}

// Per point (camera frame p, depth z = p.z): s = mu / sigma(z); with a valid sample r at
// q = R p + t and e = s r, the point adds W J J^T and W r J with W = huber_weight(e) s^2 (one
// Gauss-Newton step of the iteratively reweighted problem), r^2, and rho(e) to the cost; a
// point without a valid sample adds rho(s), the cost of a saturated residual r = 1.
__device__ __forceinline__ void gn_solve_wg(const double* partials, int grid, const GnState* st_in, GnState* st_out, double* s_cand,
                            double* cand_out, int final_step, double* stats_row);
__device__ __forceinline__ void gn_step(const double* tot, GnState& S, double* s_cand, double* cand_out, int final_step,
                                        double* stats_row);

// One evaluation. With a previous evaluation (prev_grid > 0) every workgroup first takes its
// Levenberg-Marquardt step itself (gn_solve_wg over the previous partials: the same sums in
// the same order, so every workgroup gets the same candidate; workgroup 0 stores the state,
// the candidate and the stats row), which saves a launch per evaluation and needs no ordering
// between workgroups; otherwise the candidate is a.pose.
// Register budget: 3 waves per SIMD (154 VGPRs, no spills; the 31 FP64 sums and the robust
// weighting of a point need more than the 128 of 4 waves, which spilled into the point loop).
// The grid is at most 512 workgroups of 4 waves, so every wave of a launch is resident at once.
template <int kStorage>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_gn_accum(GnArgs a) {
  __shared__ double s_pose[12];
  if (a.prev_grid > 0) {
    const bool w0 = blockIdx.x == 0;
    gn_solve_wg(a.prev_partials, a.prev_grid, a.st_in, w0 ? a.st_out : nullptr, s_pose, w0 ? a.cand_out : nullptr, 0,
                w0 ? a.stats_row : nullptr);
  } else {
    if (threadIdx.x < 12) s_pose[threadIdx.x] = a.pose[threadIdx.x];
    __syncthreads();
  }
  double acc[kGnAcc];
#pragma unroll
  for (int k = 0; k < kGnAcc; ++k) acc[k] = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * blockDim.x) {
    const float px = a.pts[3 * i], py = a.pts[3 * i + 1], pz = a.pts[3 * i + 2];
    if (!(isfinite(px) && isfinite(py) && isfinite(pz))) continue;  // no depth at this pixel
    acc[29] += 1.0;
    const double zt = (double)pz - 0.4;
    const double sc = (double)a.mu / (0.0012 + 0.0019 * (zt * zt));
    double qx, qy, qz;
    pose_apply_d(s_pose, px, py, pz, qx, qy, qz);
    float r, g[3], w[8];
    uint32_t vid[8];
    if (!sample_sdf<kStorage>(a.q, qx, qy, qz, r, g, vid, w)) {
      acc[30] += huber_rho(sc);
      continue;
    }
    const double rd = (double)r, e = sc * rd;
    const double W = huber_weight(e) * sc * sc;
    const double gx = g[0], gy = g[1], gz = g[2];
    const double j[6] = {gx, gy, gz, qy * gz - qz * gy, qz * gx - qx * gz, qx * gy - qy * gx};
    double wj[6];
#pragma unroll
    for (int p = 0; p < 6; ++p) wj[p] = W * j[p];
    int m = 0;
#pragma unroll
    for (int p = 0; p < 6; ++p)
#pragma unroll
      for (int c = p; c < 6; ++c) acc[m++] += wj[p] * j[c];
#pragma unroll
    for (int p = 0; p < 6; ++p) acc[21 + p] += wj[p] * rd;
    acc[27] += rd * rd;
    acc[28] += 1.0;
    acc[30] += huber_rho(e);
  }
  wg_reduce_store_rs<kGnAcc>(acc, a.partials + kGnAcc * (size_t)blockIdx.x);
}

// One workgroup, FP64: deterministic sum of the partials at the candidate pose, then a
// Levenberg-Marquardt decision and step.
//   cost(candidate) = the robust cost of k_gn_accum (every point with finite coordinates
//   counts, an invalid one as a saturated residual), so costs at different poses compare;
//   skipped points count neither way. A tracked frame scales that term (and its A, b) by
//   n_eff / (points considered) and adds its motion prior (motion_prior). The candidate is
//   accepted if its cost does not exceed
//   the accepted one (lambda /= 3) and rejected otherwise (lambda *= 4). Unless `final`, the
//   next candidate is exp(delta^) T_acc with (A + lambda diag(A)) delta = -b, A = J^T W J and
//   b = J^T W r at the accepted pose, delta = (rho, phi), T <- exp(delta^) T.
//   stats_row: [n_valid, rms, |delta|, lambda, accepted, accepted pose (12)].
__device__ __forceinline__ void se3_left_update(const double* x, const double* pose, double* out) {
  const double rho[3] = {x[0], x[1], x[2]}, ph[3] = {x[3], x[4], x[5]};
  const double th2 = ph[0] * ph[0] + ph[1] * ph[1] + ph[2] * ph[2];
  const double th = sqrt(th2);
  double A1, B1, C1;  // sin/th, (1-cos)/th^2, (th - sin)/th^3
  if (th < 1e-5) {
    A1 = 1.0 - th2 / 6.0;
    B1 = 0.5 - th2 / 24.0;
    C1 = 1.0 / 6.0 - th2 / 120.0;
  } else {
    A1 = sin(th) / th;
    B1 = (1.0 - cos(th)) / th2;
    C1 = (th - sin(th)) / (th2 * th);
  }
  const double K[3][3] = {{0.0, -ph[2], ph[1]}, {ph[2], 0.0, -ph[0]}, {-ph[1], ph[0], 0.0}};
  double K2[3][3], R[3][3], V[3][3];
  for (int p = 0; p < 3; ++p)
    for (int c = 0; c < 3; ++c) {
      double acc = 0.0;
      for (int k = 0; k < 3; ++k) acc += K[p][k] * K[k][c];
      K2[p][c] = acc;
    }
  for (int p = 0; p < 3; ++p)
    for (int c = 0; c < 3; ++c) {
      const double I = p == c ? 1.0 : 0.0;
      R[p][c] = I + A1 * K[p][c] + B1 * K2[p][c];
      V[p][c] = I + B1 * K[p][c] + C1 * K2[p][c];
    }
  for (int p = 0; p < 3; ++p) {
    for (int c = 0; c < 3; ++c) out[4 * p + c] = R[p][0] * pose[c] + R[p][1] * pose[4 + c] + R[p][2] * pose[8 + c];
    out[4 * p + 3] = R[p][0] * pose[3] + R[p][1] * pose[7] + R[p][2] * pose[11] +
                     (V[p][0] * rho[0] + V[p][1] * rho[1] + V[p][2] * rho[2]);
  }
}

// Motion prior of a tracked frame at pose T (row-major 3x4): xi = log(S^-1 T) with S the
// frame's starting pose, translation first; the angle from atan2(|w|/2, (tr - 1)/2) (no loss of
// precision near zero), phi = theta / (2 sin theta) w, rho = V(phi)^-1 D_t with
// V^-1 = I - K/2 + (1/theta^2 - (1 + cos theta) / (2 theta sin theta)) K^2 (series below 1e-5).
// Its Jacobian for T <- exp(delta^) T is taken as M = Ad(T^-1) (the exact one is
// Jr^-1(xi) M; xi is a few noise standard deviations), so with Lam = diag(lam_t I3, lam_r I3) it
// adds M^T Lam M to J^T W J, M^T Lam xi to J^T W r and xi^T Lam xi / 2 to the cost.
__device__ __forceinline__ void motion_prior(const GnState& S, double* Ab, double& cost) {
  const double* P = S.start;
  const double* T = S.cand;
  double D[3][3], dt[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 3; ++j) D[i][j] = P[i] * T[j] + P[4 + i] * T[4 + j] + P[8 + i] * T[8 + j];
    dt[i] = P[i] * (T[3] - P[3]) + P[4 + i] * (T[7] - P[7]) + P[8 + i] * (T[11] - P[11]);
  }
  const double wv[3] = {D[2][1] - D[1][2], D[0][2] - D[2][0], D[1][0] - D[0][1]};
  const double sn = 0.5 * sqrt(wv[0] * wv[0] + wv[1] * wv[1] + wv[2] * wv[2]);
  const double cs = 0.5 * (D[0][0] + D[1][1] + D[2][2] - 1.0);
  const double th = atan2(sn, cs);
  double f, c2;
  if (th < 1e-5) {
    f = 0.5 + th * th / 12.0;
    c2 = 1.0 / 12.0 + th * th / 720.0;
  } else {
    f = th / (2.0 * sn);
    c2 = 1.0 / (th * th) - (1.0 + cs) / (2.0 * th * sn);
  }
  const double ph[3] = {f * wv[0], f * wv[1], f * wv[2]};
  const double K[3][3] = {{0.0, -ph[2], ph[1]}, {ph[2], 0.0, -ph[0]}, {-ph[1], ph[0], 0.0}};
  double xi[6];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double a = dt[i] - 0.5 * (K[i][0] * dt[0] + K[i][1] * dt[1] + K[i][2] * dt[2]);
#pragma unroll
    for (int k = 0; k < 3; ++k) a += c2 * (K[i][0] * K[0][k] + K[i][1] * K[1][k] + K[i][2] * K[2][k]) * dt[k];
    xi[i] = a;
    xi[3 + i] = ph[i];
  }
  // With M = Ad(T^-1) = [[R^T, [p]x R^T], [0, R^T]], p = -R^T t, and R R^T = I:
  // M^T Lam M = [[lt I, -lt [t]x], [lt [t]x, lt (|t|^2 I - t t^T) + lr I]] and
  // M^T Lam xi = (g, t x g + lr R xi_phi) with g = lt R xi_rho.
  const double lt = S.lam_t, lr = S.lam_r;
  const double tv[3] = {T[3], T[7], T[11]};
  const double tt = tv[0] * tv[0] + tv[1] * tv[1] + tv[2] * tv[2];
  const double Kt[3][3] = {{0.0, -tv[2], tv[1]}, {tv[2], 0.0, -tv[0]}, {-tv[1], tv[0], 0.0}};
  cost += 0.5 * (lt * (xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2]) +
                 lr * (xi[3] * xi[3] + xi[4] * xi[4] + xi[5] * xi[5]));
#pragma unroll
  for (int a = 0, m = 0; a < 6; ++a)
#pragma unroll
    for (int b = a; b < 6; ++b, ++m) {
      double v;
      if (a < 3 && b < 3) v = a == b ? lt : 0.0;
      else if (a < 3) v = -lt * Kt[a][b - 3];
      else v = lt * ((a == b ? tt : 0.0) - tv[a - 3] * tv[b - 3]) + (a == b ? lr : 0.0);
      Ab[m] += v;
    }
  double g[3], gr[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    g[i] = lt * (T[4 * i] * xi[0] + T[4 * i + 1] * xi[1] + T[4 * i + 2] * xi[2]);
    gr[i] = lr * (T[4 * i] * xi[3] + T[4 * i + 1] * xi[4] + T[4 * i + 2] * xi[5]);
  }
  Ab[21] += g[0];
  Ab[22] += g[1];
  Ab[23] += g[2];
  Ab[24] += (tv[1] * g[2] - tv[2] * g[1]) + gr[0];
  Ab[25] += (tv[2] * g[0] - tv[0] * g[2]) + gr[1];
  Ab[26] += (tv[0] * g[1] - tv[1] * g[0]) + gr[2];
}

// The partials are summed by 248 lanes at once: lane (v, j) (v < kGnAcc, j < 8) sums
// accumulator v over the workgroups g = j, j + 8, ... in order, then lanes v < kGnAcc add the 8
// sums in a fixed tree, so every workgroup gets the same sums in the same order. Meanwhile
// lanes below the state's size stage the solver state in LDS, so the step (one lane) reads it
// from there. All 256 lanes call it; the new candidate is in s_cand on return. st_out,
// cand_out and stats_row (global, each may be null) receive the new state, the candidate and
// the evaluation's row.
__device__ __forceinline__ void gn_solve_wg(const double* partials, int grid, const GnState* st_in, GnState* st_out, double* s_cand,
                            double* cand_out, int final_step, double* stats_row) {
  constexpr int kWays = 8, kStateWords = (int)(sizeof(GnState) / sizeof(double));
  static_assert(kWays * kGnAcc <= 256 && kStateWords <= 256 && sizeof(GnState) % sizeof(double) == 0, "layout");
  __shared__ double s_sum[kWays][kGnAcc];
  __shared__ double tot[kGnAcc];
  __shared__ double s_st[kStateWords];
  const int t = threadIdx.x;
  if (t < kStateWords) s_st[t] = reinterpret_cast<const double*>(st_in)[t];
  if (t < kWays * kGnAcc) {
    const int v = t % kGnAcc, j = t / kGnAcc;
    // in order g = j, j + 8, ...; the loads of 8 terms in flight together, then their sum
    double acc = 0.0;
    int g = j;
    for (; g + 7 * kWays < grid; g += 8 * kWays) {
      double p[8];
#pragma unroll
      for (int r = 0; r < 8; ++r) p[r] = partials[kGnAcc * (size_t)(g + r * kWays) + v];
#pragma unroll
      for (int r = 0; r < 8; ++r) acc += p[r];
    }
    for (; g < grid; g += kWays) acc += partials[kGnAcc * (size_t)g + v];
    s_sum[j][v] = acc;
  }
  __syncthreads();
  if (t < kGnAcc)
    tot[t] = ((s_sum[0][t] + s_sum[1][t]) + (s_sum[2][t] + s_sum[3][t])) +
             ((s_sum[4][t] + s_sum[5][t]) + (s_sum[6][t] + s_sum[7][t]));
  __syncthreads();
  if (t == 0) gn_step(tot, *reinterpret_cast<GnState*>(s_st), s_cand, cand_out, final_step, stats_row);
  __syncthreads();
  if (st_out && t < kStateWords) reinterpret_cast<double*>(st_out)[t] = s_st[t];
}

// The Levenberg-Marquardt decision and step of one evaluation (one lane), on the solver state
// staged in LDS (updated in place; the workgroup writes it out afterwards).
__device__ __forceinline__ void gn_step(const double* tot, GnState& S, double* s_cand, double* cand_out, int final_step,
                                        double* stats_row) {
  double c[12], row[kGnStats];
  const double cnt = tot[28];
  // tracking: depth term scaled to n_eff points, plus the motion prior
  const double al = S.n_eff > 0.0 && tot[29] > 0.0 ? S.n_eff / tot[29] : 1.0;
  double cost = al * tot[30], Ab[27];
#pragma unroll
  for (int k = 0; k < 27; ++k) Ab[k] = al * tot[k];
  if (S.lam_t > 0.0 || S.lam_r > 0.0) motion_prior(S, Ab, cost);
  row[0] = cnt;
  row[1] = cnt > 0.0 ? sqrt(tot[27] / cnt) : 0.0;
  row[2] = 0.0;
  const bool accept = !S.has || cost <= S.cost;
  if (accept) {
#pragma unroll
    for (int k = 0; k < 12; ++k) S.pose[k] = S.cand[k];  // the pose this evaluation was made at
#pragma unroll
    for (int k = 0; k < 27; ++k) S.Ab[k] = Ab[k];
    S.cost = cost;
    S.lambda = S.has ? fmax(S.lambda / 3.0, 1e-9) : S.lambda;
    S.has = 1;
  } else {
    S.lambda = fmin(S.lambda * 4.0, 1e9);
  }
  row[3] = S.lambda;
  row[4] = accept ? 1.0 : 0.0;
#pragma unroll
  for (int k = 0; k < 12; ++k) row[5 + k] = S.pose[k];
  // damped normal equations, Cholesky, the two triangular solves (all loops unrolled: the
  // matrices stay in registers); a matrix that is not positive definite proposes the
  // accepted pose again
  double A[6][6], b[6], L[6][6], x[6], y[6];
#pragma unroll
  for (int p = 0, m = 0; p < 6; ++p)
#pragma unroll
    for (int q = p; q < 6; ++q, ++m) { A[p][q] = S.Ab[m]; A[q][p] = S.Ab[m]; }
#pragma unroll
  for (int p = 0; p < 6; ++p) {
    b[p] = -S.Ab[21 + p];
    A[p][p] += S.lambda * A[p][p] + 1e-12;
  }
  bool pd = true;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    double d = A[q][q];
#pragma unroll
    for (int k = 0; k < q; ++k) d -= L[q][k] * L[q][k];
    pd = pd && d > 0.0;
    L[q][q] = sqrt(d);
#pragma unroll
    for (int p = q + 1; p < 6; ++p) {
      double a = A[p][q];
#pragma unroll
      for (int k = 0; k < q; ++k) a -= L[p][k] * L[q][k];
      L[p][q] = a / L[q][q];
    }
  }
#pragma unroll
  for (int p = 0; p < 6; ++p) {
    double a = b[p];
#pragma unroll
    for (int k = 0; k < p; ++k) a -= L[p][k] * y[k];
    y[p] = a / L[p][p];
  }
#pragma unroll
  for (int p = 5; p >= 0; --p) {
    double a = y[p];
#pragma unroll
    for (int k = p + 1; k < 6; ++k) a -= L[k][p] * x[k];
    x[p] = a / L[p][p];
  }
  if (!final_step && pd) {
    se3_left_update(x, S.pose, c);
    row[2] = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2] + x[3] * x[3] + x[4] * x[4] + x[5] * x[5]);
  } else {
#pragma unroll
    for (int k = 0; k < 12; ++k) c[k] = S.pose[k];
  }
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    S.cand[k] = c[k];
    s_cand[k] = c[k];
  }
  if (cand_out) {
#pragma unroll
    for (int k = 0; k < 12; ++k) cand_out[k] = c[k];
  }
  if (stats_row) {
#pragma unroll
    for (int k = 0; k < kGnStats; ++k) stats_row[k] = row[k];
  }
}

// The last evaluation's step alone (one workgroup): with final_step the accepted pose is the
// candidate written to cand.
__global__ __launch_bounds__(256) void k_gn_solve(const double* partials, int grid, const GnState* st_in,
                                                  GnState* st_out, double* cand, int final_step, double* stats_row) {
  __shared__ double s_cand[12];
  gn_solve_wg(partials, grid, st_in, st_out, s_cand, cand, final_step, stats_row);
}

// Start of a refinement: the candidate is the starting pose and the solver state is cleared
// with the initial damping. The pose comes in by value (kernel argument), so starting a
// refinement needs no host-to-device copy.
__global__ __launch_bounds__(64) void k_gn_start(GnStart s, GnState* st, double* cand) {
  const int k = threadIdx.x;
  if (k == 0) {
    GnState z{};
    z.lambda = s.lambda;
#pragma unroll
    for (int q = 0; q < 12; ++q) z.cand[q] = s.pose[q];
    *st = z;
  }
  if (k < 12) cand[k] = s.pose[k];
}

// ============================================================================ tracking
__device__ __forceinline__ uint32_t track_slot(const TrackArgs& t) { return t.ctl->cursor[1] % t.n_seq; }

// The frame's starting pose becomes the candidate; the solver state is cleared (a kernel,
// not a memset node, so the captured graph replays it exactly). With increments, start =
// out[f-1] * prior[f] in FP64, each entry summed in column order ((a0 b0 + a1 b1) + a2 b2) +
// a3 b3. Lanes 0..11 of workgroup 0 of k_track_begin.
__device__ void track_start(const TrackArgs& t, uint32_t fs) {
  const int k = threadIdx.x;
  const float* P = t.prior + 16 * (size_t)fs;
  if (k < 12) {
    double p;
    if (t.increment && fs > 0) {
      const float* A = t.out + 16 * (size_t)(fs - 1);
      const int r = k >> 2, c = k & 3;
      p = (((double)A[4 * r] * (double)P[c] + (double)A[4 * r + 1] * (double)P[4 + c]) +
           (double)A[4 * r + 2] * (double)P[8 + c]) + (double)A[4 * r + 3] * (double)P[12 + c];
    } else {
      p = (double)P[k];
    }
    t.pose[k] = p;
    t.state->cand[k] = p;   // the first evaluation's pose
    t.state->start[k] = p;  // the motion prior's centre
  }
  if (k == 0) {  // no accepted pose yet (the other fields are set by the first acceptance)
    t.state->has = 0;
    t.state->lambda = 1e-4;
    const double st = t.prior_sigma_t, sr = t.prior_sigma_r;
    t.state->lam_t = st > 0.0 ? 1.0 / (st * st) : 0.0;
    t.state->lam_r = sr > 0.0 ? 1.0 / (sr * sr) : 0.0;
    t.state->n_eff = t.n_eff;
  }
}

// Start of a tracked frame: workgroup 0 sets the starting pose (track_start), and every lane
// back-projects one stride-th pixel to its camera-frame point ((u - cx) d / fx,
// (v - cy) d / fy, d) in FP32 with this operation order; pixels without a valid depth give
// NaN (skipped by the solver).
template <bool kU16>
__global__ __launch_bounds__(256) void k_track_begin(TrackArgs t) {
  const uint32_t fs = track_slot(t);
  if (blockIdx.x == 0) track_start(t, fs);
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= t.pw * t.ph) return;
  const int u = (i % t.pw) * t.stride, v = (i / t.pw) * t.stride;
  const size_t elem = kU16 ? 2 : 4;
  const void* base = reinterpret_cast<const char*>(t.depth) + (size_t)fs * t.W * t.H * elem;
  const float d = depth_at<kU16>(base, t.depth_scale, v * t.W + u);
  float x = __builtin_nanf(""), y = x, z = x;
  if (depth_valid(d, t.max_depth)) {
    x = (((float)u - t.K.cx) * d) / t.K.fx;
    y = (((float)v - t.K.cy) * d) / t.K.fy;
    z = d;
  }
  t.pts[3 * i] = x;
  t.pts[3 * i + 1] = y;
  t.pts[3 * i + 2] = z;
}

// One workgroup: the accepted pose (the candidate after the final step), rounded to FP32, is
// the frame's pose; the solver's rows of this frame go to the per-frame stats.
__global__ __launch_bounds__(64) void k_track_end(TrackArgs t) {
  const uint32_t fs = track_slot(t);
  const int k = threadIdx.x;
  if (k < 16) t.out[16 * (size_t)fs + k] = k < 12 ? (float)t.pose[k] : (k == 15 ? 1.0f : 0.0f);
  if (t.stats) {
    const int rows = (t.iters + 1) * kGnStats;
    for (int j = k; j < rows; j += 64) t.stats[(size_t)fs * rows + j] = t.eval_stats[j];
  }
}

}  // namespace

// ============================================================================ launchers
template <class F>
static void with_input(bool u16, bool color, F f) {
  using T = std::true_type;
  using N = std::false_type;
  if (u16) {
    if (color) f(T(), T());
    else f(T(), N());
  } else {
    if (color) f(N(), T());
    else f(N(), N());
  }
}

hipError_t launch_alloc(const FrameArgs& a, int depth_u16, int k, hipStream_t st) {
  const int rows = k > 1 ? kAllocRows : 1;
  const uint32_t tiles = alloc_tiles(a.W, a.H, rows);
  const dim3 g(tiles * k), b(256);
  with_input(depth_u16 != 0, a.rgba != nullptr, [&](auto U, auto C) {
    constexpr bool u16 = decltype(U)::value, col = decltype(C)::value;
    if (rows > 1) hipLaunchKernelGGL((k_alloc<u16, kAllocRows, col>), g, b, 0, st, a, tiles);
    else hipLaunchKernelGGL((k_alloc<u16, 1, col>), g, b, 0, st, a, tiles);
  });
  return hipGetLastError();
}

// Calls f(std::integral_constant<int, storage>) for the runtime storage format.
template <class F>
static void with_storage(int storage, F f) {
  switch (storage) {
    case VH_STORAGE_F16: f(std::integral_constant<int, VH_STORAGE_F16>()); break;
    case VH_STORAGE_F16R: f(std::integral_constant<int, VH_STORAGE_F16R>()); break;
    default: f(std::integral_constant<int, VH_STORAGE_F32>()); break;
  }
}

// Calls f(std::integral_constant<bool, u16>(), std::integral_constant<bool, colour>()).

hipError_t launch_integrate(const FrameArgs& a, int storage, int grid, int k, hipStream_t st) {
  with_storage(storage, [&](auto S) {
    const dim3 g(grid / kIntegrateWaves), b(64 * kIntegrateWaves);
    constexpr int kS = decltype(S)::value;
    const auto kern = a.colors ? (k > 1 ? k_integrate<kS, true, true> : k_integrate<kS, true, false>)
                               : (k > 1 ? k_integrate<kS, false, true> : k_integrate<kS, false, false>);
    hipLaunchKernelGGL(kern, g, b, 0, st, a, (uint32_t)k);
  });
  return hipGetLastError();
}

hipError_t launch_carve(const FrameArgs& a, int k, hipStream_t st) {
  const uint32_t g = min(kCarveGroups, (a.pool_cap + 255u) / 256u);
  hipLaunchKernelGGL(k_carve, dim3(g), dim3(256), 0, st, a, (uint32_t)k);
  return hipGetLastError();
}

hipError_t launch_integrate_alloc(const FrameArgs& a, int storage, int depth_u16, int grid, int k_int, int k_alloc,
                                  hipStream_t st) {
  const uint32_t ig = (uint32_t)(grid / kIntegrateWaves);
  const uint32_t tiles = alloc_tiles(a.W, a.H, kAllocRows);
  const uint32_t ag = tiles * (uint32_t)k_alloc;
  with_storage(storage, [&](auto S) {
    with_input(depth_u16 != 0, a.colors != nullptr, [&](auto U, auto C) {
      const dim3 g(ig + ag), b(64 * kIntegrateWaves);
      hipLaunchKernelGGL((k_integrate_alloc<decltype(S)::value, decltype(U)::value, decltype(C)::value>), g, b, 0,
                         st, a, ig, ag, tiles, (uint32_t)k_int);
    });
  });
  return hipGetLastError();
}

// Grid of the per-block extraction kernels: one wave per block, kMcWaves per workgroup, so
// the dispatcher balances blocks of unequal work (empty, dense, near-corner). Measured on the
// final room_medium_5mm volume: 0.74 ms per extraction with one-wave workgroups against
// 0.87 ms with 8,192 block-striding workgroups and 0.85 ms with one resident round of them.
template <class Kernel>
static uint32_t block_grid(Kernel, uint32_t n_blocks) {
  return (n_blocks + kMcWaves - 1) / kMcWaves;
}

hipError_t launch_mc(const McArgs& a, int storage, hipStream_t st) {
  if (a.n_blocks == 0) return hipSuccess;
  const uint32_t n_chunks = (a.n_blocks + kScanChunk - 1) / kScanChunk;
  hipLaunchKernelGGL(k_mc_neighbours, dim3((uint32_t)(((size_t)a.n_blocks * 8 + 255) / 256)), dim3(256), 0, st, a);
  with_storage(storage, [&](auto S) {
    auto kc = k_mc_count<decltype(S)::value>;
    hipLaunchKernelGGL(kc, dim3(block_grid(kc, a.n_blocks)), dim3(64 * kMcWaves), 0, st, a);
  });
  hipLaunchKernelGGL(k_scan_chunks, dim3(n_chunks), dim3(256), 0, st, a);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(256), 0, st, a, n_chunks);
  with_storage(storage, [&](auto S) {
    auto kw = a.colors ? k_mc_write<decltype(S)::value, true> : k_mc_write<decltype(S)::value, false>;
    hipLaunchKernelGGL(kw, dim3(block_grid(kw, a.n_blocks)), dim3(64 * kMcWaves), 0, st, a);
  });
  return hipGetLastError();
}

hipError_t launch_mesh_index_count(const McArgs& a, int storage, hipStream_t st) {
  if (a.n_blocks == 0) return hipSuccess;
  const uint32_t n_chunks = (a.n_blocks + kScanChunk - 1) / kScanChunk;
  hipLaunchKernelGGL(k_mc_neighbours, dim3((uint32_t)(((size_t)a.n_blocks * 8 + 255) / 256)), dim3(256), 0, st, a);
  with_storage(storage, [&](auto S) {
    auto km = k_ix_mark<decltype(S)::value>;
    hipLaunchKernelGGL(km, dim3(block_grid(km, a.n_blocks)), dim3(64 * kMcWaves), 0, st, a);
  });
  hipLaunchKernelGGL(k_ix_count, dim3(min((a.n_blocks + 3u) / 4u, 16384u)), dim3(256), 0, st, a);
  hipLaunchKernelGGL(k_scan_chunks, dim3(n_chunks), dim3(256), 0, st, a);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(256), 0, st, a, n_chunks);
  McArgs v = a;  // the same scans over the vertex counts
  v.counts = a.vcounts;
  v.chunk_sums = a.vchunks;
  v.total = a.vtotal;
  hipLaunchKernelGGL(k_scan_chunks, dim3(n_chunks), dim3(256), 0, st, v);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(256), 0, st, v, n_chunks);
  return hipGetLastError();
}

hipError_t launch_mesh_index_write(const McArgs& a, int storage, hipStream_t st) {
  if (a.n_blocks == 0) return hipSuccess;
  with_storage(storage, [&](auto S) {
    auto kw = (a.colors && a.vrgb) ? k_ix_write<decltype(S)::value, true> : k_ix_write<decltype(S)::value, false>;
    hipLaunchKernelGGL(kw, dim3(block_grid(kw, a.n_blocks)), dim3(64 * kMcWaves), 0, st, a);
  });
  return hipGetLastError();
}

static uint32_t stream_grid(uint32_t n) { return min((n + 3u) / 4u, 16384u); }

hipError_t launch_stream_out(const StreamArgs& a, hipStream_t st) {
  if (a.n) hipLaunchKernelGGL(k_st_out, dim3(stream_grid(a.n)), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_stream_move(const StreamArgs& a, hipStream_t st) {
  if (a.n) hipLaunchKernelGGL(k_st_move, dim3(stream_grid(a.n)), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_stream_rehash(const StreamArgs& a, hipStream_t st) {
  if (a.n) hipLaunchKernelGGL(k_st_rehash, dim3((a.n + 255u) / 256u), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_stream_insert(const StreamArgs& a, int storage, hipStream_t st) {
  if (!a.n) return hipSuccess;
  hipLaunchKernelGGL(k_st_insert, dim3((a.n + 255u) / 256u), dim3(256), 0, st, a);
  with_storage(storage, [&](auto S) {
    hipLaunchKernelGGL(k_st_place<decltype(S)::value>, dim3(stream_grid(a.n)), dim3(256), 0, st, a);
  });
  return hipGetLastError();
}

static int query_grid(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g > 4096) g = 4096;
  if (g < 1) g = 1;
  return (int)g;
}

hipError_t launch_query(const QueryArgs& q, int storage, const float* pts, int64_t n, float* sdf, uint8_t* valid,
                        hipStream_t st) {
  const int g = query_grid(n);
  with_storage(storage, [&](auto S) {
    hipLaunchKernelGGL(k_query<decltype(S)::value>, dim3(g), dim3(256), 0, st, q, pts, n, sdf, valid);
  });
  return hipGetLastError();
}

hipError_t launch_backward(const QueryArgs& q, int storage, const float* pts, int64_t n, const float* T12,
                           const float* grad_r, float* r, uint8_t* valid, float* J, double* partials, int grid,
                           float* voxel_grad, hipStream_t st) {
  with_storage(storage, [&](auto S) {
    hipLaunchKernelGGL(k_backward<decltype(S)::value>, dim3(grid), dim3(256), 0, st, q, pts, n, T12, grad_r, r,
                       valid, J, partials, voxel_grad);
  });
  return hipGetLastError();
}

hipError_t launch_reduce6(const double* partials, int grid, double* out6, hipStream_t st) {
  hipLaunchKernelGGL(k_reduce6, dim3(1), dim3(64), 0, st, partials, grid, out6);
  return hipGetLastError();
}

hipError_t launch_gn_accum(const GnArgs& a, int storage, int grid, hipStream_t st) {
  with_storage(storage, [&](auto S) {
    hipLaunchKernelGGL(k_gn_accum<decltype(S)::value>, dim3(grid), dim3(256), 0, st, a);
  });
  return hipGetLastError();
}

hipError_t launch_gn_solve(const double* partials, int grid, const GnState* st_in, GnState* st_out, double* cand,
                           int final_step, double* stats_row, hipStream_t st) {
  hipLaunchKernelGGL(k_gn_solve, dim3(1), dim3(256), 0, st, partials, grid, st_in, st_out, cand, final_step, stats_row);
  return hipGetLastError();
}

hipError_t launch_gn_start(const GnStart& s, GnState* state, double* cand, hipStream_t st) {
  hipLaunchKernelGGL(k_gn_start, dim3(1), dim3(64), 0, st, s, state, cand);
  return hipGetLastError();
}

hipError_t launch_track_begin(const TrackArgs& t, hipStream_t st) {
  const int n = t.pw * t.ph;
  const dim3 g((n + 255) / 256), b(256);
  if (t.depth_u16) hipLaunchKernelGGL(k_track_begin<true>, g, b, 0, st, t);
  else hipLaunchKernelGGL(k_track_begin<false>, g, b, 0, st, t);
  return hipGetLastError();
}

hipError_t launch_track_end(const TrackArgs& t, hipStream_t st) {
  hipLaunchKernelGGL(k_track_end, dim3(1), dim3(64), 0, st, t);
  return hipGetLastError();
}

}  // namespace vh
