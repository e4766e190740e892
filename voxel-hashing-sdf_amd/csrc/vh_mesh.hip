// vh_mesh.hip - indexed mesh from the marching-cubes triangle soup (gfx950).
//
// The soup (k_mc_write) carries, per triangle corner, the key of the lattice edge the corner
// lies on (mc_edge_key). Corners are sorted by key (radix sort over the 38 key bits), the
// corners of the k-th distinct key get vertex id k (inclusive scan of first-of-key flags,
// minus one),
// every corner's face entry takes its key's id, and each vertex is computed once from the
// voxel values at the edge's two end points, interpolated from the lower end point to the
// upper one (VertexInterp, src/tsdf.cu:1640-1660, with that orientation). All triangles that
// use an edge therefore share one vertex: the mesh has no cracks where neighbouring cubes
// interpolate the same edge in opposite directions (the reference welds float-identical
// coordinates instead, src/tsdf.cu:1772-1850).
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include "vh.h"
#include "vh_internal.h"
#include "vh_math.h"

namespace vh {

namespace {

__global__ __launch_bounds__(256) void k_iota(uint32_t* idx, uint32_t n) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i < n) idx[i] = i;
}

__global__ __launch_bounds__(256) void k_first_of_key(const unsigned long long* keys, uint32_t* flag, uint32_t n) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i < n) flag[i] = (i == 0u || keys[i] != keys[i - 1]) ? 1u : 0u;
}

// Voxel (tsdf, weight) and colour at pool block `pool`, in-block index `vi`.
template <int kStorage>
__device__ __forceinline__ void voxel_at(const IndexArgs& a, uint32_t pool, uint32_t vi, float& s, float& w,
                                         uint32_t& c) {
  using IO = VoxelIO<kStorage>;
  const size_t v = (size_t)pool * kB3 + vi;
  IO::unpack(reinterpret_cast<const typename IO::Word*>(a.voxels)[v], s, w);
  c = a.colors ? a.colors[v] : 0u;
}

// Faces and vertices: corner i (sorted order) writes its face entry; the first corner of a
// key computes the vertex.
template <int kStorage>
__global__ __launch_bounds__(256) void k_emit_indexed(IndexArgs a) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= a.n) return;
  const uint32_t id = a.vid[i] - 1u;  // vid: inclusive count of distinct keys up to i
  a.faces[a.idx_sorted[i]] = (int32_t)id;
  if (i == a.n - 1u) *a.n_vertices = a.vid[i];
  if (!a.flag[i] || id >= a.vertex_cap) return;
  const unsigned long long key = a.keys_sorted[i];
  const uint32_t axis = (uint32_t)(key & 3u), vi = (uint32_t)((key >> 2) & 511u), pool = (uint32_t)(key >> 11);
  int bx, by, bz;
  unpack_key(a.block_keys[pool], bx, by, bz);
  int lo[3] = {bx * kB + (int)(vi >> 6), by * kB + (int)((vi >> 3) & 7u), bz * kB + (int)(vi & 7u)};
  // upper end point: the next lattice point along the axis, in this block or the next one
  const uint32_t c_axis = axis == 0 ? (vi >> 6) : (axis == 1 ? ((vi >> 3) & 7u) : (vi & 7u));
  const uint32_t stride = axis == 0 ? 64u : (axis == 1 ? 8u : 1u);
  uint32_t hpool = pool, hvi = vi + stride;
  if (c_axis == 7u) {
    hvi = vi - 7u * stride;
    hpool = find_block(a.hash, pack_key(bx + (axis == 0), by + (axis == 1), bz + (axis == 2)));
  }
  float s0, w0, s1 = 0.0f, w1 = 0.0f;
  uint32_t c0, c1 = 0u;
  voxel_at<kStorage>(a, pool, vi, s0, w0, c0);
  if (hpool < a.pool_cap) voxel_at<kStorage>(a, hpool, hvi, s1, w1, c1);
  const float p0[3] = {(float)lo[0], (float)lo[1], (float)lo[2]};
  const float p1[3] = {(float)(lo[0] + (axis == 0)), (float)(lo[1] + (axis == 1)), (float)(lo[2] + (axis == 2))};
  float x, y, z;
  uint32_t c;
  vertex_interp_rgb(p0[0], p0[1], p0[2], p1[0], p1[1], p1[2], s0, s1, c0, c1, x, y, z, c);
  float* o = a.vertices + (size_t)id * 3;
  o[0] = x * a.vox;
  o[1] = y * a.vox;
  o[2] = z * a.vox;
  if (a.vrgb) {
    uint8_t* q = a.vrgb + (size_t)id * 3;
    q[0] = (uint8_t)c;
    q[1] = (uint8_t)(c >> 8);
    q[2] = (uint8_t)(c >> 16);
  }
}

}  // namespace

size_t mesh_index_temp_bytes(uint32_t n) {
  size_t sort_bytes = 0, scan_bytes = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (const unsigned long long*)nullptr,
                                           (unsigned long long*)nullptr, (const uint32_t*)nullptr,
                                           (uint32_t*)nullptr, (int)n, 0, kEdgeKeyBits);
  (void)hipcub::DeviceScan::InclusiveSum(nullptr, scan_bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)n);
  return sort_bytes > scan_bytes ? sort_bytes : scan_bytes;
}

hipError_t launch_mesh_index(const IndexArgs& a, int storage, hipStream_t st) {
  if (a.n == 0) return hipMemsetAsync(a.n_vertices, 0, sizeof(uint32_t), st);
  const dim3 g((a.n + 255u) / 256u), b(256);
  hipLaunchKernelGGL(k_iota, g, b, 0, st, a.idx, a.n);
  size_t bytes = a.temp_bytes;
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(a.temp, bytes, a.keys, a.keys_sorted, a.idx, a.idx_sorted,
                                                    (int)a.n, 0, kEdgeKeyBits, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_first_of_key, g, b, 0, st, a.keys_sorted, a.flag, a.n);
  bytes = a.temp_bytes;
  e = hipcub::DeviceScan::InclusiveSum(a.temp, bytes, a.flag, a.vid, (int)a.n, st);
  if (e != hipSuccess) return e;
  switch (storage) {
    case VH_STORAGE_F16: hipLaunchKernelGGL(k_emit_indexed<VH_STORAGE_F16>, g, b, 0, st, a); break;
    case VH_STORAGE_F16R: hipLaunchKernelGGL(k_emit_indexed<VH_STORAGE_F16R>, g, b, 0, st, a); break;
    default: hipLaunchKernelGGL(k_emit_indexed<VH_STORAGE_F32>, g, b, 0, st, a); break;
  }
  return hipGetLastError();
}

}  // namespace vh
