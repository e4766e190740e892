/* vh.h - C ABI of libvh, an MI355X (gfx950) voxel-hashing TSDF engine.
 *
 * The reference exports one C++ class, ark::GpuTsdfGenerator (include/tsdf.cuh:548-683),
 * from libTSDF.so (CMakeLists.txt:15-16). libvh exports the same three operations -
 * construct, integrate a depth frame at a pose, extract the mesh - as plain extern "C"
 * functions over pointers and sizes, plus the extensions the reference lacks (trilinear
 * SDF query, its backward pass, Gauss-Newton pose refinement, FP16 voxel storage).
 * A header-only C++ class with the reference's constructor and method names sits on top
 * of this ABI (voxel-hashing-sdf_amd/host/GpuTsdfGenerator.hpp).
 *
 * Conventions (kept from the reference):
 *   - depth is H x W row-major, metres (float) or uint16 scaled by depth_scale
 *     (SaveFrame.cpp:179-180 uses 0.001); 0 marks an invalid pixel; a pixel is used iff
 *     0 < d <= max_depth (src/tsdf.cu:713), so NaN is never used (the reference's form
 *     of the test lets NaN through; its 16-bit depth inputs never hold one);
 *   - c2w is a row-major 4x4 camera-to-world matrix (PointCloudGenerator.cpp:118-125);
 *   - voxel (i,j,k) sits at (i,j,k)*vox_size, blocks are 8^3 voxels, in-block order
 *     (x*8+y)*8+z (src/tsdf.cu:24-64);
 *   - the TSDF is normalised: min(1, (d - z)/trunc), voxels with d - z <= -trunc are not
 *     updated (src/tsdf.cu:718-738).
 *
 * Every entry point returns a vh_status. No entry point exits, throws or traps; the text of
 * the last failure on the calling thread is returned by vh_last_error().
 * A volume is externally synchronised: calls on one volume must not overlap.
 */
#ifndef VH_H
#define VH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VH_ABI_VERSION 7

typedef enum vh_status {
  VH_OK = 0,
  VH_ERR_INVALID_ARG = 1,
  VH_ERR_OOM = 2,
  VH_ERR_POOL_FULL = 3,  /* a block could not get pool memory; it was not integrated */
  VH_ERR_HASH_FULL = 4,  /* a key found no free slot within the probe limit */
  VH_ERR_DEVICE = 5,     /* a HIP runtime call failed */
  VH_ERR_RANGE = 6,      /* a truncation band left the representable block range */
  VH_ERR_TRUNCATED = 7,  /* an output buffer was too small; the required size is reported */
  VH_ERR_IO = 8,
  VH_ERR_NO_DEVICE = 9
} vh_status;

typedef enum vh_storage {
  VH_STORAGE_F32 = 0,  /* float tsdf + float weight, 8 B per voxel */
  VH_STORAGE_F16 = 1,  /* binary16 tsdf + uint16 weight, 4 B per voxel; updates in FP32 */
  VH_STORAGE_F16R = 2  /* binary16 tsdf + uint8 weight + int8 rounding residual (units of
                          ulp/256 of the binary16 value), 4 B per voxel; the residual carries
                          the rounding error of each store into the next update */
} vh_storage;

typedef enum vh_depth_format { VH_DEPTH_U16 = 0, VH_DEPTH_F32 = 1 } vh_depth_format;

typedef struct vh_volume vh_volume;

/* Engine parameters. Replaces the 15-argument constructor
 * GpuTsdfGenerator(width, height, fx, fy, cx, cy, max_depth, origin_x/y/z, vox_size,
 * trunc_m, vox_dim_x/y/z) (include/tsdf.cuh:605-607, src/tsdf.cu:1248-1372) and the hash
 * sizes the reference hard-codes (src/tsdf.cu:1488,1492). origin and vox_dim are accepted
 * for compatibility and ignored, as the reference's hash path ignores them. */
typedef struct vh_config {
  int32_t width, height;
  float fx, fy, cx, cy;
  float max_depth;       /* metres */
  float min_depth;       /* metres: near end of a truncation band (reference: 0.1) */
  float vox_size;        /* metres */
  float trunc;           /* metres (mu) */
  int32_t log2_buckets;  /* 2^log2_buckets hash buckets of 4 slots each (1..30) */
  int32_t pool_blocks;   /* voxel-block pool capacity */
  int32_t storage;       /* vh_storage */
  int32_t weight_cap;    /* 0 = uncapped; VH_STORAGE_F16 requires 1..65535, F16R 1..255 */
  int32_t device;        /* HIP device ordinal */
  float origin[3];       /* ignored */
  int32_t vox_dim[3];    /* ignored */
  int32_t color;         /* 1: integrate colour (Voxel::sdf_color, include/tsdf.cuh:122): every
                            frame then needs its RGB image, and meshes carry vertex colours;
                            0: depth only (4 B per voxel less traffic each way) */
  int32_t seq_batch;     /* frames per integration launch on bound sequences (1..8; 0 = 8):
                            each block's voxels are read and written once per batch and
                            updated by every frame of the batch that touched it, in frame
                            order (the same result as frame by frame) */
  int32_t carve;         /* 1: each frame updates every allocated voxel it observes (projects into
                            the image at a depth 0 < d <= max_depth with d - z > -trunc), so
                            free space in front of the surface is carved toward +1, as the
                            reference's whole-ray block selection does (src/tsdf.cu:2128-2177,
                            isBlockInCameraFrustum :2056-2064); 0: only the blocks crossed by the
                            frame's truncation bands are updated (faster; other blocks keep
                            their values) */
  int32_t reserved[5];   /* zero */
} vh_config;

/* One input frame in host memory (replaces processFrame's float* depth, uchar* rgb,
 * float* c2w, include/tsdf.cuh:609-610). */
typedef struct vh_frame {
  const void* depth;     /* height rows of pitch bytes */
  int32_t depth_format;  /* vh_depth_format */
  float depth_scale;     /* metres per unit for VH_DEPTH_U16; ignored for VH_DEPTH_F32 */
  int64_t pitch;         /* bytes per depth row; 0 = tightly packed */
  const uint8_t* rgb;    /* colour volumes: height x width x 3 bytes, tightly packed, the
                            channels in the order the mesh colours are to have (the reference
                            passes its image's bytes through unchanged); ignored otherwise */
  float c2w[16];         /* row-major camera-to-world */
} vh_frame;

typedef struct vh_stats {
  int64_t frames;             /* frames integrated since create/reset */
  int64_t blocks_allocated;   /* pool blocks in use */
  int64_t pool_capacity;
  int64_t slots_used;         /* hash slots holding a key (>= blocks_allocated) */
  int64_t touched_last;       /* distinct blocks crossed by truncation bands, last frame */
  int64_t touched_total;      /* sum of touched_last over all frames */
  int64_t new_total;          /* blocks allocated since create/reset */
  int32_t storage;
  int32_t flags;              /* bit 0 pool full, bit 1 hash full, bit 2 range */
  double integrate_ms;        /* profiling mode: summed integration-kernel time */
  double alloc_ms;            /* profiling mode: summed allocation-kernel time */
  int64_t profiled_frames;
  double track_ms;            /* profiling mode: summed tracking time (back-projection + solver) */
  int64_t profiled_launches;  /* profiling mode: integration launches timed (one per batch) */
  int64_t blocks_integrated;  /* sum over integration launches of the blocks read and written
                                 (a block touched by several frames of a batch counts once) */
  int64_t host_blocks;        /* blocks in the host store (vh_stream_out) */
  int64_t streamed_out_total; /* blocks moved to the host store since create/reset */
  int64_t streamed_in_total;  /* blocks moved back into the pool since create/reset */
  int64_t stream_merged_total;/* of those, blocks merged into a resident block of the same key */
  int64_t carved_total;       /* carving volumes: block-frame updates added by carving (a frame
                                 observes a block that its truncation bands did not cross) */
  double carve_ms;            /* profiling mode: summed carve-pass time */
  double fused_ms;            /* profiling mode 2: summed time of the fused batch launches (one
                                 captured graph: integration of a batch and allocation of the
                                 next), HIP events around each on the volume's stream */
  int64_t fused_launches;     /* profiling mode 2: fused batch launches timed */
} vh_stats;

typedef enum vh_prior_mode {
  VH_PRIOR_ABSOLUTE = 0,  /* d_prior_c2w[f] is frame f's starting pose */
  VH_PRIOR_INCREMENT = 1  /* d_prior_c2w[f] is the motion from frame f-1 (f > 0) */
} vh_prior_mode;

/* Tracking of a bound sequence (vh_bind_tracking). Device buffers stay owned by the caller
 * and must outlive the frames that use them. */
typedef struct vh_track_config {
  const float* d_prior_c2w; /* n_frames x 16 row-major (device) */
  float* d_out_c2w;         /* n_frames x 16 row-major refined poses (device) */
  double* d_stats;          /* n_frames x (gn_iters + 1) x 17 (device) or NULL */
  int32_t gn_iters;         /* 0..63 */
  int32_t point_stride;     /* 1..64 */
  float max_point_depth;    /* metres; <= 0: the volume's max_depth */
  int32_t prior_mode;       /* vh_prior_mode */
  /* Motion prior (0: off; each >= 0): the solver adds 1/2 xi^T diag(I3/st^2, I3/sr^2) xi,
   * xi = log(T_start^-1 T) (translation first), to the cost, T_start being the frame's starting
   * pose; st in metres and sr in radians are the per-frame noise of the prior's motion. */
  float prior_sigma_t;
  float prior_sigma_rot;
  /* > 0: the depth term is scaled by depth_points / (points with a depth), so that the depth
   * image counts as this many independent points against the motion prior (the residuals of
   * one frame share the map's errors); 0: unscaled. */
  float depth_points;
  int32_t reserved[1];      /* zero */
} vh_track_config;

/* Triangle soup in host memory, owned by the library (freed by vh_mesh_free). */
typedef struct vh_mesh {
  int64_t n_tris;
  float* xyz;       /* n_tris * 9 floats: three vertices in metres */
  int32_t* cubes;   /* n_tris * 4: cube's lattice x,y,z and (cube_case << 8 | tri_in_cube) */
  uint8_t* rgb;     /* colour volumes: n_tris * 9 bytes, 3 channels per vertex; else NULL */
} vh_mesh;

typedef struct vh_gn_stats {
  int32_t iters;
  int32_t n_valid[64];  /* points with a valid trilinear sample at each evaluated candidate */
  double rms[64];       /* residual RMS (normalised TSDF units) over those points */
  double step_norm[64]; /* |delta xi| proposed after the evaluation (0 for the last) */
  double pose[64][12];  /* accepted pose after each evaluation (row-major 3x4) */
} vh_gn_stats;

/* ------------------------------------------------------------------ lifetime */
void vh_config_default(vh_config* cfg);
/* GpuTsdfGenerator::GpuTsdfGenerator (src/tsdf.cu:1248-1372). */
int vh_create(const vh_config* cfg, vh_volume** out);
/* GpuTsdfGenerator::Shutdown (src/tsdf.cu:1620-1638), which the reference never calls. */
int vh_destroy(vh_volume* vol);
/* Empties the volume (all blocks freed). No reference counterpart: the reference rebuilds
 * its per-frame tables in processFrame (src/tsdf.cu:1488-1492). */
int vh_reset(vh_volume* vol);
const char* vh_last_error(void);
int vh_abi_version(void);

/* ------------------------------------------------------------------ integration */
/* GpuTsdfGenerator::processFrame(depth, rgb, c2w) (src/tsdf.cu:1484-1598): allocates
 * every block crossed by a valid pixel's truncation band and integrates those blocks (and,
 * in a colour volume, the frame's RGB image at the same pixels: VH_ERR_INVALID_ARG without
 * one). Host buffers are copied before return; the call is synchronous. */
int vh_integrate(vh_volume* vol, const vh_frame* frame);
/* Asynchronous variant of vh_integrate (no reference counterpart; processFrame blocks,
 * src/tsdf.cu:1495-1497): the depth (and RGB) is copied on a second stream into one of 4
 * device slots while earlier frames integrate, and the frame's allocation and integration
 * wait for the copy. The call returns after queuing; the buffers must stay unchanged until vh_sync
 * (page-locked memory, e.g. from hipHostMalloc, gives an asynchronous copy; pageable memory
 * is staged by the runtime before the call returns). Overflows are reported by vh_sync. */
int vh_integrate_async(vh_volume* vol, const vh_frame* frame);

/* Device-resident input: binds n_frames depth images (tightly packed, device memory,
 * frame f at d_depth + f*width*height*elem) and n_frames row-major 4x4 poses
 * (device memory). Later vh_integrate_next calls consume them in order. A colour volume
 * also needs the sequence's RGB images (vh_bind_sequence_rgb) before vh_integrate_next. */
int vh_bind_sequence(vh_volume* vol, const void* d_depth, int32_t depth_format,
                     float depth_scale, const float* d_c2w, int32_t n_frames);
/* RGB images of the sequence bound last (by vh_bind_sequence or vh_bind_tracking): frame f
 * at d_rgb + f*width*height*3 (device memory, vh_frame.rgb's layout). Colour volumes only;
 * each bind call clears it. */
int vh_bind_sequence_rgb(vh_volume* vol, const uint8_t* d_rgb);
/* Device-resident input with tracking: like vh_bind_sequence, but each frame's pose is
 * first refined against the current volume (gn_iters Levenberg-Marquardt steps of
 * vh_gn_refine over the frame's depth back-projected on every point_stride-th pixel of every
 * point_stride-th row; pixels with depth outside (0, min(max_depth, max_point_depth)] are
 * skipped, max_point_depth <= 0 meaning max_depth), then the frame is integrated at the
 * refined pose, which is also written (FP32, row-major 4x4) to d_out_c2w[f]. The starting
 * pose of frame f is d_prior_c2w[f] (VH_PRIOR_ABSOLUTE), or, with VH_PRIOR_INCREMENT and
 * f > 0, the refined pose of frame f-1 times d_prior_c2w[f] (a camera-frame motion, e.g.
 * odometry); frame 0 always starts from d_prior_c2w[0]. d_stats (device, may be NULL)
 * receives per frame (gn_iters + 1) rows of 17 doubles: n_valid, rms, |delta xi|, lambda,
 * accepted, accepted pose (12). On an empty volume no point has a sample and the starting
 * pose is kept.
 * No reference counterpart (the reference takes poses as given, PointCloudGenerator.cpp:93-128). */
int vh_bind_tracking(vh_volume* vol, const void* d_depth, int32_t depth_format, float depth_scale,
                     int32_t n_frames, const vh_track_config* track);
/* Enqueues the next n frames of the bound sequence on the volume's stream (one captured
 * HIP graph launch per frame unless profiling is on). Asynchronous; see vh_sync. */
int vh_integrate_next(vh_volume* vol, int32_t n);
/* Waits for queued work and reports any pool/hash/range overflow raised by it. */
int vh_sync(vh_volume* vol);
/* Profiling mode 1: frames launch eagerly with HIP events around each kernel; the summed
 * kernel times appear in vh_stats. Mode 2: bound sequences run their captured graphs as
 * without profiling, with HIP events around each fused batch launch (vh_stats.fused_ms).
 * 0: off. Switching clears the sums. */
int vh_set_profiling(vh_volume* vol, int32_t on);
int vh_get_stats(vh_volume* vol, vh_stats* out);
/* The volume's HIP stream (hipStream_t), for callers that order their own work on it. */
void* vh_stream(vh_volume* vol);

/* ------------------------------------------------------------------ out-of-core streaming */
/* streamOutGPU2CPU (src/tsdf.cu:469-596): every resident block whose centre (the mean of its
 * 8^3 lattice points) lies farther than radius metres from center moves from the GPU pool
 * into the volume's host store, its voxel words (and colours) byte for byte. The pool is then
 * compacted (blocks above the new count take the vacated indices below it) and the hash
 * table rebuilt. *n_blocks receives the blocks moved. Waits for queued work first; meshes,
 * exports and queries see resident blocks only. */
int vh_stream_out(vh_volume* vol, const float center[3], float radius, int64_t* n_blocks);
/* streamInCPU2GPU (src/tsdf.cu:277-457): every host-store block whose centre lies within
 * radius metres of center moves back into the pool (radius = INFINITY: all of them). A block
 * whose key was allocated again while it was out (a frame touched its region) is merged into
 * the resident block voxel by voxel: weights add (capped by weight_cap), the TSDF value and
 * each colour channel take the weighted mean; a side with weight 0 leaves the other as it
 * is. VH_ERR_POOL_FULL, with nothing moved, when the pool cannot hold all selected blocks
 * without merging. A frame only changes blocks that are resident: to reproduce a volume
 * that never streamed, stream in every block within reach of the cameras of the frames to
 * come: (max_depth + trunc) times the longest pixel ray per unit depth (the image corner's,
 * sqrt(1 + ((W-1-cx)/fx)^2 + ((H-1-cy)/fy)^2) at the farther corner), plus one block
 * diagonal. */
int vh_stream_in(vh_volume* vol, const float center[3], float radius, int64_t* n_blocks);
/* Checkpoint (no reference counterpart: the reference's volume lives only in its host
 * chunks): the configuration, the frame counters and every block, resident and in the host
 * store, voxel words and colours byte for byte. Layout (native byte order): char[8]
 * "VHVOLUME", int32 ABI version, int32 32-bit words per block (1024 FP32, 512 16-bit
 * formats), vh_config, uint32 frame id, uint32 0, uint64 frames, int64 n; then n packed
 * block keys (uint64), n blocks of voxel words (in-block order), and for colour volumes n
 * blocks of 512 colour words. */
int vh_save_volume(vh_volume* vol, const char* path);
/* Creates a volume from a checkpoint on device `device` (-1: the saved one). All blocks go
 * into the pool if they fit, otherwise they stay in the host store (vh_stream_in brings them
 * in by region). Integrating after a load gives what integrating without the save would. */
int vh_load_volume(const char* path, int32_t device, vh_volume** out);
/* The volume's configuration (as created or loaded). No reference counterpart. */
int vh_get_config(vh_volume* vol, vh_config* out);

/* ------------------------------------------------------------------ mesh */
/* GpuTsdfGenerator::SavePLY -> tsdf2mesh (src/tsdf.cu:1696-1708,1760-1888), but computed
 * by marching cubes over every allocated block on the GPU rather than gathered from
 * per-frame caches. Cubes need all 8 corners allocated and observed (weight > 0). */
int vh_extract_mesh(vh_volume* vol, vh_mesh* out);
int vh_mesh_free(vh_mesh* mesh);
/* Device variant: writes up to capacity triangles (9 floats each) to d_xyz and, if
 * d_cubes is not NULL, 4 ints each to d_cubes; *n_tris receives the full count
 * (VH_ERR_TRUNCATED if it exceeds capacity). Synchronous. */
int vh_extract_mesh_device(vh_volume* vol, float* d_xyz, int32_t* d_cubes, int64_t capacity,
                           int64_t* n_tris);
/* vh_extract_mesh_device with vertex colours (the reference's Vertex::r,g,b, interpolated as
 * VertexInterp does, src/tsdf.cu:1641-1660): d_rgb (device, may be NULL) receives 9 bytes per
 * triangle. A non-NULL d_rgb needs a colour volume. */
int vh_extract_mesh_device_rgb(vh_volume* vol, float* d_xyz, int32_t* d_cubes, uint8_t* d_rgb,
                               int64_t capacity, int64_t* n_tris);
/* Indexed mesh: one vertex per lattice edge crossed by the surface, shared by every face
 * that uses the edge and interpolated along it from its lower lattice point to its upper one
 * (VertexInterp with that orientation), and per face the indices of its three vertices. The
 * faces are the triangles of vh_extract_mesh in its order and winding, plus the zero-area
 * triangles vh_extract_mesh drops (two corners coincide where the surface passes exactly
 * through a lattice point): with them the mesh of a closed surface is closed, each edge in
 * two faces. Device buffers: d_vertices (vertex_capacity x 3 floats, metres), d_faces
 * (face_capacity x 3 int32), d_vrgb (vertex_capacity x 3 bytes, colour volumes, may be
 * NULL). *n_vertices and *n_faces receive the counts; VH_ERR_TRUNCATED if either exceeds its
 * capacity (nothing is written then). Vertices are numbered by owning block (the block of
 * the edge's lower end point, in pool order), then by axis and in-block index.
 * Synchronous. Replaces the welding step of tsdf2mesh (src/tsdf.cu:1772-1850), which merges
 * vertices with identical float coordinates and so leaves the two interpolations of one edge
 * by neighbouring cubes (in opposite directions) apart when they differ in the last bit. */
int vh_extract_mesh_indexed_device(vh_volume* vol, float* d_vertices, int64_t vertex_capacity,
                                   int32_t* d_faces, int64_t face_capacity, uint8_t* d_vrgb,
                                   int64_t* n_vertices, int64_t* n_faces);
/* Host variant: arrays owned by the library (freed by vh_indexed_mesh_free). */
typedef struct vh_indexed_mesh {
  int64_t n_vertices;
  int64_t n_faces;
  float* vertices;   /* n_vertices x 3 floats, metres */
  int32_t* faces;    /* n_faces x 3 vertex indices */
  uint8_t* rgb;      /* colour volumes: n_vertices x 3 bytes; else NULL */
} vh_indexed_mesh;
int vh_extract_mesh_indexed(vh_volume* vol, vh_indexed_mesh* out);
int vh_indexed_mesh_free(vh_indexed_mesh* mesh);
/* Writes the indexed mesh (vh_extract_mesh_indexed) as binary little-endian PLY: float x, y,
 * z per vertex (colour volumes add uchar red, green, blue), then per face the count 3 and
 * three int32 vertex indices. */
int vh_save_ply(vh_volume* vol, const char* path);

/* ------------------------------------------------------------------ block export */
/* Copies out every allocated block: keys (3 ints each), tsdf and weight (512 floats each, in
 * (x*8+y)*8+z order; 16-bit formats are decoded exactly as the kernels read them). *n receives the block count
 * (VH_ERR_TRUNCATED if it exceeds capacity; nothing is written then). Any pointer may be
 * NULL to skip that array. */
int vh_export_blocks(vh_volume* vol, int32_t* keys, float* tsdf, float* weight,
                     int64_t capacity, int64_t* n);
/* Colour of every allocated block in vh_export_blocks' block and voxel order: 512 x 3 bytes
 * per block (zero where the voxel was never observed). Colour volumes only; *n and
 * VH_ERR_TRUNCATED as vh_export_blocks. */
int vh_export_block_colors(vh_volume* vol, uint8_t* rgb, int64_t capacity, int64_t* n);

/* The marching-cubes tables of MarchingCubeParam (include/tsdf.cuh:243-534): edge_table[c]
 * (256 ints) is the mask of the edges crossed in case c (bit e: edge e's end corners differ
 * in sign), tri_table (256 x 16 ints) case c's triangles as edge indices, padded with -1.
 * Corners and edges are numbered as in include/tsdf.cuh:234-241 and src/tsdf.cu:1001-1037.
 * Either pointer may be NULL. Needs no device. */
int vh_mc_tables(int32_t* edge_table, int32_t* tri_table);

/* ------------------------------------------------------------------ query (extension) */
/* Trilinear SDF at n world points (host arrays, 3 floats each). valid[i] = 1 iff all 8
 * lattice corners are allocated and observed. No reference counterpart. */
int vh_query_sdf(vh_volume* vol, const float* pts, int64_t n, float* sdf, uint8_t* valid);
/* Device variant; asynchronous on the volume's stream. */
int vh_query_sdf_device(vh_volume* vol, const float* d_pts, int64_t n, float* d_sdf,
                        uint8_t* d_valid);
/* Backward of r_i = sdf(R p_i + t) with respect to a left se(3) perturbation
 * xi = (rho, phi) of the pose (T <- exp(xi^) T) and to the voxel tsdf values.
 * d_pts are camera-frame points, c2w the pose (host, row-major 4x4), d_grad_r the upstream
 * gradient dL/dr_i. Outputs (device, each may be NULL):
 *   d_r           n residuals (0 where invalid), d_valid n flags,
 *   d_J           n x 6 per-point dr/dxi,
 *   d_pose_grad   6 doubles: sum_i dL/dr_i * dr_i/dxi,
 *   d_voxel_grad  pool_blocks*512 floats, accumulated (+=) dL/dtsdf per pool voxel.
 * Asynchronous on the volume's stream. */
int vh_query_sdf_backward_device(vh_volume* vol, const float* d_pts, int64_t n,
                                 const float* c2w, const float* d_grad_r, float* d_r,
                                 uint8_t* d_valid, float* d_J, double* d_pose_grad,
                                 float* d_voxel_grad);
/* Gauss-Newton refinement of c2w (host, row-major 4x4, updated in place, FP64) so that the
 * camera-frame points p_i (host, n x 3) land on the zero level set: minimises
 * sum_i sdf(R p_i + t)^2, counting a point without a valid sample as a residual of 1;
 * points with a non-finite coordinate are skipped.
 * Each of the `iters` (<= 63) steps reduces J^T J, J^T r and the cost in FP64 on the GPU
 * and solves the damped 6x6 system (Levenberg-Marquardt, Cholesky) there; a last
 * evaluation picks the best accepted pose. stats (may be NULL) holds iters + 1 entries.
 * No reference counterpart (the reference takes poses as given). */
int vh_gn_refine(vh_volume* vol, const float* pts, int64_t n, double c2w[16], int32_t iters,
                 vh_gn_stats* stats);
/* vh_gn_refine over device-resident points (d_pts, n x 3 floats) that allocates no device
 * memory: no reference counterpart; for callers that cannot allocate (code under HIP graph
 * capture, tight serving or tracking loops, harnesses that forbid hipMalloc). The solver's
 * scratch (partial sums, candidate pose, state, stats rows; 264 KB) is a static device array,
 * one copy per GPU held until the process ends. Limit: iters <= 63 (VH_ERR_INVALID_ARG
 * beyond); any n. Same kernels and results as vh_gn_refine on the same points. Every call on
 * one GPU shares that GPU's scratch: calls that overlap on one GPU (from several threads or
 * volumes) race and must be serialised by the caller. Synchronous. */
int vh_gn_refine_device_noalloc(vh_volume* vol, const float* d_pts, int64_t n, double c2w[16],
                                int32_t iters, vh_gn_stats* stats);

/* Pool layout for callers that wrap the voxel memory (e.g. as a tensor): device pointer of
 * the voxel pool (pool_blocks*512 voxels; FP32: float2 {tsdf, weight}; FP16: uint32
 * {binary16 tsdf | uint16 weight << 16}; F16R: uint32 {binary16 tsdf | uint8 weight << 16 |
 * int8 residual << 24}) and of the block keys (pool_blocks packed keys). */
int vh_pool_view(vh_volume* vol, void** d_voxels, void** d_block_keys, int64_t* n_blocks);

#ifdef __cplusplus
}
#endif

#endif /* VH_H */
