/* vh_synth.h - seeded synthetic depth sequences, part of libvh (host code only).
 *
 * Stands in for the reference's disk frame loader SaveFrame::frameLoad
 * (src/SaveFrame.cpp:154-218), which reads depth/<id>.png (uint16 mm) and tcw/<id>.txt;
 * the reference ships no frames (its .gitignore excludes the scene directory), so the
 * workloads are generated: analytic ray casting of a box room (or a walled multi-room floor
 * plan) with spheres and axis-aligned boxes, seen by a 640x480 pinhole camera
 * (fx = fy = 525, cx = 319.5, cy = 239.5) that moves 2 cm and turns 1 degree per frame.
 * Depth noise is Gaussian with sigma(z) = 0.0012 + 0.0019 (z - 0.4)^2 m, a fixed fraction of
 * pixels drops to 0, and depth is quantised to uint16 millimetres.
 *
 * All randomness comes from a counter-based generator (SplitMix64 of seed, frame, pixel),
 * so any frame can be produced independently and identically on any host.
 */
#ifndef VH_SYNTH_H
#define VH_SYNTH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vh_scene vh_scene;

typedef enum vh_scene_preset {
  VH_SCENE_ROOM_SMALL = 1, /* 4 x 3 x 2.5 m room, 3 spheres + 2 boxes, 40 frames, 1% dropout */
  VH_SCENE_ROOM_MEDIUM = 2,/* 8 x 6 x 3 m room, 12 primitives, 500 frames, 1% dropout */
  VH_SCENE_FLOOR_LARGE = 5 /* 20 x 20 x 4 m, 2 x 2 rooms with 10 cm walls and doorways,
                              40 primitives, 2000 frames, 1.5% dropout */
} vh_scene_preset;

typedef struct vh_scene_info {
  int32_t width, height, n_frames;
  float fx, fy, cx, cy;
  float extent[3];      /* metres */
  int32_t n_spheres, n_boxes;
  float dropout;
} vh_scene_info;

int vh_scene_create(int32_t preset, uint64_t seed, vh_scene** out);
int vh_scene_destroy(vh_scene* s);
int vh_scene_info_get(const vh_scene* s, vh_scene_info* out);
/* Camera pose of frame f (row-major 4x4 camera-to-world, float). */
int vh_scene_pose(const vh_scene* s, int32_t f, float c2w[16]);
/* Noisy uint16 millimetre depth of frame f (width*height values), at the scene's pose or,
 * when c2w is not NULL, at that pose (the frame index still seeds the noise). */
int vh_scene_render_u16(const vh_scene* s, int32_t f, const float* c2w, uint16_t* depth);
/* Noise-free, dropout-free camera-z depth in metres (FP32) at pose c2w (NULL = frame f's). */
int vh_scene_render_clean(const vh_scene* s, int32_t f, const float* c2w, float* depth);
/* Noise-free colour image of frame f (width*height*3 bytes, RGB) at pose c2w (NULL = frame
 * f's): each surface (room face, sphere, box) has a base colour drawn from the seed, on a
 * 10 cm world-aligned checkerboard whose dark cells carry 5/8 of it; 0 where nothing is hit.
 * The reference's frames carry an RGB image beside the depth (SaveFrame.cpp:154-218). */
int vh_scene_render_rgb(const vh_scene* s, int32_t f, const float* c2w, uint8_t* rgb);
/* Frames [f0, f0+n) into depth (n*width*height) and c2w (n*16), multithreaded. */
int vh_scene_render_range(const vh_scene* s, int32_t f0, int32_t n, uint16_t* depth,
                          float* c2w);

#ifdef __cplusplus
}
#endif

#endif /* VH_SYNTH_H */
