/* vh_io.h - frame and settings loading of libvh (host only, no GPU needed).
 *
 * Replaces the reference's disk loader and settings reader, which go through OpenCV:
 *   - SaveFrame::frameLoad (src/SaveFrame.cpp:154-218) reads <dir>/depth/<id>.png (16-bit
 *     millimetres, cv::imread(-1)), scales it by 0.001 into metres (convertTo + *= 0.001), and
 *     reads <dir>/tcw/<id>.txt (16 numbers, row-major 4x4) which it inverts into mTcw;
 *     PointCloudGenerator::PushFrame inverts it back (src/PointCloudGenerator.cpp:139-143)
 *     before processFrame, so the file's matrix is the camera-to-world pose integrated.
 *     The colour image <dir>/RGB/<id>.jpg is not read (no JPEG decoder here): callers with
 *     decoded colour pass it in vh_frame.rgb.
 *   - PointCloudGenerator's constructor (src/PointCloudGenerator.cpp:20-45) reads
 *     Camera.fx/fy/cx/cy/width/height, DepthMapFactor, MaxDepth, Voxel.Origin.x/y/z,
 *     Voxel.Size, Voxel.TruncMargin and Voxel.Dim.x/y/z with cv::FileStorage from an
 *     OpenCV YAML file (scene0220_02/scene0220_02.yaml).
 * Every function returns a vh_status (vh.h); the text of a failure is in vh_last_error().
 */
#ifndef VH_IO_H
#define VH_IO_H

#include <stdint.h>

#include "vh.h"

#ifdef __cplusplus
extern "C" {
#endif

/* bits of vh_settings.found */
enum {
  VH_SET_FX = 1 << 0, VH_SET_FY = 1 << 1, VH_SET_CX = 1 << 2, VH_SET_CY = 1 << 3,
  VH_SET_WIDTH = 1 << 4, VH_SET_HEIGHT = 1 << 5, VH_SET_DEPTH_FACTOR = 1 << 6, VH_SET_MAX_DEPTH = 1 << 7,
  VH_SET_ORIGIN = 1 << 8, VH_SET_VOX_SIZE = 1 << 9, VH_SET_TRUNC = 1 << 10, VH_SET_VOX_DIM = 1 << 11
};

typedef struct vh_settings {
  float fx, fy, cx, cy;        /* Camera.fx, Camera.fy, Camera.cx, Camera.cy */
  int32_t width, height;       /* Camera.width, Camera.height */
  float depth_factor;          /* DepthMapFactor */
  float max_depth;             /* MaxDepth */
  float origin[3];             /* Voxel.Origin.x/y/z (the hash path ignores it) */
  float vox_size;              /* Voxel.Size */
  float trunc;                 /* Voxel.TruncMargin */
  int32_t vox_dim[3];          /* Voxel.Dim.x/y/z (the hash path ignores it) */
  uint32_t found;              /* VH_SET_* bits of the keys present (ORIGIN / VOX_DIM: all 3) */
} vh_settings;

/* Reads the OpenCV-YAML settings file (a "%YAML:1.0" header, "Key: value" scalar lines, '#'
 * comments). Keys not listed above are ignored; a listed key with a non-numeric value is
 * VH_ERR_INVALID_ARG. Missing keys leave their fields 0 and their bit clear. */
int vh_settings_load(const char* path, vh_settings* out);

/* Fills the camera and voxel fields of cfg from the settings the way the reference's
 * constructor maps them (width, height, fx, fy, cx, cy, max_depth, origin, vox_size, trunc,
 * vox_dim); the other fields of cfg are left as they are. Requires the camera keys,
 * MaxDepth, Voxel.Size and Voxel.TruncMargin. */
int vh_settings_to_config(const vh_settings* s, vh_config* cfg);

/* 16-bit greyscale PNG (bit depth 16, colour type 0, not interlaced): size, then pixels
 * (row-major, host byte order). Other PNG formats are VH_ERR_INVALID_ARG. */
int vh_png16_info(const char* path, int32_t* width, int32_t* height);
int vh_png16_load(const char* path, uint16_t* px, int32_t width, int32_t height);
/* Writes a 16-bit greyscale PNG (filter 0 on every row, zlib level 6). */
int vh_png16_save(const char* path, const uint16_t* px, int32_t width, int32_t height);

/* Reads 16 whitespace-separated numbers (row-major 4x4) from a text file. */
int vh_pose_load(const char* path, float c2w[16]);

/* <dir>/depth/<id>.png into depth (width x height uint16) and <dir>/tcw/<id>.txt into c2w.
 * A missing depth image is VH_ERR_IO (the end of a sequence). */
int vh_frame_load(const char* dir, int32_t id, uint16_t* depth, int32_t width, int32_t height,
                  float c2w[16]);

#ifdef __cplusplus
}
#endif

#endif /* VH_IO_H */
