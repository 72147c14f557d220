/*
 * oracle.h — CPU restatement of the reference's per-frame dense SLAM hot path.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker or
 * the CPU baseline; the product library (libsicp.so) never links it.
 *
 * Every function is a scalar restatement of the reference file:line it cites,
 * written from reading that source. Images are contiguous row-major host
 * arrays (no pitch); float4 maps are rows*cols*4 floats; RGB is rows*cols*3.
 * The map, fusion, raycast and tracking functions exist in two precisions:
 * `fp64 = 0` keeps the reference's FP32 operation order (including its FP64
 * promotions), `fp64 = 1` computes every floating-point value in double.
 */
#ifndef SICP_ORACLE_H_
#define SICP_ORACLE_H_

#include <stddef.h>
#include <stdint.h>
#include "../include/sicp.h"

#ifdef __cplusplus
extern "C" {
#endif

int oracle_set_threads(int n);       /* OpenMP threads, returns the count in use */
int oracle_max_threads(void);

/* ---- pyramid (Tracking/Pyrdown.cu) ---- */
float oracle_expf(float x);
void oracle_filter_depth(const uint16_t* depth, int cols, int rows, float depth_scale,
                         float depth_cutoff, float* raw, float* filtered);
void oracle_compute_vmap(const float* depth, int cols, int rows, float fx, float fy, float cx,
                         float cy, float depth_cutoff, float* vmap);
void oracle_compute_nmap(const float* vmap, int cols, int rows, float* nmap);
void oracle_pyrdown_f32(const float* src, int scols, int srows, float* dst, int dcols, int drows);
void oracle_pyrdown_u8(const uint8_t* src, int scols, int srows, uint8_t* dst, int dcols,
                       int drows);
void oracle_image_to_intensity(const uint8_t* rgb, int cols, int rows, uint8_t* gray);
void oracle_derivative_image(const uint8_t* gray, int cols, int rows, int16_t* dx, int16_t* dy);
void oracle_resize_map(const float* vsrc, const float* nsrc, int scols, int srows, float* vdst,
                       float* ndst, int dcols, int drows);

/* ---- ICP (Tracking/Reduction.cu:213-385) ---- */
/* Per-pixel rows in FP32 (same order); sums accumulated in FP64 into
 * sums_fp64 (29), and optionally the per-pixel 29-vectors into rows_out
 * (rows*cols*29 floats). */
void oracle_icp_sums(const float* vcurr, const float* vlast, const float* ncurr,
                     const float* nlast, int cols, int rows, const sicp_pose_f* curr,
                     const sicp_pose_f* last, sicp_intrinsics K, double* sums_fp64,
                     float* rows_out);
/* The precision form's sums of per-pixel rows (29 floats a pixel as
 * oracle_icp_sums writes them, slot 28 = 1 where the pixel contributes): FP32
 * with compensation, in the GPU reduction's order (the single-step entry
 * points' partition: small images spread one pixel per thread); compensate =
 * 0: plain FP32 in the same order. */
void oracle_sums_k32(const float* rows29, int n, int compensate, float* sums);
/* The same computed entirely in double from double poses (4x4 row-major). */
void oracle_icp_sums_fp64(const double* vcurr, const double* vlast, const double* ncurr,
                          const double* nlast, int cols, int rows, const double* curr_pose,
                          const double* last_pose, const double* K4, double* sums);
/* RGBStep (Tracking/Reduction.cu:387-642): images u8, vertex maps float4,
 * derivatives int16, one shape. The 29 photometric sums (FP32 rows, FP64
 * accumulation) into sums_fp64; corresp = {count, int sum of squared
 * differences}; sigma as the reference forms it. */
void oracle_rgb_sums(const uint8_t* nimg, const uint8_t* limg, const float* nvmap,
                     const float* lvmap, const int16_t* dx, const int16_t* dy, int cols, int rows,
                     const sicp_pose_f* curr, const sicp_pose_f* last, sicp_intrinsics K,
                     double* sums_fp64, int32_t* corresp, float* sigma);
/* SO3Step (Tracking/Reduction.cu:67-211): 11 sums (FP32 rows, FP64
 * accumulation); rotations curr->Rinv and last->R. */
void oracle_so3_sums(const uint8_t* nimg, const uint8_t* limg, const int16_t* dx,
                     const int16_t* dy, int cols, int rows, const sicp_pose_f* curr,
                     const sicp_pose_f* last, sicp_intrinsics K, double* sums);
void oracle_create_matrix(const float* sums29, double* A, double* b);
int oracle_cholesky_solve6(const double* A, const double* b, double* x);
void oracle_se3_exp(const double* xi, double* T);
int oracle_se3_update(const double* A, const double* b, double* delta, const double* last_pose,
                      double* next_pose, double* x_out);
void oracle_euler_sin_norm(const double* pose_a, const double* pose_b, double* rot_norm,
                           double* trans_norm);

/* ---- map (Mapping/DeviceMap.cu, FuseMap.cu, RenderScene.cu) ---- */
void* oracle_map_create(const sicp_map_config* cfg, int fp64);
void oracle_map_destroy(void* map);
void oracle_map_reset(void* map);
uint32_t oracle_map_hash(const void* map, int x, int y, int z);
/* FuseMapColor (FuseMap.cu:1039-1162). Returns the visible count. */
uint32_t oracle_fuse_map_color(void* map, const float* depth, const uint8_t* rgb,
                               const float* nmap, const uint8_t* mask, int cols, int rows,
                               int num_detection, const sicp_pose_f* pose, sicp_intrinsics K);
/* The same with a double pose for the fp64 map. */
uint32_t oracle_fuse_map_color_fp64(void* map, const double* depth, const uint8_t* rgb,
                                    const double* nmap, const uint8_t* mask, int cols, int rows,
                                    int num_detection, const double* pose, const double* K4);
uint32_t oracle_check_visibility(void* map, const sicp_pose_f* pose, int cols, int rows,
                                 sicp_intrinsics K, float depth_max, float depth_min);
/* CreateRenderingBlocks (RenderScene.cu:210-289). Returns 1 if any tile was touched. */
int oracle_create_rendering_blocks(void* map, float* zmin, float* zmax, int zcols, int zrows,
                                   float depth_max, float depth_min, const sicp_pose_f* pose,
                                   sicp_intrinsics K);
/* Raycast (RenderScene.cu:380-537). */
void oracle_raycast(void* map, float* vmap, float* nmap, int cols, int rows, const float* zmin,
                    const float* zmax, int zcols, const sicp_pose_f* pose, sicp_intrinsics K);
void oracle_map_get_stats(void* map, sicp_map_stats* out);
void oracle_map_export(void* map, sicp_hash_entry* table, sicp_hash_entry* excess,
                       sicp_block* blocks, sicp_voxel_geom* geom, sicp_voxel_sem* sem);
uint32_t oracle_map_visible(void* map, sicp_block* out, uint32_t cap);
/* UnifyNeighbor (Mapping/FuseMap.cu:757-948) over the last visible list,
 * every voxel seeing the labels before the pass. */
void oracle_unify_neighbor(void* map);
/* IEEE binary16 rounding (nearest even) of the precision form, and back. */
uint16_t oracle_f32_to_f16(float f);
float oracle_f16_to_f32(uint16_t h);
/* A new FP64 map with exactly the state of the FP32 map `src`. */
void* oracle_map_clone_fp64(void* src);
/* Raycast of an FP64 map at a double pose (4x4 camera -> world). */
void oracle_raycast_fp64(void* map, double* vmap, double* nmap, int cols, int rows,
                         const float* zmin, const float* zmax, int zcols, const double* pose16,
                         double invfx, double invfy, double cx, double cy);
/* Test hooks: one voxel's sdf. */
double oracle_map_get_sdf(void* map, int slot, int idx);
void oracle_map_set_sdf(void* map, int slot, int idx, double v);
/* Test hook: one voxel's weight and label state on an FP32 map. */
void oracle_map_set_label(void* map, int slot, int idx, int weight, int label, int count,
                          int label2, int count2);
/* MeshScene (Mapping/MeshScene.cu:4-356) on an FP32 map, with the given
 * tables (edge[256], tri[256][16], nvert[256]); returns the triangles
 * written (at most max_tris); `total` gets the number found. */
uint32_t oracle_mesh_scene(void* map, const int32_t* edge, const int8_t* tri,
                           const int32_t* nvert, uint32_t max_tris, float* vertices,
                           float* normals, uint8_t* colors, uint32_t* total);
/* Counters of the last passes: table probes of the allocation, block
 * lookups and march steps of the raycast, and the raycast outcomes (no depth
 * range, back-face exit, no zero crossing, normal rejected, surface found). */
void oracle_map_counters(void* map, uint64_t* out8);

/* ---- per-frame system (Core/System.cc:131-172) ---- */
void* oracle_system_create(const sicp_system_config* cfg, int fp64);
void oracle_system_destroy(void* sys);
int oracle_system_grab(void* sys, const uint16_t* depth, const uint8_t* rgb,
                       const uint8_t* labels, int num_detection, sicp_frame_result* out);
void* oracle_system_map(void* sys);
/* Work summed over the frames grabbed: raycasts, rays, march steps,
 * interpolated reads, raycast block lookups, allocation probes, voxels fused,
 * visible blocks fused. */
void oracle_system_work(void* sys, uint64_t* out8);
/* ComputeSE3's loop decisions (Tracking.cc:398-503) on a given stream of
 * per-iteration ICP errors / counts (and, when rgb_error is non-NULL, the
 * photometric error / count checked first), with the pose replaced by a
 * marker: 0 the initial pose, k + 1 the pose after iteration k's update.
 * out4 = {final marker, failed, updates applied, iterations run};
 * out_error = the final ICP error. Returns -1 if the stream is too short. */
int oracle_se3_control_trace(const int* iters3, const int* min_count3, const float* icp_error,
                             const int* icp_count, const float* rgb_error, const int* rgb_count,
                             float w, int n, int* out4, float* out_error);
/* One ComputeSE3 (point-to-plane only) on given pyramids: next_* the next
 * frame's measured vertex / normal maps, last_* the last frame's model maps,
 * three levels each (float4, cols/2^l x rows/2^l), from `last_pose`
 * (4x4 row-major); `fp64` selects the double transcription (else FP32 in
 * the reference's order, with cfg->icp_sums_fp32 honoured). The pose after
 * tracking (the last pose when lost) into pose_out; returns 1 if tracked. */
int oracle_track_frame(const sicp_system_config* cfg, int fp64, const float* const* next_vmap,
                       const float* const* next_nmap, const float* const* last_vmap,
                       const float* const* last_nmap, const double* last_pose, double* pose_out,
                       sicp_frame_result* res);
/* The final gate of ComputeSE3 (Tracking.cc:505-516): 1 = tracked. */
int oracle_se3_gate(float icp_error, float thresh, const double* pose0, const double* pose);
/* Copies the last frame's level-l model vertex / normal maps (after raycast). */
void oracle_system_model(void* sys, int level, float* vmap, float* nmap);

#ifdef __cplusplus
}
#endif
#endif
