/*
 * sicp.h — C ABI of the MI355X-native dense SLAM hot path.
 *
 * Every entry point is `extern "C"`, takes plain pointers, sizes and POD
 * descriptors, runs on the caller's HIP stream (`stream`, may be NULL for the
 * null stream) and returns an int status (SICP_OK = 0) instead of exiting the
 * process the way the reference's `SafeCall` does (Utility/SafeCall.h:13-21).
 *
 * Image descriptors (`sicp_image`) describe pitched device buffers the same way
 * the reference's `DeviceArray2D<T>` / `PtrStepSz<T>` do
 * (Utility/DeviceArray.h:31-41,153-162): row y starts at
 * `(char*)ptr + y * pitch`.
 *
 * Poses follow the reference convention (Core/Frame.cc:626-645): a camera to
 * world transform with row-major rotation R, its transpose Rinv and
 * translation t, all float. Entry points that fuse, raycast or track also
 * accept the pose through a device pointer (`sicp_pose_f`), so a whole frame
 * can run with no host round trip.
 */
#ifndef SICP_H_
#define SICP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
    SICP_OK = 0,
    SICP_ERR_INVALID_ARGUMENT = 1, /* shape / pitch / config mismatch        */
    SICP_ERR_HIP = 2,              /* a HIP runtime call failed               */
    SICP_ERR_CAPACITY = 3,         /* heap, request set or buffer exhausted   */
    SICP_ERR_NO_VISIBLE = 4,       /* no visible block (FuseMap.cu:1117-1120)  */
    SICP_ERR_NOTHING_TO_RENDER = 5,/* CreateRenderingBlocks returned false    */
    SICP_ERR_NO_DEVICE = 6         /* no HIP device present                   */
};

/* Last HIP / argument error message of the calling thread (never NULL). */
const char* sicp_last_error(void);

/* ---- POD descriptors ---------------------------------------------------- */
typedef struct sicp_image {
    void* ptr;      /* device pointer to row 0                           */
    size_t pitch;   /* bytes between rows                                */
    int cols;       /* width in elements                                 */
    int rows;       /* height in rows                                    */
} sicp_image;

typedef struct sicp_intrinsics {
    float fx, fy, cx, cy; /* Utility/Intrinsics.h:3-18 */
} sicp_intrinsics;

/* Float view of one pose: R row-major, Rinv = R^T, t. */
typedef struct sicp_pose_f {
    float R[9];
    float Rinv[9];
    float t[3];
} sicp_pose_f;

/* ---- map layout (public so callers can inspect a downloaded map) -------- */

/* Geometry plane, one record per voxel, 8 B: what the raycast reads.
 * Fields of the reference's `Voxel` sdf / weight / color
 * (Mapping/DeviceMap.h:109-111). */
typedef struct sicp_voxel_geom {
    float sdf;        /* normalised by the truncation, NaN when empty      */
    uint8_t weight;   /* 0 = empty                                          */
    uint8_t r, g, b;  /* running-mean colour                                */
} sicp_voxel_geom;

/* Semantic plane, one record per voxel, 12 B: the reference's label fields
 * (Mapping/DeviceMap.h:113-118) with the counts widened to uint16. */
typedef struct sicp_voxel_sem {
    uint16_t count, count2;
    uint8_t label, label2;
    uint8_t r2, g2, b2;
    uint8_t pad[3];
} sicp_voxel_sem;

/* One hash-table entry, 16 B. `key` packs the block coordinate
 * (21 bits per axis, two's complement); `slot` is the heap block (-1 =
 * empty); `next` links the bucket's overflow chain (-1 = end). The entry
 * replaces the reference's `HashEntry {int3 pos; int ptr; int offset}`
 * (Mapping/DeviceMap.h:20-60). */
typedef struct sicp_hash_entry {
    uint64_t key;
    int32_t slot;
    int32_t next;
} sicp_hash_entry;

/* One allocated / visible block: block coordinate and heap slot. Plays the
 * role of the reference's `visibleEntries` records (DeviceMap.h:198). */
typedef struct sicp_block {
    int32_t x, y, z, slot;
} sicp_block;

/* Runtime map configuration. The reference hard-codes these as
 * `DeviceMap` constants (Mapping/DeviceMap.h:166-189). */
typedef struct sicp_map_config {
    float voxel_size;       /* metres (VoxelSize, DeviceMap.h:183)          */
    float truncation;       /* metres (TruncateDist, DeviceMap.h:184)       */
    float depth_min;        /* metres (DepthMin, DeviceMap.h:168)           */
    float depth_max;        /* metres (DepthMax, DeviceMap.h:169)           */
    int32_t num_buckets;    /* buckets of 2 entries each                    */
    int32_t num_blocks;     /* heap capacity in 8^3 blocks                  */
    int32_t max_weight;     /* TSDF weight cap, 255 in the reference        */
    int32_t max_label_count;/* label vote cap: 255 reference, 65535 widened */
    int32_t request_capacity; /* per-frame new-block request set, pow2     */
    int32_t max_cols;       /* largest image width the map is used with     */
    int32_t max_rows;       /* largest image height                         */
    int32_t unify_labels;   /* 1: every fusion with labels is followed by the
                               neighbour label pass (UnifyNeighbor, as the
                               reference's FuseMapColor runs it,
                               FuseMap.cu:1140-1160); 0 (default): not run */
    int32_t tsdf_half;      /* 1: the precision form: each voxel's sdf is
                               stored as IEEE half (round to nearest even)
                               beside its weight in one 4-byte word, the
                               colour in the semantic record; every update and
                               interpolation is FP32. 0 (default): FP32 sdf.
                               Downloads return the FP32 record either way. */
} sicp_map_config;

/* Fills `cfg` with the reference's constants scaled to the given voxel size
 * and truncation; buckets/blocks as given. */
void sicp_map_config_default(sicp_map_config* cfg, float voxel_size, float truncation,
                             int32_t num_buckets, int32_t num_blocks);

typedef struct sicp_map_stats {
    uint32_t allocated_blocks;  /* heap slots handed out so far              */
    uint32_t visible_blocks;    /* last visibility pass                      */
    uint32_t new_blocks;        /* blocks created by the last fusion         */
    uint32_t dropped_heap;      /* requests lost to heap exhaustion (total)  */
    uint32_t dropped_requests;  /* requests lost to request-set overflow     */
    uint32_t dropped_range;     /* block coordinates outside 21-bit range    */
    uint32_t updated_voxels;    /* voxels written by the last fusion         */
    uint32_t render_tiles;      /* blocks that produced tiles, last prepass  */
    uint32_t ray_handover;      /* rays the last raycast marched in its second pass */
    uint64_t updated_total;     /* voxels written, summed since reset        */
    uint64_t visible_total;     /* visible blocks fused, summed since reset  */
    uint64_t fusions;           /* fusion passes since reset                 */
    uint64_t heap_scanned_total;/* allocated blocks at each fusion (the heap
                                   slots its visibility pass scanned), summed
                                   since reset                              */
} sicp_map_stats;

typedef struct sicp_map sicp_map; /* opaque; owns every map buffer (Mapping.cc:11-51) */

/* ---- device / library --------------------------------------------------- */
int sicp_device_count(int* count);
const char* sicp_version(void);

/* ---- depth / intensity pyramid (Tracking/Reduction.h:18-42) -------------- */

/* FilterDepth (Tracking/Pyrdown.cu:72-121): uint16 depth -> metres in `raw`
 * (NaN at or beyond the cutoff) and the 5x5 bilateral `filtered`. */
int sicp_filter_depth(sicp_image depth_u16, sicp_image raw_f32, sicp_image filtered_f32,
                      float depth_scale, float depth_cutoff, void* stream);

/* ComputeVMap (Pyrdown.cu:123-158): depth -> camera-frame float4 vertices. */
int sicp_compute_vmap(sicp_image depth_f32, sicp_image vmap_f4, float fx, float fy,
                      float cx, float cy, float depth_cutoff, void* stream);

/* ComputeNMap (Pyrdown.cu:160-192): forward-difference normals. */
int sicp_compute_nmap(sicp_image vmap_f4, sicp_image nmap_f4, void* stream);

/* PyrDownGauss (Pyrdown.cu:194-248), float and uint8 variants. */
int sicp_pyrdown_gauss_f32(sicp_image src, sicp_image dst, void* stream);
int sicp_pyrdown_gauss_u8(sicp_image src, sicp_image dst, void* stream);

/* ImageToIntensity (Pyrdown.cu:250-272): uchar3 RGB -> uint8. */
int sicp_image_to_intensity(sicp_image rgb_u8x3, sicp_image gray_u8, void* stream);

/* ComputeDerivativeImage (Pyrdown.cu:274-316): flipped Sobel -> int16 dx, dy. */
int sicp_compute_derivative_image(sicp_image gray_u8, sicp_image dx_i16, sicp_image dy_i16,
                                  void* stream);

/* ResizeMap (Pyrdown.cu:318-358): 2x2 mean of vertex / normal maps. */
int sicp_resize_map(sicp_image vsrc, sicp_image nsrc, sicp_image vdst, sicp_image ndst,
                    void* stream);

/* UnifyNeighbor (Fusion::UnifyNeighbor / UpdateNeighborLabel,
 * Mapping/FuseMap.cu:757-948): every labelled voxel of the map's visible
 * blocks takes the most frequent label among its 26 neighbours (those with
 * a non-zero count), keeping its previous label and count as the secondary
 * pair when the label changes. Deterministic: every voxel sees the labels of
 * the state before the pass. Runs on the map's last visible list; `stream`
 * ordered. Device scratch of num_blocks x 2 KiB is allocated on the first
 * call unless the map was created with unify_labels = 1. */
int sicp_unify_neighbor_labels(sicp_map* map, void* stream);

/* ---- Raycast adjoint ------------------------------------------------------ */

/* The derivative of Raycast's output (RenderScene.cu:380-488) for sampled
 * pixels. No counterpart in the reference: it serves callers that optimise
 * a pose or the TSDF through the renderer. The vertex and normal of each
 * pixel are functions of the camera pose and of the TSDF values, with the
 * zRange images and the march's discrete decisions held fixed; the march is
 * recomputed per pixel (nothing is stored by the forward pass). Given
 * d loss / d vertex (`grad_vertex_dev`, n x 3, camera frame, metres) and
 * d loss / d normal (`grad_normal_dev`, n x 3, may be NULL) of the pixels
 * `pixels_dev` (n indices y * cols + x; others contribute nothing):
 *  - `grad_pose_dev` (6 doubles, device) gets d loss / d xi for the camera
 *    perturbation T' = T exp(xi), xi = (translation, rotation), camera frame;
 *  - `grad_pose_px_dev` (n x 6 floats, may be NULL) the same per pixel;
 *  - `grad_sdf_dev` (num_blocks x 512 floats, heap-slot order, may be NULL)
 *    accumulates d loss / d sdf (normalised units); the caller zeroes it.
 * Pixels whose forward output is NaN contribute 0. Arguments as
 * sicp_raycast_inv; `scratch_dev` holds sicp_raycast_backward_scratch_size()
 * bytes, 8-byte aligned. TSDF gradients are summed with atomics (their
 * rounding depends on the order); the pose sums are in a fixed order. */
int sicp_raycast_backward(sicp_map* map, sicp_image zmin, sicp_image zmax,
                          const sicp_pose_f* pose, float invfx, float invfy, float cx, float cy,
                          int32_t cols, int32_t rows, const int32_t* pixels_dev,
                          int32_t n_pixels, const float* grad_vertex_dev,
                          const float* grad_normal_dev, float* grad_sdf_dev,
                          float* grad_pose_px_dev, double* grad_pose_dev, void* scratch_dev,
                          size_t scratch_bytes, void* stream);
size_t sicp_raycast_backward_scratch_size(void);

/* ---- Mesh extraction (Mapping/MeshScene.cu:4-356) ------------------------ */

/* MeshScene + Mapping::CreateModel (MeshScene.cu:308-356, Mapping.cc:183-192):
 * marching cubes over every allocated block. A voxel cell makes triangles
 * when its eight corners are neither empty (NaN) nor exactly 1 and the
 * voxels either side of every corner along each axis exist. Per triangle,
 * three vertices of `vertices_dev` / `normals_dev` (float x, y, z each, in
 * metres / unit length) and `colors_dev` (uchar r, g, b each). The order is
 * canonical: heap slot, voxel z*64 + y*8 + x, triangle table order.
 * Triangles past `max_triangles` are dropped (MaxTriangles,
 * DeviceMap.h:181); `n_triangles` gets the number written and `n_total`
 * (may be NULL) the number found. With max_triangles = 0 the call only
 * counts. Synchronises `stream`. */
int sicp_mesh_scene(sicp_map* map, float* vertices_dev, float* normals_dev, uint8_t* colors_dev,
                    uint32_t max_triangles, uint32_t* n_triangles, uint32_t* n_total,
                    void* stream);
/* The three tables MeshScene takes (edge, triangle and vertex-count tables,
 * Mapping/Constant.h; Mapping.cc builds them): edge_table[256] (bit j: edge j
 * cut), triangle_table[256][16] (edge triples, -1 terminated),
 * vertex_table[256] (3 x triangles). Host copies; any may be NULL. */
int sicp_mesh_tables(int32_t* edge_table, int8_t* triangle_table, int32_t* vertex_table);
/* System::WriteMeshToDisk (Core/System.cc:272-325): ASCII PLY with one
 * vertex per triangle corner (x y z nx ny nz r g b) and one face per
 * triangle, from host arrays as sicp_mesh_scene writes them. */
int sicp_mesh_write_ply(const char* path, const float* vertices, const float* normals,
                        const uint8_t* colors, uint32_t n_triangles);

/* ---- ICP reduction (Tracking/Reduction.h:74-79) -------------------------- */

/* ICPStep (Tracking/Reduction.cu:328-385). Computes the 29 sums (27 upper
 * A|b entries, sum r^2, count) of the point-to-plane Gauss-Newton system into
 * `sums_dev` (29 floats, device). `partials_dev` is device scratch of
 * `sicp_icp_partials_size()` bytes, 16-byte aligned (per-workgroup partial
 * sums, FP64). The
 * sums are bitwise reproducible for a given image size. When `host_sums` is non-NULL the call
 * synchronises `stream` and copies the 29 sums to the host; when `A_host`,
 * `b_host`, `residual_host` are non-NULL it also expands them exactly like
 * the reference's CreateMatrix<6,7> (Reduction.cu:6-16): A row-major 6x6
 * (mirrored), b[6], residual = {sum r^2, count}. */
int sicp_icp_step(sicp_image vcurr, sicp_image vlast, sicp_image ncurr, sicp_image nlast,
                  const sicp_pose_f* curr, const sicp_pose_f* last, sicp_intrinsics K,
                  void* partials_dev, float* sums_dev, float* host_sums,
                  double* A_host, double* b_host, float* residual_host, void* stream);
size_t sicp_icp_partials_size(void);
/* ICPStep with no workspace argument and no allocation: the same kernels,
 * arguments (less `partials_dev`) and results as sicp_icp_step, with the
 * workgroups' partial rows in a 116 KiB array the library holds on each GPU
 * for the life of the process. No counterpart in the reference; it is for
 * callers that cannot allocate device memory (code under hipGraph capture,
 * serving loops, harnesses that forbid hipMalloc). Every image size up to
 * 2^30 pixels fits; larger ones are rejected (SICP_ERR_INVALID_ARGUMENT).
 * The call makes no hipMalloc / hipFree and no allocating library call; the
 * array is found once per device (hipGetSymbolAddress) on the first call
 * there. All calls on one GPU share its array: calls that overlap in time on
 * one GPU (other streams, other threads) race and must be serialised by the
 * caller. */
int sicp_icp_step_static(sicp_image vcurr, sicp_image vlast, sicp_image ncurr, sicp_image nlast,
                         const sicp_pose_f* curr, const sicp_pose_f* last, sicp_intrinsics K,
                         float* sums_dev, float* host_sums, double* A_host, double* b_host,
                         float* residual_host, void* stream);
/* ICPStep in the precision form: the same 29 sums, arguments and outputs as
 * sicp_icp_step, accumulated in FP32 with compensated summation instead of
 * FP64: every addition (a thread adding a product, partial sums meeting in
 * the workgroup and across workgroups) is Knuth's TwoSum, the error terms
 * are summed in FP32 and each value is rounded once at the end (cascaded
 * summation) — the reference's own sums are plain FP32 (BlockReduce /
 * Reduce<float, 29>, Reduction.cu:29-60, 315, 374). Bitwise reproducible for
 * a given image size. */
int sicp_icp_step_fp32(sicp_image vcurr, sicp_image vlast, sicp_image ncurr, sicp_image nlast,
                       const sicp_pose_f* curr, const sicp_pose_f* last, sicp_intrinsics K,
                       void* partials_dev, float* sums_dev, float* host_sums,
                       double* A_host, double* b_host, float* residual_host, void* stream);

/* RGBStep (Tracking/Reduction.h:81-90, Reduction.cu:575-642): the
 * photometric term of one Gauss-Newton iteration. A correspondence pass
 * (FindCorresp) gives the count and the int sum of squared intensity
 * differences (`corresp_host`, the reference's outRes) and sigma =
 * sqrt(sum / max(count, 1)) (`sigma_host`); the row pass (GetRow) then gives
 * the 29 robust-weighted sums into `sums_dev` (29 floats, device), as for
 * sicp_icp_step. Images: intensities (u8) and derivatives (int16) of the next
 * frame, the last frame's intensity, both frames' vertex maps; one shape.
 * `scratch_dev` holds `scratch_bytes` >= sicp_rgb_scratch_size(cols, rows)
 * bytes, 16-byte aligned. Host outputs may be NULL; any non-NULL one
 * synchronises `stream`. */
int sicp_rgb_step(sicp_image next_image, sicp_image last_image, sicp_image next_vmap,
                  sicp_image last_vmap, sicp_image didx, sicp_image didy,
                  const sicp_pose_f* curr, const sicp_pose_f* last, sicp_intrinsics K,
                  void* scratch_dev, size_t scratch_bytes, float* sums_dev, float* host_sums,
                  int32_t* corresp_host, float* sigma_host, double* A_host, double* b_host,
                  float* residual_host, void* stream);
size_t sicp_rgb_scratch_size(int32_t cols, int32_t rows);

/* SO3Step (Tracking/Reduction.h:67-72, Reduction.cu:67-211): the
 * rotation-only photometric alignment of Tracker::ComputeSO3
 * (Tracking.cc:322-366). Every pixel's bearing in the last frame is rotated
 * into the current camera (curr->Rinv * last->R) and matched by rounded
 * projection; the 11 sums (A 3x3 upper and b interleaved, sum r^2, count) go
 * to `sums_dev` (11 floats, device). `partials_dev` is device scratch of
 * sicp_icp_partials_size() bytes, 16-byte aligned. Host outputs as for
 * sicp_icp_step, with CreateMatrix<3,4>: A row-major 3x3 (mirrored), b[3]. */
int sicp_so3_step(sicp_image next_image, sicp_image last_image, sicp_image didx,
                  sicp_image didy, const sicp_pose_f* curr, const sicp_pose_f* last,
                  sicp_intrinsics K, void* partials_dev, float* sums_dev, float* host_sums,
                  double* A_host, double* b_host, float* residual_host, void* stream);

/* FP64 6x6 solve + SE(3) update of one Gauss-Newton iteration, on the host
 * arrays of Tracking/Tracking.cc:493-498: x = A^-1 b (Cholesky),
 * delta = exp(x) * delta, next = last * delta^-1. Poses are 4x4 row-major. */
int sicp_se3_update_host(const double* A, const double* b, double* delta,
                         const double* last_pose, double* next_pose, double* x_out);

/* ---- map (Mapping/RenderScene.h:3-55) ------------------------------------ */
int sicp_map_create(const sicp_map_config* cfg, sicp_map** out);
int sicp_map_destroy(sicp_map* map);
int sicp_map_get_config(const sicp_map* map, sicp_map_config* cfg);

/* ResetMap (FuseMap.cu:1291-1304). */
int sicp_map_reset(sicp_map* map, void* stream);

/* FuseMapColor (FuseMap.cu:1039-1162): allocate the blocks of the +-T/2 band
 * around every valid raw-depth pixel, compact the visible blocks, then fuse
 * TSDF + colour and, when num_detection > 0, the label votes from `mask`.
 * Pose through `pose_host` (by value) or `pose_dev` (device pointer, used when
 * non-NULL). `visible_host`, when non-NULL, receives the visible count (the
 * call then synchronises `stream`). Returns SICP_ERR_NO_VISIBLE when that
 * count is zero and `visible_host` was given. */
int sicp_fuse_map_color(sicp_map* map, sicp_image depth_f32, sicp_image rgb_u8x3,
                        sicp_image nmap_f4, sicp_image mask_u8, int num_detection,
                        const sicp_pose_f* pose_host, const sicp_pose_f* pose_dev,
                        sicp_intrinsics K, uint32_t* visible_host, void* stream);

/* CheckBlockVisibility (FuseMap.cu:991-1037): visibility compaction only. */
int sicp_check_block_visibility(sicp_map* map, const sicp_pose_f* pose_host,
                                const sicp_pose_f* pose_dev, int cols, int rows,
                                sicp_intrinsics K, float depth_max, float depth_min,
                                uint32_t* visible_host, void* stream);

/* CreateRenderingBlocks (RenderScene.cu:210-289): per-8x8-tile depth range of
 * the visible blocks into zmin/zmax (ceil(W/8) x ceil(H/8) float images).
 * `nonempty_host`, when non-NULL, receives 1 if any tile was touched (the
 * call then synchronises). */
int sicp_create_rendering_blocks(sicp_map* map, sicp_image zmin, sicp_image zmax,
                                 float depth_max, float depth_min,
                                 const sicp_pose_f* pose_host, const sicp_pose_f* pose_dev,
                                 sicp_intrinsics K, int* nonempty_host, void* stream);

/* Raycast (RenderScene.cu:495-537): march every pixel through the TSDF and
 * write camera-frame vertices / normals (NaN where no surface). Takes fx, fy
 * like its siblings; the reference's invfx = 1/fx is derived inside. */
int sicp_raycast(sicp_map* map, sicp_image vmap_f4, sicp_image nmap_f4, sicp_image zmin,
                 sicp_image zmax, const sicp_pose_f* pose_host, const sicp_pose_f* pose_dev,
                 sicp_intrinsics K, void* stream);
/* Raycast with the reference's own argument set (Mapping/RenderScene.h:13-18):
 * the inverse focal lengths exactly as the caller computed them. */
int sicp_raycast_inv(sicp_map* map, sicp_image vmap_f4, sicp_image nmap_f4, sicp_image zmin,
                     sicp_image zmax, const sicp_pose_f* pose_host, const sicp_pose_f* pose_dev,
                     float invfx, float invfy, float cx, float cy, void* stream);
/* Device pointer and capacity (records) of the map's visible list, which
 * CheckBlockVisibility / FuseMapColor fill and CreateRenderingBlocks reads
 * (the reference's DeviceMap::visibleEntries, Mapping/DeviceMap.h:198). */
int sicp_map_visible_list(sicp_map* map, sicp_block** list_dev, uint32_t* capacity);

int sicp_map_get_stats(sicp_map* map, sicp_map_stats* out, void* stream);

/* Map inspection (synchronous device->host copies). Buffers sized by the
 * config: table = 2*num_buckets entries, excess / blocks = num_blocks,
 * voxels = num_blocks*512. Any pointer may be NULL. */
int sicp_map_download(sicp_map* map, sicp_hash_entry* table, sicp_hash_entry* excess,
                      sicp_block* blocks, sicp_voxel_geom* geom, sicp_voxel_sem* sem,
                      uint32_t n_blocks_limit);
/* Visible list of the last visibility pass (up to `cap` records). */
int sicp_map_download_visible(sicp_map* map, sicp_block* visible, uint32_t cap,
                              uint32_t* count);

/* Map snapshot (System::WriteMapToDisk / ReadMapFromDisk,
 * Core/System.cc:327-403): a versioned file holding the hash table, the used
 * heap blocks (overflow entries, block records, both voxel planes), the
 * visible list and the counters. Loading resets `map` first and requires
 * the same voxel size, truncation, depth range, caps and bucket count, and
 * a heap at least as large as the snapshot's used part. Synchronous. */
int sicp_map_save(sicp_map* map, const char* path);
int sicp_map_load(sicp_map* map, const char* path);

/* ---- per-frame system (Core/System.cc:131-172) --------------------------- */
typedef struct sicp_system_config {
    int32_t cols, rows;
    sicp_intrinsics K;       /* level-0 intrinsics (SysDesc, MainTum.cc:40-51) */
    float depth_scale;       /* raw units per metre (Frame::mDepthScale)       */
    float depth_cutoff;      /* metres (Frame::mDepthCutoff)                   */
    int32_t icp_iters[3];    /* ITERATIONS_SE3 = {10, 5, 3} (Tracking.h:110)   */
    int32_t icp_min_count[3];/* MIN_ICP_COUNT = {2000,1000,100} (Tracking.h:112) */
    float icp_thresh;        /* THRESH_ICP_SE3 = 1e-4 (Tracking.h:117)         */
    sicp_map_config map;
    int32_t icp_only;        /* 1: ComputeSE3(true, ...), point-to-plane only;
                                0: ComputeSE3(false, ...), the photometric term
                                of RGBStep added (TrackFrame, Tracking.cc:170) */
    float rgb_weight;        /* w of the photometric term, 1e-4 (Tracking.cc:392) */
    int32_t icp_sums_fp32;   /* 0: the Gauss-Newton sums (ICP and photometric)
                                accumulated in FP64; 1: the precision form, FP32
                                with compensated summation (TwoSum at every
                                addition, as sicp_icp_step_fp32). The reference
                                sums in plain FP32 (BlockReduce / Reduce<float,
                                29>, Reduction.cu:29-60, 315, 374). */
    int32_t icp_spread;      /* 1 (default): pyramid levels of at most 320 x 256
                                pixels take one pixel per thread in the ICP
                                reduction (four times the workgroups on a GPU
                                the level alone would leave mostly idle); 0:
                                four pixels per thread at every level, for
                                systems run together (sicp_batch_*) whose
                                launches already fill the GPU. The sums' order
                                depends on it, so runs compared bit for bit
                                use the same value. */
} sicp_system_config;

typedef struct sicp_system sicp_system;

typedef struct sicp_frame_result {
    int32_t frame_index;
    int32_t tracked;        /* 1 tracked (or first frame), 0 lost              */
    int32_t icp_count;      /* valid correspondences of the last iteration     */
    int32_t iterations;     /* Gauss-Newton iterations actually applied        */
    float icp_error;        /* sqrt(sum r^2) / count                           */
    uint32_t visible_blocks;
    uint32_t allocated_blocks;
    double pose[16];        /* camera -> world, row-major                      */
    float rgb_error;        /* photometric term: sqrt(sum r^2) / count of the
                               last iteration (0 when the term is off)         */
    int32_t rgb_count;      /* its correspondences                             */
} sicp_frame_result;

void sicp_system_config_default(sicp_system_config* cfg, int cols, int rows, float fx,
                                float fy, float cx, float cy, float voxel_size,
                                float truncation, int32_t num_buckets, int32_t num_blocks);
int sicp_system_create(const sicp_system_config* cfg, sicp_system** out);
int sicp_system_destroy(sicp_system* sys);
/* GrabImage: track the frame against the model, fuse it, raycast the model.
 * Inputs are device buffers (uint16 depth, uchar3 RGB, uint8 labels or NULL).
 * Asynchronous on `stream`; results through sicp_system_last_result. */
int sicp_system_grab_image(sicp_system* sys, sicp_image depth_u16, sicp_image rgb_u8x3,
                           sicp_image labels_u8, int num_detection, void* stream);
int sicp_system_last_result(sicp_system* sys, sicp_frame_result* out, void* stream);
sicp_map* sicp_system_map(sicp_system* sys);
/* Stage timing with HIP events on the frame's stream (no counterpart in the
 * reference, whose only timing is std::clock around GrabImage,
 * Core/System.cc:142-146). Stages: 0 pyramid, 1 track, 2 allocate,
 * 3 visibility, 4 fuse kernel, 5 render prepass, 6 raycast (both passes),
 * 7 fusion count, 8 whole frame, 9 each ICP iteration launch (nested in 1),
 * 10 tracking finish (nested in 1). read_timing synchronises `stream`, returns per-stage
 * summed milliseconds and instance counts of the frames grabbed since
 * set_timing(1), and clears them; read_timing_detail (call it first) gives
 * one stage's per-instance milliseconds in grab order without clearing. */
int sicp_system_set_timing(sicp_system* sys, int enable);
int sicp_system_read_timing(sicp_system* sys, double* total_ms, int* counts, int n_stages,
                            void* stream);
int sicp_system_read_timing_detail(sicp_system* sys, int stage, double* ms_out, int cap,
                                   int* count, void* stream);
/* Graph form of GrabImage (no counterpart in the reference). The system owns
 * staged input images (uint16 depth, uchar3 RGB, uint8 labels); the caller
 * fills them, e.g. with device-to-device copies on `stream`, and calls
 * grab_staged. From the second frame on, with `use_graph`, the whole frame
 * (pyramid, 18 ICP launches, allocation, fusion, prepass, raycast) is one
 * hipGraph launch: two graphs, one per frame-buffer parity, recorded on first
 * use and again when num_detection changes. Results equal grab_image's. */
int sicp_system_staging(sicp_system* sys, sicp_image* depth_u16, sicp_image* rgb_u8x3,
                        sicp_image* labels_u8);
int sicp_system_grab_staged(sicp_system* sys, int num_detection, int use_graph, void* stream);
/* Fills the staged inputs from a frame with one copy kernel on `stream` (no
 * counterpart in the reference): depth (u16) and rgb (u8x3) of the system's
 * size, and labels (u8) unless labels_u8 or its ptr is NULL. Rows are copied
 * 16 bytes per lane where their addresses, pitches and length allow. */
int sicp_system_stage(sicp_system* sys, const sicp_image* depth_u16, const sicp_image* rgb_u8x3,
                      const sicp_image* labels_u8, void* stream);

/* GrabImage of several independent systems (sequences of any resolutions)
 * as one batch, on one stream (no counterpart in the reference, whose System
 * serves one camera). Each system's pyramid, then the tracking of every
 * system past its first frame in shared launches — one per Gauss-Newton
 * iteration, each system on its own range of workgroups with exactly the
 * arithmetic of its own launches — then each system's fusion and raycast.
 * Results equal n separate sicp_system_grab_image calls. At most 8 systems,
 * each distinct, point-to-plane tracking (icp_only = 1) with one form of the
 * sums and the same iteration counts; labels_u8 and num_detection may be NULL
 * (no labels). Stage timing is not recorded for batched frames. */
int sicp_system_grab_batch(sicp_system* const* systems, int32_t n, const sicp_image* depth_u16,
                           const sicp_image* rgb_u8x3, const sicp_image* labels_u8,
                           const int32_t* num_detection, void* stream);
/* The graph form of batches: a batch holds the systems (not owned) and, once
 * every system is past its first frame, runs a round — one frame of every
 * system, from each system's staged inputs (sicp_system_staging) — as one
 * hipGraph launch (two graphs, one per frame-buffer parity pattern, recorded
 * on first use and again when the patterns or num_detection change). Results
 * equal the eager rounds. */
typedef struct sicp_batch sicp_batch;
int sicp_batch_create(sicp_system* const* systems, int32_t n, sicp_batch** out);
int sicp_batch_destroy(sicp_batch* batch);
int sicp_batch_grab_staged(sicp_batch* batch, const int32_t* num_detection, int use_graph,
                           void* stream);
/* sicp_system_stage for every system of the batch, one kernel for the round:
 * system i's frame from entry i of each array (labels_u8 NULL, or an entry
 * with a NULL ptr: no labels for that system). */
int sicp_batch_stage(sicp_batch* batch, const sicp_image* depth_u16, const sicp_image* rgb_u8x3,
                     const sicp_image* labels_u8, void* stream);
/* shared = 1 (default): the tracking iterations of the batch's systems share
 * launches, as above. shared = 0: each system's whole frame, its own
 * tracking launches included, runs on a stream of its own, so a round is n
 * independent chains (parallel branches of the round's graph) joined at its
 * end. Results are the same either way; the graphs are recorded again. */
int sicp_batch_set_shared_tracking(sicp_batch* batch, int shared);

/* Launch shape of the system's raycast (Mapping::RayTrace, Mapping.cc:142-161;
 * the reference's RayCastKernel launch is fixed at 4 x 8 threads per block,
 * RenderScene.cu:528-533) and its block lookup cache: the samples one
 * first-pass lane marches before handing its ray to the second pass, and the
 * second pass's 256-thread workgroups. Every valid choice gives bit-identical
 * model maps. Defaults 48, 1024 and 1; set_tuning takes 1..4096, 1..65536, 0/1.
 * A change re-records the system's (and its batches') graphs on next use. */
typedef struct sicp_tuning {
    int32_t ray_first_pass_samples;
    int32_t ray_second_pass_workgroups;
    int32_t ray_lds_block_table; /* 1 (default): the first pass's block lookups go
                                    through a per-wavefront LDS table; 0: only each
                                    lane's one-entry cache (DeviceMap.cu:115-133).
                                    The autotuner keeps the value set here. */
} sicp_tuning;
int sicp_system_get_tuning(const sicp_system* sys, sicp_tuning* out);
int sicp_system_set_tuning(sicp_system* sys, const sicp_tuning* tuning);
/* The autotuner (no counterpart in the reference). Each call re-runs the
 * last frame's raycast into scratch maps under each of 28 candidate shapes
 * (the default 48 / 1024 first, then 16..128 samples x 256..2048
 * workgroups), `repeats` (1..64) interleaved rounds, and adds each
 * candidate's median time to the system's running sums; the candidate with
 * the lowest sum becomes the tuning. Call it on several frames of a sequence
 * to choose for the sequence. Reports the chosen and the default shape's
 * mean time per call (microseconds; either pointer may be NULL). The
 * frame's own maps are untouched. Needs a processed frame that was tracked
 * and rendered; synchronises `stream`; allocates scratch for the call.
 * autotune_reset clears the sums, as set_tuning does when it changes the
 * block-table setting; autotune_results lists every candidate with its mean
 * (up to `cap`; `count` = 28, or 0 before any call). */
int sicp_system_autotune(sicp_system* sys, int32_t repeats, sicp_tuning* chosen,
                         float* chosen_us, float* default_us, void* stream);
int sicp_system_autotune_reset(sicp_system* sys);
int sicp_system_autotune_results(const sicp_system* sys, sicp_tuning* shapes, float* mean_us,
                                 int32_t cap, int32_t* count);

/* Frame buffers of the system, for inspection (the public members of Frame,
 * Core/Frame.h): which = 0 vmap, 1 nmap, 2 raw depth (level 0 only),
 * 3 filtered depth, 4 intensity, 5 d/dx, 6 d/dy at pyramid `level`, of the
 * frame just grabbed (its level-0 maps hold the raycast model); add 16 for the
 * frame before it, whose levels 1-2 hold its model pyramid. */
int sicp_system_frame_image(sicp_system* sys, int which, int level, sicp_image* out);

/* ---- seeded synthetic room (test and benchmark input) -------------------- */
typedef struct sicp_synth_config {
    int32_t cols, rows;
    float fx, fy, cx, cy;
    uint64_t seed;
    float depth_units_per_m; /* 1000: 1 mm per unit */
    float dropout;           /* fraction of pixels set to 0 */
    int32_t noise;           /* 0 = noiseless depth */
} sicp_synth_config;
typedef struct sicp_synth sicp_synth;
int sicp_synth_create(const sicp_synth_config* cfg, sicp_synth** out);
int sicp_synth_destroy(sicp_synth* s);
/* Camera->world pose of frame i (row-major 4x4 double). */
int sicp_synth_pose(const sicp_synth* s, int frame, double pose[16]);
/* Renders frame i into host buffers (depth rows*cols, rgb rows*cols*3, labels). */
int sicp_synth_render(const sicp_synth* s, int frame, uint16_t* depth, uint8_t* rgb,
                      uint8_t* labels);

#ifdef __cplusplus
}
#endif
#endif /* SICP_H_ */
