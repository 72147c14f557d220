// sicp_kernels.hpp — internal launch API shared by the HIP translation
// units, the C ABI (abi.cpp) and the per-frame driver (system.cpp).
#pragma once

#include "sicp_common.hpp"

namespace sicp {

// ------------------------------------------------------------- pyramid ----
int launch_filter_depth(View<uint16_t> depth, View<float> raw, View<float> filtered,
                        float scale_inv, float cutoff, hipStream_t s);
int launch_vmap(View<float> depth, View<float4> vmap, float invfx, float invfy, float cx,
                float cy, float cutoff, hipStream_t s);
int launch_nmap(View<float4> vmap, View<float4> nmap, hipStream_t s);
int launch_pyrdown_f32(View<float> src, View<float> dst, hipStream_t s);
int launch_pyrdown_u8(View<uint8_t> src, View<uint8_t> dst, hipStream_t s);
int launch_intensity(View<uint8_t> rgb, View<uint8_t> gray, hipStream_t s);
int launch_derivative(View<uint8_t> img, View<int16_t> dx, View<int16_t> dy, hipStream_t s);
int launch_resize_map(View<float4> vsrc, View<float4> nsrc, View<float4> vdst, View<float4> ndst,
                      hipStream_t s);
// The photometric candidate key of every pixel (RgbLevel::key).
int launch_rgb_keys(View<uint8_t> img, View<int16_t> dx, View<int16_t> dy, View<float4> vmap,
                    View<int16_t> key, hipStream_t s);

// The images of one frame at three pyramid levels (Frame, Core/Frame.h).
struct FramePyramid {
    View<float> depth[3];
    View<float4> vmap[3], nmap[3];
    View<uint8_t> image[3];
    View<int16_t> dx[3], dy[3];
    View<int16_t> key[3];  // photometric candidate keys (RgbLevel::key); ptr null: not made
    float invfx[3], invfy[3], cx[3], cy[3];
};
// A frame's model maps (raycast at level 0, 2x2 means below).
struct ModelPyramid {
    View<float4> v[3], n[3];
};
// What the frame's first launch also resets for the frame's fusion count and
// render: the zRange images (zmin 100, zmax 0), the render tile count, the
// raycast queue count and the count of voxels the fusion updated.
struct FrameReset {
    View<float> zmin, zmax;
    uint32_t* render_tiles;
    uint32_t* rq_count;
    uint32_t* updated_voxels;
};
// FillImages of the new frame and ResizeImages of the last one, in three
// launches; pixel for pixel the same as the single-image launches above.
int launch_frame_pyramid(View<uint16_t> depth_u16, View<uint8_t> rgb, View<float> range,
                         const FramePyramid& P, float scale_inv, float cutoff,
                         const ModelPyramid& last_model, const FrameReset& reset, hipStream_t s);

// ---------------------------------------------------------------- ICP -----
constexpr int kIcpThreads = 256;
constexpr int kIcpMaxBlocks = 512;
constexpr int kIcpSums = 29;

// Pose and Gauss-Newton state of one frame's tracking, resident on the device
// so every iteration runs without a host round trip.
struct FrameMeta {
    double pose[16];   // camera -> world, row-major (Frame::pose, Core/Frame.h)
    sicp_pose_f pose_f;
};

struct TrackState {
    double last_pose[16];
    double next_pose[16];
    double pose0[16];
    double delta[16];
    double candidate[16];
    sicp_pose_f curr_f, last_f;
    float icp_error, last_error, candidate_icp, thresh;
    int icp_count, break_count, applied, done;
    int failed, tracked, frame_index, pad;
    float rgb_error, rgb_weight;  // photometric error of the last iteration; its weight w
    int rgb_count, rgb_on;        // its correspondences; 1 when the term is used
};

// Arguments of ComputeSE3 (Tracking.cc:369) that are not images.
struct TrackParams {
    float thresh;      // thresh_icp
    float rgb_weight;  // w (Tracking.cc:392)
    int rgb;           // !icpOnly: RGBStep runs before ICPStep in every iteration
    int sums_k32;      // the precision form: FP32 compensated sums (else FP64)
    int spread;        // one pixel per thread on small levels (the ICP partition)
    int launch;        // the launch's index in the frame's tracking loop (phase stamps)
};

struct IcpScratch {
    double* partials;  // kIcpSums * kIcpMaxBlocks, column-major
};

// ICPStep: the 29 sums of one level into `sums_out` (device); `k32`: the
// precision form's FP32 compensated sums instead of FP64 accumulation.
int launch_icp_sums(View<float4> vcurr, View<float4> vlast, View<float4> ncurr,
                    View<float4> nlast, PoseArg curr, PoseArg last, sicp_intrinsics K,
                    IcpScratch scratch, float* sums_out, bool k32, hipStream_t s);

// The photometric inputs of one level (RGBStep, Reduction.h:81-90): the
// current frame's candidate key (its intensity where the pose-independent
// tests of FindCorresp pass, else -1) and derivatives, the last frame's
// intensity.
struct RgbLevel {
    View<int16_t> key, dx, dy;
    View<uint8_t> ilast;
};
// Device memory of the photometric pass.
struct RgbScratch {
    int2* rec;         // per pixel: last-frame pixel index (-1: none), intensity difference
    uint32_t* ipart;   // per workgroup: count, then sum of squared differences (2 * kIcpMaxBlocks)
    double* partials;  // kIcpSums * kIcpMaxBlocks, column-major
    int32_t* totals;   // count, sum of squared differences, sigma (float bits)
};

// The maps of one pyramid level and its correspondence minimum.
struct IcpLevel {
    View<float4> vcurr, vlast, ncurr, nlast;
    sicp_intrinsics K;
    int min_count;
    RgbLevel rgb;  // used when the photometric term is on
};
// One Gauss-Newton iteration of ComputeSE3: the update of the previous
// iteration (`prev` = its level; null for the first iteration of a frame,
// whose state is built from `last`'s pose), then this iteration's pixel pass
// into `partials`. The state moves from `st_in` to `st_out` (two buffers,
// alternating).
// With `rgb` set (the photometric term on), the pixel pass also runs the
// correspondence pass of RGBStep into `rgb->rec` / `rgb->ipart`, and the
// update folds the previous iteration's photometric rows from
// `rgb->partials`.
int launch_icp_iteration(const IcpLevel& cur, const IcpLevel* prev, const FrameMeta* last,
                         TrackParams tp, const TrackState* st_in, TrackState* st_out,
                         const double* partials_prev, double* partials, const RgbScratch* rgb,
                         hipStream_t s);
// RGBStep's second pass for the iteration whose state is `st`: sigma from the
// correspondence counts, then the 29 weighted photometric sums into
// `rgb.partials`.
int launch_rgb_rows(const IcpLevel& cur, const TrackState* st, RgbScratch rgb, bool k32,
                    bool spread, hipStream_t s);
// The last iteration's update, the final gate, the new frame's pose and,
// when the frame is not tracked, the copy of the last frame's level-0 model
// maps (and, with the photometric term, its intensity pyramid) into the new
// frame. The frame's final state goes to `st_final`.
struct FrameImages {
    View<uint8_t> last[3], next[3];
};
int launch_track_finish(const IcpLevel* prev, const FrameMeta* last, TrackParams tp,
                        const TrackState* st_in, const double* partials_prev,
                        const RgbScratch* rgb, TrackState* st_final, FrameMeta* next,
                        View<float4> vlast, View<float4> nlast, View<float4> vnext,
                        View<float4> nnext, const FrameImages* images, hipStream_t s);

// Several sequences' tracking in the same launches (point-to-plane term):
// at Gauss-Newton iteration k every sequence of a batch is at the same level
// and iteration, so one launch serves them all, each on its own range of
// workgroups with exactly its own launch's arithmetic.
constexpr int kMaxBatch = 8;
struct IcpBatchEntry {
    IcpLevel cur;
    const IcpLevel* prev;  // the previous iteration's level; null for the first
    const FrameMeta* last;
    TrackParams tp;
    const TrackState* st_in;
    TrackState* st_out;
    const double* partials_prev;
    double* partials;
};
int launch_icp_iteration_batch(const IcpBatchEntry* e, int n, hipStream_t s);
struct FinishBatchEntry {
    const IcpLevel* prev;  // the last iteration's level
    const FrameMeta* last;
    TrackParams tp;
    const TrackState* st_in;
    const double* partials_prev;
    TrackState* st_final;
    FrameMeta* next;
    View<float4> vlast, nlast, vnext, nnext;
};
int launch_track_finish_batch(const FinishBatchEntry* e, int n, hipStream_t s);

// RGBStep on its own (the C ABI): key pass, correspondence pass, row pass,
// then the 29 sums into `sums_out` (device) and the correspondence count,
// sum of squared differences and sigma into `rgb.totals`.
int launch_rgb_sums(View<uint8_t> inext, View<uint8_t> ilast, View<float4> vnext,
                    View<float4> vlast, View<int16_t> dx, View<int16_t> dy, PoseArg curr,
                    PoseArg last, sicp_intrinsics K, int16_t* key, RgbScratch rgb,
                    float* sums_out, hipStream_t s);
int launch_track_init(FrameMeta* next, FrameMeta* last, TrackState* st, hipStream_t s);

// SO3Step: the 11 sums of the rotation-only photometric alignment (the 6
// upper entries of A and b interleaved row by row, sum r^2, count) into
// `sums_out`.
int launch_so3_sums(View<uint8_t> inext, View<uint8_t> ilast, View<int16_t> dx,
                    View<int16_t> dy, PoseArg curr, PoseArg last, sicp_intrinsics K,
                    IcpScratch scratch, float* sums_out, hipStream_t s);

// Host FP64 solve + SE(3) update (same code path as the device).
void se3_update_host(const double* A, const double* b, double* delta, const double* last_pose,
                     double* next_pose, double* x_out);

// ---------------------------------------------------------------- map -----
// A ray handed from the first raycast pass to the second: its pixel and the
// march state at the hand-over (the rest is recomputed from the pixel).
// A ray handed from the raycast's first pass to its second: the pixel, the
// march's position, distance and last sdf, and the ray's direction and end.
struct RayState {
    int px;
    float rx, ry, rz, dist_s, sdf;
    float dx, dy, dz, dist_e;
    int pad[2];
};

struct MapDev {
    sicp_hash_entry* table;   // 2 * nb
    sicp_hash_entry* excess;  // num_blocks, indexed by heap slot
    sicp_block* blocks;       // num_blocks
    sicp_voxel_geom* geom;    // num_blocks * 512 (FP32 form)
    uint32_t* geom_h;         // num_blocks * 512 (half form): sdf half | weight << 16
    int half;                 // 1: the half form (geom_h; colour in sem.pad)
    sicp_voxel_sem* sem;      // num_blocks * 512
    uint32_t nb, nb_mask;     // nb_mask = nb-1 when nb is a power of two, else 0
    int num_blocks;
    float voxel, voxel_inv, trunc, step_scale, block_width, dmin, dmax;
    int max_weight, max_count;
    RayState* rq;        // raycast hand-over queue, rq_cap entries
    uint32_t* rq_count;  // entries queued by the first pass
    uint32_t rq_cap;
    // A device word holding 1, read in place of an absent gate or flag, so
    // that kernels load their flags without testing the pointers first.
    const int* one;
};

struct MapCounters {
    uint32_t heap_used;
    uint32_t visible;
    uint32_t new_total;      // distinct new blocks requested this frame
    uint32_t new_blocks;     // of which created
    uint32_t dropped_heap;
    uint32_t dropped_requests;
    uint32_t dropped_range;
    uint32_t updated_voxels;
    uint32_t render_tiles;
    uint32_t alloc_base;     // heap_used when this frame's allocation began (its first slot)
    unsigned long long updated_total;  // voxels written, summed since reset
    unsigned long long visible_total;  // visible blocks, summed since reset
    unsigned long long fusions;        // fusion passes since reset
    unsigned long long heap_total;     // allocated blocks at each fusion, summed since reset
};

struct AllocScratch {
    uint64_t* req_keys;        // request set keys (EMPTY = ~0)
    unsigned long long* req_first; // first requester id per request-set slot
    uint32_t req_mask;         // capacity - 1 (power of two)
    unsigned long long* pix_mask;  // per pixel: steps that missed the table
    uint32_t* pix_count;       // per pixel: new blocks first seen there
    uint32_t* scan_partials;   // scan workspace
    uint64_t* new_keys;        // rank -> key
    uint32_t* new_req_slot;    // rank -> request-set slot
    int32_t* bucket_leader;    // per bucket, smallest rank this frame (INT_MAX idle)
    int32_t* bucket_head;      // per bucket, pending list head (-1 idle)
    int32_t* pending_next;     // rank -> next pending rank
    uint8_t* vis_flags;        // per heap slot
    uint32_t* fuse_counts;     // per fusion wavefront: voxels written (kFuseBlocks * 4)
    uint32_t cap_pixels;
};

constexpr int kFuseBlocks = 2048;  // fusion workgroups (4 wavefronts each), grid-stride

int launch_map_reset(MapDev m, MapCounters* ctr, AllocScratch a, hipStream_t s);
// The camera frustum of a visibility pass: image size and depth range.
struct VisFrustum {
    int cols, rows;
    float dmax, dmin;
};
// The allocation (Fusion::CreateBlocks) and, in its last launch, the
// visibility flags of the heap it leaves (CheckBlockVisibility) for the
// camera `K` and `vf`; launch_visible_compact then completes the
// visibility pass (the visible list and count).
int launch_allocate(MapDev m, MapCounters* ctr, AllocScratch a, View<float> depth, PoseArg pose,
                    sicp_intrinsics K, VisFrustum vf, const int* gate, hipStream_t s);
int launch_visible_compact(MapDev m, MapCounters* ctr, AllocScratch a, sicp_block* visible,
                           const int* gate, hipStream_t s);
// A whole visibility pass (flags, then the visible list).
int launch_visibility(MapDev m, MapCounters* ctr, AllocScratch a, sicp_block* visible,
                      PoseArg pose, sicp_intrinsics K, int cols, int rows, float dmax,
                      float dmin, const int* gate, hipStream_t s);
int launch_fuse(MapDev m, MapCounters* ctr, const sicp_block* visible, View<float> depth,
                View<uint8_t> rgb, View<float4> nmap, View<uint8_t> mask, int num_detection,
                PoseArg pose, sicp_intrinsics K, uint32_t* wave_counts, const int* gate,
                hipStream_t s);
// The per-wavefront counts of launch_fuse into the map's counters (always run
// after it). With `zmin` set, the same launch also clears the zRange images,
// the rendering-tile count and the raycast queue count (`rq_count`), so the
// prepass and the raycast that follow can skip their own clearing launches.
int launch_fuse_count(MapCounters* ctr, const uint32_t* wave_counts, const int* gate,
                      hipStream_t s, View<float> zmin = View<float>{},
                      View<float> zmax = View<float>{}, uint32_t* rq_count = nullptr);
// `zrange_cleared`: the zRange images and the tile count were cleared earlier
// on the same stream (FrameReset). `fuse_counts` (kFuseBlocks * 4 per-wavefront
// counts of the fusion before it, with updated_voxels reset by FrameReset): the
// prepass also does launch_fuse_count's sums.
int launch_render_prepass(MapDev m, MapCounters* ctr, const sicp_block* visible, View<float> zmin,
                          View<float> zmax, PoseArg pose, sicp_intrinsics K, float dmax,
                          float dmin, const int* gate, hipStream_t s,
                          bool zrange_cleared = false, const uint32_t* fuse_counts = nullptr);
// Launch shape of the two-pass raycast. Any valid choice gives the same
// vertex and normal maps bit for bit: the first pass hands a ray over with
// its complete loop state, and the second pass continues that loop exactly.
struct RayTuning {
    int first_pass_samples = 48;     // samples a first-pass lane marches before the hand-over
    int second_pass_workgroups = 1024;  // 256-thread workgroups draining the hand-over queue
    int lds_block_table = 1;  // first pass: block lookups through the wavefront's LDS table
};
constexpr int kRayTuneMaxSamples = 4096;
constexpr int kRayTuneMaxWorkgroups = 65536;
int launch_raycast(MapDev m, View<float4> vmap, View<float4> nmap, View<float> zmin,
                   View<float> zmax, PoseArg pose, float invfx, float invfy, float cx, float cy,
                   const int* gate, hipStream_t s);

// The raycast's adjoint for `n` sampled pixels (pixel index y * cols + x):
// per-pixel pose gradients (n x 6, may be null), their sum into
// `grad_pose` (6 doubles, device), TSDF gradients accumulated into
// `grad_sdf` (num_blocks x 512, may be null). `partials`: scratch of
// raycast_adjoint_scratch_bytes().
int launch_raycast_adjoint(MapDev m, View<float> zmin, View<float> zmax, PoseArg pose,
                           float invfx, float invfy, float cx, float cy, int cols, int rows,
                           const int32_t* pixels, int n, const float* grad_v,
                           const float* grad_n, float* grad_sdf, float* grad_px,
                           double* partials, double* grad_pose, hipStream_t s);
size_t raycast_adjoint_scratch_bytes();

// Device exclusive scan of `n_dev` (device count, at most `cap`) u32 values.
int launch_scan_u32(const uint32_t* in, uint32_t* out, const uint32_t* n_dev, uint32_t cap,
                    uint32_t* partials, uint32_t* total_out, hipStream_t s);

}  // namespace sicp
