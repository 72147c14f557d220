// system.hip — the per-frame driver: the C++ counterpart of the reference's
// Frame / Tracker / Mapping / System callers of the hot path
// (Core/Frame.cc:33-98, Tracking/Tracking.cc:65-119,157-173,246-280,369-517,
// Mapping/Mapping.cc:76-97,142-161, Core/System.cc:131-172).
//
// A frame is one ordered sequence of launches on one stream with no host
// synchronisation: the poses, the Gauss-Newton state, the visible count and
// the tracked / lost decision all stay in device memory, and the fusion /
// raycast kernels read the decision from there. The host only swaps its
// frame pointers, unconditionally; when a frame is lost, a device-side copy
// puts the last model maps and pose into the frame that becomes "last", so
// the next frame tracks against the same model as in the reference, where
// the swap is skipped.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "se3_math.hpp"
#include "sicp_internal.hpp"

namespace sicp {

struct FrameBuf {
    int cols[3], rows[3];
    DevImage<float> range;  // raw depth in metres, level 0 (Frame::range)
    DevImage<float> depth[3];
    DevImage<float4> vmap[3], nmap[3];
    DevImage<uint8_t> image[3];
    DevImage<int16_t> dx[3], dy[3];
    DevImage<int16_t> key[3];   // photometric candidate keys (with the photometric term)
    FrameMeta* meta = nullptr;  // device

    int create(int c, int r, bool rgb) {
        int rc;
        if ((rc = range.create(c, r))) return rc;
        for (int i = 0; i < 3; ++i) {
            cols[i] = c / (1 << i);
            rows[i] = r / (1 << i);
            if ((rc = depth[i].create(cols[i], rows[i])) || (rc = vmap[i].create(cols[i], rows[i])) ||
                (rc = nmap[i].create(cols[i], rows[i])) || (rc = image[i].create(cols[i], rows[i])) ||
                (rc = dx[i].create(cols[i], rows[i])) || (rc = dy[i].create(cols[i], rows[i])))
                return rc;
            if (rgb && (rc = key[i].create(cols[i], rows[i]))) return rc;
        }
        hipError_t e = hipMalloc((void**)&meta, sizeof(FrameMeta));
        if (e != hipSuccess) return hip_fail(e, "hipMalloc(FrameMeta)");
        FrameMeta h{};
        set_identity(h.pose);
        pose_to_float(h.pose, &h.pose_f);
        e = hipMemcpy(meta, &h, sizeof(h), hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_fail(e, "hipMemcpy(FrameMeta)");
        return SICP_OK;
    }
    void release() {
        range.release();
        for (int i = 0; i < 3; ++i) {
            depth[i].release();
            vmap[i].release();
            nmap[i].release();
            image[i].release();
            dx[i].release();
            dy[i].release();
            key[i].release();
        }
        if (meta) (void)hipFree(meta);
        meta = nullptr;
    }
};

}  // namespace sicp

struct sicp_system {
    sicp_system_config cfg;
    sicp_map* map = nullptr;
    sicp::FrameBuf fa, fb;
    sicp::FrameBuf* next = &fa;
    sicp::FrameBuf* last = &fb;
    sicp::TrackState* st = nullptr;    // [0] the frame's final state, [1..2] iterations
    double* partials = nullptr;        // two sets of ICP partial rows, alternating
    void* rgb_mem = nullptr;           // photometric scratch (records, counts, rows)
    sicp::RgbScratch rgb{};
    sicp::DevImage<float> zmin, zmax;
    sicp_intrinsics K[3];
    int frame_index = 0;
    // Stage timing: when enabled, every stage of every frame is bracketed by
    // a pair of events recorded on the frame's stream.
    bool timing = false;
    std::vector<hipEvent_t> pool;
    size_t pool_used = 0;
    std::vector<std::pair<int, size_t>> marks;  // (stage, index of the start event)
    // Graph form: staged inputs owned by the system and, per frame parity
    // (which buffer is "next"), one instantiated graph of the whole frame.
    sicp::DevImage<uint16_t> in_depth;
    sicp::DevImage<uint8_t> in_rgb, in_labels;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    int graph_nd = -1;  // num_detection the graphs were captured with
    hipStream_t capture_stream = nullptr;
    // Launch shape of the raycast (set or autotuned); graphs recorded under
    // an older tuning epoch are recorded again.
    sicp::RayTuning tuning;
    uint64_t tuning_epoch = 0;
    uint64_t graph_epoch = ~0ull;
    std::vector<double> tune_sum;  // per autotuner candidate: summed median ms over calls
    int tune_calls = 0;
    // Batched frames: the system's own stages run on its side stream, forked
    // from and joined back to the batch's stream with these events.
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
};

namespace sicp {

enum Stage { kStagePyramid = 0, kStageTrack, kStageAlloc, kStageVisible, kStageFuse,
             kStagePrepass, kStageRaycast, kStageFuseCount, kStageFrame, kStageIcpIter,
             kStageTrackFinish, kStageCount };

// Returns the index of the start event, or SIZE_MAX when timing is off.
static size_t stage_begin(sicp_system* s, hipStream_t st) {
    if (!s->timing) return (size_t)-1;
    while (s->pool.size() < s->pool_used + 2) {
        hipEvent_t e;
        // Timing-only events: no system-scope fence (cache writeback and
        // invalidate) at each record; read_timing synchronises the stream.
        if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess)
            return (size_t)-1;
        s->pool.push_back(e);
    }
    size_t i = s->pool_used;
    s->pool_used += 2;
    (void)hipEventRecord(s->pool[i], st);
    return i;
}

static void stage_end(sicp_system* s, int stage, size_t i, hipStream_t st) {
    if (i == (size_t)-1) return;
    (void)hipEventRecord(s->pool[i + 1], st);
    s->marks.push_back({stage, i});
}

// Frame::FillImages (Core/Frame.cc:68-92).
// Frame::Create / FillImages of the new frame (Core/Frame.cc:33-92) and
// ResizeImages of the last frame's model maps (:94-98).
static int frame_pyramid(sicp_system* s, FrameBuf& f, FrameBuf& last,
                         const sicp_image& depth_u16, const sicp_image& rgb, hipStream_t st) {
    FramePyramid P;
    ModelPyramid mp;
    for (int i = 0; i < 3; ++i) {
        const sicp_intrinsics& k = s->K[i];
        P.depth[i] = f.depth[i].v();
        P.vmap[i] = f.vmap[i].v();
        P.nmap[i] = f.nmap[i].v();
        P.image[i] = f.image[i].v();
        P.dx[i] = f.dx[i].v();
        P.dy[i] = f.dy[i].v();
        P.key[i] = f.key[i].ptr ? f.key[i].v() : View<int16_t>{};
        P.invfx[i] = (float)(1.0 / (double)k.fx);
        P.invfy[i] = (float)(1.0 / (double)k.fy);
        P.cx[i] = k.cx;
        P.cy[i] = k.cy;
        mp.v[i] = last.vmap[i].v();
        mp.n[i] = last.nmap[i].v();
    }
    // the frame's fusion count and render start from these
    const FrameReset fr{s->zmin.v(), s->zmax.v(), &s->map->ctr->render_tiles,
                        s->map->dev.rq_count, &s->map->ctr->updated_voxels};
    return launch_frame_pyramid(view<uint16_t>(depth_u16), view<uint8_t>(rgb), f.range.v(), P,
                                (float)(1.0 / (double)s->cfg.depth_scale), s->cfg.depth_cutoff,
                                mp, fr, st);
}

static void system_free(sicp_system* s) {
    if (!s) return;
    for (hipEvent_t e : s->pool) (void)hipEventDestroy(e);
    for (hipGraphExec_t& g : s->graph)
        if (g) (void)hipGraphExecDestroy(g);
    if (s->capture_stream) (void)hipStreamDestroy(s->capture_stream);
    if (s->side) (void)hipStreamDestroy(s->side);
    if (s->ev_fork) (void)hipEventDestroy(s->ev_fork);
    if (s->ev_join) (void)hipEventDestroy(s->ev_join);
    s->in_depth.release();
    s->in_rgb.release();
    s->in_labels.release();
    s->fa.release();
    s->fb.release();
    s->zmin.release();
    s->zmax.release();
    if (s->st) (void)hipFree(s->st);
    if (s->partials) (void)hipFree(s->partials);
    if (s->rgb_mem) (void)hipFree(s->rgb_mem);
    map_free(s->map);
    delete s;
}

}  // namespace sicp

using namespace sicp;

namespace sicp {
// Staged inputs: the images of a round (three per system, at most eight
// systems) copied by one kernel. Image i owns workgroups wg0[i] ..
// wg0[i + 1] - 1, each copying kStageUnits units (16, 4 or 1 bytes, the
// widest that the rows' addresses, pitches and length allow) per thread.
constexpr int kStageMax = 3 * kMaxBatch, kStageUnits = 4, kStageThreads = 256;
struct StageCopy {
    const unsigned char* src;
    unsigned char* dst;
    size_t spitch, dpitch;
    uint32_t row_units, units, shift, wg0;
};
struct StageArgs {
    StageCopy c[kStageMax];
    int n;
};
__global__ __launch_bounds__(kStageThreads) void stage_copy_kernel(StageArgs a) {
    int i = 0;
    while (i + 1 < a.n && blockIdx.x >= a.c[i + 1].wg0) ++i;
    const StageCopy& c = a.c[i];
    const uint32_t u0 = (blockIdx.x - c.wg0) * (kStageThreads * kStageUnits) + threadIdx.x;
#pragma unroll
    for (int j = 0; j < kStageUnits; ++j) {
        const uint32_t u = u0 + j * kStageThreads;
        if (u >= c.units) break;
        const uint32_t row = u / c.row_units, col = u - row * c.row_units;
        const unsigned char* s = c.src + row * c.spitch + ((size_t)col << c.shift);
        unsigned char* d = c.dst + row * c.dpitch + ((size_t)col << c.shift);
        if (c.shift == 4)
            *(uint4*)d = *(const uint4*)s;
        else if (c.shift == 2)
            *(uint32_t*)d = *(const uint32_t*)s;
        else
            *d = *s;
    }
}

// Adds the copy of `src` into `dst` (same shape, `elem` bytes per element).
static void stage_add(StageArgs& a, uint32_t& wgs, const sicp_image& src, const sicp_image& dst,
                      size_t elem) {
    const size_t row_bytes = (size_t)dst.cols * elem;
    const uintptr_t any = (uintptr_t)src.ptr | (uintptr_t)dst.ptr | src.pitch | dst.pitch | row_bytes;
    const uint32_t shift = (any & 15u) == 0 ? 4u : ((any & 3u) == 0 ? 2u : 0u);
    StageCopy& c = a.c[a.n++];
    c.src = (const unsigned char*)src.ptr;
    c.dst = (unsigned char*)dst.ptr;
    c.spitch = src.pitch;
    c.dpitch = dst.pitch;
    c.row_units = (uint32_t)(row_bytes >> shift);
    c.units = c.row_units * (uint32_t)dst.rows;
    c.shift = shift;
    c.wg0 = wgs;
    wgs += div_up(c.units, (uint32_t)(kStageThreads * kStageUnits));
}

// The system's staged inputs, allocated on first use.
static int staged_buffers(sicp_system* s) {
    if (s->in_depth.ptr) return SICP_OK;
    int rc;
    const int c = s->cfg.cols, r = s->cfg.rows;
    if ((rc = s->in_depth.create(c, r)) || (rc = s->in_rgb.create(c, r, 3)) ||
        (rc = s->in_labels.create(c, r)))
        return rc;
    return SICP_OK;
}

// Checks one system's frame and adds its copies.
static int stage_frame(sicp_system* s, const sicp_image* depth, const sicp_image* rgb,
                       const sicp_image* labels, StageArgs& a, uint32_t& wgs) {
    const int c = s->cfg.cols, r = s->cfg.rows;
    if (!depth || !rgb || !image_ok(*depth, 2, c, r) || !image_ok(*rgb, 3, c, r))
        return invalid("stage: depth / rgb shape mismatch");
    if (labels && labels->ptr && !image_ok(*labels, 1, c, r))
        return invalid("stage: label image shape mismatch");
    const int rc = staged_buffers(s);
    if (rc) return rc;
    stage_add(a, wgs, *depth, s->in_depth.img(), 2);
    stage_add(a, wgs, *rgb, s->in_rgb.img(), 3);
    if (labels && labels->ptr) stage_add(a, wgs, *labels, s->in_labels.img(), 1);
    return SICP_OK;
}

static int stage_launch(const StageArgs& a, uint32_t wgs, hipStream_t st) {
    if (wgs == 0) return SICP_OK;
    stage_copy_kernel<<<wgs, kStageThreads, 0, st>>>(a);
    SICP_LAUNCH_CHECK("stage_copy_kernel");
    return SICP_OK;
}
}  // namespace sicp

extern "C" {

void sicp_system_config_default(sicp_system_config* cfg, int cols, int rows, float fx, float fy,
                                float cx, float cy, float voxel_size, float truncation,
                                int32_t num_buckets, int32_t num_blocks) {
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->cols = cols;
    cfg->rows = rows;
    cfg->K = sicp_intrinsics{fx, fy, cx, cy};
    cfg->depth_scale = 1000.f;  // Frame::mDepthScale (Core/Frame.cc:20)
    cfg->depth_cutoff = 3.0f;   // Frame::mDepthCutoff (Core/Frame.cc:19)
    cfg->icp_iters[0] = 10;     // ITERATIONS_SE3 (Tracking.h:110)
    cfg->icp_iters[1] = 5;
    cfg->icp_iters[2] = 3;
    cfg->icp_only = 1;             // the survey's tracking decision: ICP only
    cfg->rgb_weight = 1e-4f;       // Tracking.cc:392
    cfg->icp_spread = 1;
    cfg->icp_min_count[0] = 2000;  // MIN_ICP_COUNT (Tracking.h:112)
    cfg->icp_min_count[1] = 1000;
    cfg->icp_min_count[2] = 100;
    cfg->icp_thresh = 1e-4f;    // THRESH_ICP_SE3 (Tracking.h:117)
    sicp_map_config_default(&cfg->map, voxel_size, truncation, num_buckets, num_blocks);
    cfg->map.max_cols = cols;
    cfg->map.max_rows = rows;
}

int sicp_system_create(const sicp_system_config* cfg, sicp_system** out) {
    if (!cfg || !out) return invalid("system_create: null argument");
    if (cfg->cols < 8 || cfg->rows < 8) return invalid("system_create: image too small");
    if ((size_t)cfg->cols * cfg->rows > (size_t)cfg->map.max_cols * cfg->map.max_rows)
        return invalid("system_create: map max_cols x max_rows smaller than the image");
    if (cfg->icp_only != 0 && cfg->icp_only != 1)
        return invalid("system_create: icp_only must be 0 or 1");
    if (!cfg->icp_only && !std::isfinite(cfg->rgb_weight))
        return invalid("system_create: rgb_weight must be finite");
    if (cfg->icp_sums_fp32 != 0 && cfg->icp_sums_fp32 != 1)
        return invalid("system_create: icp_sums_fp32 must be 0 or 1");
    if (cfg->icp_spread != 0 && cfg->icp_spread != 1)
        return invalid("system_create: icp_spread must be 0 or 1");
    const bool rgb = !cfg->icp_only;
    sicp_system* s = new sicp_system();
    s->cfg = *cfg;
    int rc = map_alloc(cfg->map, &s->map);
    if (rc) {
        delete s;
        return rc;
    }
    if ((rc = s->fa.create(cfg->cols, cfg->rows, rgb)) ||
        (rc = s->fb.create(cfg->cols, cfg->rows, rgb)) ||
        (rc = s->zmin.create((cfg->cols + 7) / 8, (cfg->rows + 7) / 8)) ||
        (rc = s->zmax.create((cfg->cols + 7) / 8, (cfg->rows + 7) / 8))) {
        system_free(s);
        return rc;
    }
    hipError_t e = hipMalloc((void**)&s->st, 3 * sizeof(TrackState));
    if (e == hipSuccess) e = hipMemset(s->st, 0, 3 * sizeof(TrackState));
    if (e == hipSuccess)
        e = hipMalloc((void**)&s->partials, 2 * sizeof(double) * kIcpMaxBlocks * kIcpSums);
    if (e == hipSuccess && rgb) {
        // records (int2 per level-0 pixel), counts, rows, totals: one allocation
        const size_t n = (size_t)cfg->cols * cfg->rows;
        const size_t rec = n * sizeof(int2), ip = 2 * kIcpMaxBlocks * sizeof(uint32_t);
        const size_t pr = (size_t)kIcpSums * kIcpMaxBlocks * sizeof(double);
        e = hipMalloc(&s->rgb_mem, rec + ip + pr + 16);
        if (e == hipSuccess) {
            char* b = (char*)s->rgb_mem;  // rows first: 8-byte aligned
            s->rgb.partials = (double*)b;
            s->rgb.rec = (int2*)(b + pr);
            s->rgb.ipart = (uint32_t*)(b + pr + rec);
            s->rgb.totals = (int32_t*)(b + pr + rec + ip);
        }
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        rc = hip_fail(e, "system_create");
        system_free(s);
        return rc;
    }
    for (int i = 0; i < 3; ++i) {
        // Frame::SetK (Core/Frame.cc:580-588): K / 2^i in float.
        float d = (float)(1 << i);
        s->K[i] = sicp_intrinsics{cfg->K.fx / d, cfg->K.fy / d, cfg->K.cx / d, cfg->K.cy / d};
    }
    *out = s;
    return SICP_OK;
}

int sicp_system_destroy(sicp_system* sys) {
    system_free(sys);
    return SICP_OK;
}

sicp_map* sicp_system_map(sicp_system* sys) { return sys ? sys->map : nullptr; }

// System::GrabImage (Core/System.cc:131-172) with Tracker::GrabFrame /
// Track / TrackFrame / ComputeSE3 (Tracking.cc:65-119,157-173,246-280,369-517).
static int grab_frame(sicp_system* s, sicp_image depth_u16, sicp_image rgb_u8x3,
                      sicp_image labels_u8, int num_detection, void* stream);

int sicp_system_grab_image(sicp_system* s, sicp_image depth_u16, sicp_image rgb_u8x3,
                           sicp_image labels_u8, int num_detection, void* stream) {
    if (!s) return invalid("system_grab_image: null system");
    const size_t ev = stage_begin(s, (hipStream_t)stream);
    const int rc = grab_frame(s, depth_u16, rgb_u8x3, labels_u8, num_detection, stream);
    if (rc == SICP_OK) stage_end(s, kStageFrame, ev, (hipStream_t)stream);
    return rc;
}

// Input checks and Tracker::GrabFrame's images: the last frame's model
// pyramid and the new frame's pyramid.
static int grab_begin(sicp_system* s, const sicp_image& depth_u16, const sicp_image& rgb_u8x3,
                      const sicp_image& labels_u8, int num_detection, hipStream_t st) {
    const int c = s->cfg.cols, r = s->cfg.rows;
    if (!image_ok(depth_u16, 2, c, r) || !image_ok(rgb_u8x3, 3, c, r))
        return invalid("system_grab_image: depth / rgb shape mismatch");
    const bool labels = num_detection > 0 && labels_u8.ptr;
    if (labels && !image_ok(labels_u8, 1, c, r))
        return invalid("system_grab_image: label image shape mismatch");
    const size_t ev = stage_begin(s, st);
    const int rc = frame_pyramid(s, *s->next, *s->last, depth_u16, rgb_u8x3, st);
    if (rc) return rc;
    stage_end(s, kStagePyramid, ev, st);
    return SICP_OK;
}

static TrackParams track_params(const sicp_system* s, int launch = 0) {
    return TrackParams{s->cfg.icp_thresh, s->cfg.rgb_weight, s->cfg.icp_only ? 0 : 1,
                       s->cfg.icp_sums_fp32, s->cfg.icp_spread, launch};
}

// The level-l maps of the tracking pair (next frame against the last).
static IcpLevel icp_level(const sicp_system* s, int l) {
    const FrameBuf& nf = *s->next;
    const FrameBuf& lf = *s->last;
    IcpLevel cur{nf.vmap[l].v(), lf.vmap[l].v(), nf.nmap[l].v(), lf.nmap[l].v(),
                 s->K[l], s->cfg.icp_min_count[l], RgbLevel{}};
    if (!s->cfg.icp_only)
        cur.rgb = RgbLevel{nf.key[l].v(), nf.dx[l].v(), nf.dy[l].v(), lf.image[l].v()};
    return cur;
}

// Tracker::TrackFrame / ComputeSE3 of one frame (Tracking.cc:157-173,
// 369-517), or InitTracking on the first frame.
static int track_single(sicp_system* s, hipStream_t st) {
    int rc;
    FrameBuf& nf = *s->next;
    FrameBuf& lf = *s->last;
    const size_t ev = stage_begin(s, st);
    if (s->frame_index == 0) {
        // InitTracking (Tracking.cc:233-240): both poses identity.
        if ((rc = launch_track_init(nf.meta, lf.meta, s->st, st))) return rc;
    } else {
        // Every launch runs the previous iteration's update and its own pixel
        // pass; states and partial rows alternate between two buffers.
        // With the photometric term, RGBStep's row pass follows each pixel
        // pass (its sigma needs the whole correspondence pass).
        TrackState* iter = s->st + 1;
        const size_t psz = (size_t)kIcpMaxBlocks * kIcpSums;
        const bool rgb = !s->cfg.icp_only;
        TrackParams tp = track_params(s);
        IcpLevel prev{};
        int k = 0;
        for (int l = 2; l >= 0; --l) {
            const IcpLevel cur = icp_level(s, l);
            for (int j = 0; j < s->cfg.icp_iters[l]; ++j, ++k) {
                const size_t ek = stage_begin(s, st);
                tp.launch = k;
                rc = launch_icp_iteration(cur, k ? &prev : nullptr, lf.meta, tp,
                                          iter + ((k + 1) & 1), iter + (k & 1),
                                          s->partials + psz * ((k + 1) & 1),
                                          s->partials + psz * (k & 1), rgb ? &s->rgb : nullptr,
                                          st);
                if (rc) return rc;
                if (rgb && (rc = launch_rgb_rows(cur, iter + (k & 1), s->rgb,
                                                 tp.sums_k32 != 0, tp.spread != 0, st)))
                    return rc;
                stage_end(s, kStageIcpIter, ek, st);
                prev = cur;
            }
        }
        FrameImages im{};
        if (rgb)
            for (int l = 0; l < 3; ++l) {
                im.last[l] = lf.image[l].v();
                im.next[l] = nf.image[l].v();
            }
        const size_t ef = stage_begin(s, st);
        if ((rc = launch_track_finish(k ? &prev : nullptr, lf.meta, tp, iter + ((k + 1) & 1),
                                      s->partials + psz * ((k + 1) & 1), rgb ? &s->rgb : nullptr,
                                      s->st, nf.meta, lf.vmap[0].v(), lf.nmap[0].v(),
                                      nf.vmap[0].v(), nf.nmap[0].v(), rgb ? &im : nullptr, st)))
            return rc;
        stage_end(s, kStageTrackFinish, ef, st);
    }
    stage_end(s, kStageTrack, ev, st);
    return SICP_OK;
}

// SwapFrame, then Mapping::FuseColor and Mapping::RayTrace of the tracked
// frame (Mapping.cc:76-97,142-161), each gated on the device-side decision.
static int grab_end(sicp_system* s, const sicp_image& rgb_u8x3, const sicp_image& labels_u8,
                    int num_detection, hipStream_t st) {
    int rc;
    size_t ev;
    const int c = s->cfg.cols, r = s->cfg.rows;
    const bool labels = num_detection > 0 && labels_u8.ptr;
    std::swap(s->next, s->last);  // SwapFrame (Tracking.cc:318-320)
    s->frame_index++;
    FrameBuf& f = *s->last;
    const int* gate = &s->st->tracked;
    PoseArg pose = make_pose_arg(nullptr, &f.meta->pose_f);
    sicp_map* m = s->map;
    // Mapping::FuseColor (Mapping.cc:76-97) -> FuseMapColor (FuseMap.cu:1039-1162)
    ev = stage_begin(s, st);
    // the allocation's last launch also flags the visible blocks
    if ((rc = launch_allocate(m->dev, m->ctr, m->alloc, f.range.v(), pose, s->K[0],
                              VisFrustum{c, r, m->cfg.depth_max, m->cfg.depth_min}, gate, st)))
        return rc;
    stage_end(s, kStageAlloc, ev, st);
    ev = stage_begin(s, st);
    if ((rc = launch_visible_compact(m->dev, m->ctr, m->alloc, m->visible, gate, st)))
        return rc;
    stage_end(s, kStageVisible, ev, st);
    View<uint8_t> mask = labels ? view<uint8_t>(labels_u8) : View<uint8_t>{};
    ev = stage_begin(s, st);
    if ((rc = launch_fuse(m->dev, m->ctr, m->visible, f.range.v(), view<uint8_t>(rgb_u8x3),
                          f.nmap[0].v(), mask, labels ? num_detection : 0, pose, s->K[0],
                          m->alloc.fuse_counts, gate, st)))
        return rc;
    stage_end(s, kStageFuse, ev, st);
    // The fusion's counts are summed by the render prepass; the zRange images,
    // tile and queue counts were reset by the frame's first launch.
    ev = stage_begin(s, st);
    if (m->cfg.unify_labels && labels && (rc = launch_unify(m, gate, st))) return rc;
    stage_end(s, kStageFuseCount, ev, st);
    // Mapping::RayTrace (Mapping.cc:142-161): prepass, then the raycast into
    // the frame's level-0 model maps when any block produced a tile.
    ev = stage_begin(s, st);
    if ((rc = launch_render_prepass(m->dev, m->ctr, m->visible, s->zmin.v(), s->zmax.v(), pose,
                                    s->K[0], m->cfg.depth_max, m->cfg.depth_min, gate, st,
                                    true, m->alloc.fuse_counts)))
        return rc;
    stage_end(s, kStagePrepass, ev, st);
    ev = stage_begin(s, st);
    rc = launch_raycast_gated(m->dev, f.vmap[0].v(), f.nmap[0].v(), s->zmin.v(), s->zmax.v(),
                              pose, (float)(1.0 / (double)s->K[0].fx),
                              (float)(1.0 / (double)s->K[0].fy), s->K[0].cx, s->K[0].cy, gate,
                              &m->ctr->render_tiles, st, true, s->tuning);
    stage_end(s, kStageRaycast, ev, st);
    return rc;
}

static int grab_frame(sicp_system* s, sicp_image depth_u16, sicp_image rgb_u8x3,
                      sicp_image labels_u8, int num_detection, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    int rc;
    if ((rc = grab_begin(s, depth_u16, rgb_u8x3, labels_u8, num_detection, st))) return rc;
    if ((rc = track_single(s, st))) return rc;
    return grab_end(s, rgb_u8x3, labels_u8, num_detection, st);
}

// The tracking of several systems' frames, each past its first frame, in
// shared launches: one per Gauss-Newton iteration and one finishing launch.
static int track_batch(sicp_system* const* sys, int n, hipStream_t st) {
    const size_t psz = (size_t)kIcpMaxBlocks * kIcpSums;
    IcpLevel prev[kMaxBatch], cur[kMaxBatch];
    IcpBatchEntry e[kMaxBatch];
    int k = 0, rc;
    for (int l = 2; l >= 0; --l) {
        for (int q = 0; q < n; ++q) cur[q] = icp_level(sys[q], l);
        for (int j = 0; j < sys[0]->cfg.icp_iters[l]; ++j, ++k) {
            for (int q = 0; q < n; ++q) {
                sicp_system* s = sys[q];
                TrackState* iter = s->st + 1;
                e[q] = IcpBatchEntry{cur[q],
                                     k ? &prev[q] : nullptr,
                                     s->last->meta,
                                     track_params(s, k),
                                     iter + ((k + 1) & 1),
                                     iter + (k & 1),
                                     s->partials + psz * ((k + 1) & 1),
                                     s->partials + psz * (k & 1)};
            }
            if ((rc = launch_icp_iteration_batch(e, n, st))) return rc;
            for (int q = 0; q < n; ++q) prev[q] = cur[q];
        }
    }
    FinishBatchEntry f[kMaxBatch];
    for (int q = 0; q < n; ++q) {
        sicp_system* s = sys[q];
        f[q] = FinishBatchEntry{&prev[q],
                                s->last->meta,
                                track_params(s, k),
                                s->st + 1 + ((k + 1) & 1),
                                s->partials + psz * ((k + 1) & 1),
                                s->st,
                                s->next->meta,
                                s->last->vmap[0].v(),
                                s->last->nmap[0].v(),
                                s->next->vmap[0].v(),
                                s->next->nmap[0].v()};
    }
    return launch_track_finish_batch(f, n, st);
}

// Checks that the systems can share tracking launches: distinct, at most
// kMaxBatch, point-to-plane tracking with one form of the sums and the same
// iteration counts.
static int batch_ok(sicp_system* const* sys, int n) {
    if (!sys || n < 1 || n > kMaxBatch) return invalid("system batch: 1 to 8 systems");
    for (int q = 0; q < n; ++q) {
        if (!sys[q]) return invalid("system batch: null system");
        for (int p = 0; p < q; ++p)
            if (sys[p] == sys[q]) return invalid("system batch: a system appears twice");
        if (!sys[q]->cfg.icp_only)
            return invalid("system batch: the photometric term is not batched (icp_only = 1)");
        if (sys[q]->cfg.icp_sums_fp32 != sys[0]->cfg.icp_sums_fp32)
            return invalid("system batch: all systems need the same form of the sums");
        for (int l = 0; l < 3; ++l)
            if (sys[q]->cfg.icp_iters[l] != sys[0]->cfg.icp_iters[l])
                return invalid("system batch: all systems need the same iteration counts");
    }
    return SICP_OK;
}

// Side stream of a batched system: its work after fork_to runs after what
// was queued on `st` before, and `st`'s work after join_from runs after it.
static int side_init(sicp_system* s) {
    if (!s->side) SICP_HIP(hipStreamCreateWithFlags(&s->side, hipStreamNonBlocking));
    if (!s->ev_fork) SICP_HIP(hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming));
    if (!s->ev_join) SICP_HIP(hipEventCreateWithFlags(&s->ev_join, hipEventDisableTiming));
    return SICP_OK;
}

static int fork_to(sicp_system* s, hipStream_t st) {
    SICP_HIP(hipEventRecord(s->ev_fork, st));
    SICP_HIP(hipStreamWaitEvent(s->side, s->ev_fork, 0));
    return SICP_OK;
}

static int join_from(sicp_system* s, hipStream_t st) {
    SICP_HIP(hipEventRecord(s->ev_join, s->side));
    SICP_HIP(hipStreamWaitEvent(st, s->ev_join, 0));
    return SICP_OK;
}

// One frame of every system of a batch: the pyramids, the tracking of the
// systems past their first frame in shared launches on `st` (the others
// initialise), then each system's swap, fusion and raycast. A system's own
// stages run on its side stream, so the systems' pyramids overlap, and so do
// their fusions and raycasts; in a captured round these are the graph's
// parallel branches.
// With `shared` false, every system's whole frame (its own tracking
// launches included) runs on its side stream instead: the round is n
// independent chains, joined at its end.
static int grab_batch(sicp_system* const* sys, int n, const sicp_image* depth,
                      const sicp_image* rgb, const sicp_image* labels, const int32_t* nd,
                      hipStream_t st, bool shared = true) {
    int rc;
    sicp_system* tracked[kMaxBatch];
    int nt = 0;
    for (int q = 0; q < n; ++q)
        if ((rc = side_init(sys[q])) || (rc = fork_to(sys[q], st))) return rc;
    if (!shared) {
        for (int q = 0; q < n; ++q) {
            sicp_system* s = sys[q];
            const sicp_image lab = labels ? labels[q] : sicp_image{nullptr, 0, 0, 0};
            if ((rc = grab_frame(s, depth[q], rgb[q], lab, nd ? nd[q] : 0, s->side)) ||
                (rc = join_from(s, st)))
                return rc;
        }
        return SICP_OK;
    }
    for (int q = 0; q < n; ++q) {
        sicp_system* s = sys[q];
        const sicp_image lab = labels ? labels[q] : sicp_image{nullptr, 0, 0, 0};
        if ((rc = grab_begin(s, depth[q], rgb[q], lab, nd ? nd[q] : 0, s->side))) return rc;
        if (s->frame_index == 0) {
            if ((rc = track_single(s, s->side))) return rc;
        } else {
            tracked[nt++] = s;
        }
        if ((rc = join_from(s, st))) return rc;
    }
    if (nt && (rc = track_batch(tracked, nt, st))) return rc;
    for (int q = 0; q < n; ++q) {
        sicp_system* s = sys[q];
        const sicp_image lab = labels ? labels[q] : sicp_image{nullptr, 0, 0, 0};
        if ((rc = fork_to(s, st)) || (rc = grab_end(s, rgb[q], lab, nd ? nd[q] : 0, s->side)) ||
            (rc = join_from(s, st)))
            return rc;
    }
    return SICP_OK;
}

// Stage timing of the frames grabbed since it was enabled. Stages: 0 pyramid,
// 1 track, 2 allocate, 3 visibility, 4 fuse kernel, 5 render prepass,
// 6 raycast (both passes), 7 fusion count, 8 whole frame, 9 each ICP
// iteration launch, 10 tracking finish.
int sicp_system_set_timing(sicp_system* s, int enable) {
    if (!s) return invalid("system_set_timing: null system");
    s->timing = enable != 0;
    s->marks.clear();
    s->pool_used = 0;
    return SICP_OK;
}

int sicp_system_read_timing(sicp_system* s, double* total_ms, int* counts, int n_stages,
                            void* stream) {
    if (!s || !total_ms || !counts) return invalid("system_read_timing: null argument");
    SICP_HIP(hipStreamSynchronize((hipStream_t)stream));
    for (int i = 0; i < n_stages; ++i) {
        total_ms[i] = 0.0;
        counts[i] = 0;
    }
    for (auto& mk : s->marks) {
        float ms = 0.f;
        SICP_HIP(hipEventElapsedTime(&ms, s->pool[mk.second], s->pool[mk.second + 1]));
        if (mk.first < n_stages) {
            total_ms[mk.first] += ms;
            counts[mk.first] += 1;
        }
    }
    s->marks.clear();
    s->pool_used = 0;
    return SICP_OK;
}

int sicp_system_read_timing_detail(sicp_system* s, int stage, double* ms_out, int cap,
                                   int* count, void* stream) {
    if (!s || !count || (cap > 0 && !ms_out)) return invalid("system_read_timing_detail: null argument");
    SICP_HIP(hipStreamSynchronize((hipStream_t)stream));
    int n = 0;
    for (auto& mk : s->marks) {
        if (mk.first != stage) continue;
        if (n < cap) {
            float ms = 0.f;
            SICP_HIP(hipEventElapsedTime(&ms, s->pool[mk.second], s->pool[mk.second + 1]));
            ms_out[n] = ms;
        }
        ++n;
    }
    *count = n;
    return SICP_OK;
}

int sicp_system_last_result(sicp_system* s, sicp_frame_result* out, void* stream) {
    if (!s || !out) return invalid("system_last_result: null argument");
    hipStream_t st = (hipStream_t)stream;
    TrackState t;
    MapCounters c;
    FrameMeta fm;
    SICP_HIP(hipMemcpyAsync(&t, s->st, sizeof(t), hipMemcpyDeviceToHost, st));
    SICP_HIP(hipMemcpyAsync(&c, s->map->ctr, sizeof(c), hipMemcpyDeviceToHost, st));
    SICP_HIP(hipMemcpyAsync(&fm, s->last->meta, sizeof(fm), hipMemcpyDeviceToHost, st));
    SICP_HIP(hipStreamSynchronize(st));
    std::memset(out, 0, sizeof(*out));
    out->frame_index = s->frame_index - 1;
    out->tracked = t.tracked;
    out->icp_count = t.icp_count;
    out->iterations = t.applied;
    out->icp_error = t.icp_error;
    out->rgb_error = t.rgb_error;
    out->rgb_count = t.rgb_count;
    out->visible_blocks = c.visible;
    out->allocated_blocks = c.heap_used;
    std::memcpy(out->pose, fm.pose, sizeof(out->pose));
    return SICP_OK;
}

int sicp_system_staging(sicp_system* s, sicp_image* depth_u16, sicp_image* rgb_u8x3,
                        sicp_image* labels_u8) {
    if (!s) return invalid("system_staging: null system");
    const int rc = staged_buffers(s);
    if (rc) return rc;
    if (depth_u16) *depth_u16 = s->in_depth.img();
    if (rgb_u8x3) *rgb_u8x3 = s->in_rgb.img();
    if (labels_u8) *labels_u8 = s->in_labels.img();
    return SICP_OK;
}

// Records the frame sequence of both frame parities into graphs: the host
// state (frame pointers, frame index) is set for each parity, the launches
// are captured on a private stream, and the host state is restored.
static int capture_graphs(sicp_system* s, int nd) {
    if (!s->capture_stream)
        SICP_HIP(hipStreamCreateWithFlags(&s->capture_stream, hipStreamNonBlocking));
    for (hipGraphExec_t& g : s->graph)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    const bool timing = s->timing;
    FrameBuf* const next0 = s->next;
    FrameBuf* const last0 = s->last;
    const int index0 = s->frame_index;
    s->timing = false;
    const sicp_image lab = nd > 0 ? s->in_labels.img() : sicp_image{nullptr, 0, 0, 0};
    int rc = SICP_OK;
    for (int p = 0; p < 2 && rc == SICP_OK; ++p) {
        s->next = p == 0 ? &s->fa : &s->fb;
        s->last = p == 0 ? &s->fb : &s->fa;
        s->frame_index = index0 > 0 ? index0 : 1;
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamBeginCapture(s->capture_stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) {
            rc = hip_fail(e, "hipStreamBeginCapture");
            break;
        }
        rc = grab_frame(s, s->in_depth.img(), s->in_rgb.img(), lab, nd, s->capture_stream);
        e = hipStreamEndCapture(s->capture_stream, &g);
        if (rc == SICP_OK && e != hipSuccess) rc = hip_fail(e, "hipStreamEndCapture");
        if (rc == SICP_OK) {
            e = hipGraphInstantiate(&s->graph[p], g, nullptr, nullptr, 0);
            if (e != hipSuccess) rc = hip_fail(e, "hipGraphInstantiate");
        }
        if (g) (void)hipGraphDestroy(g);
    }
    s->next = next0;
    s->last = last0;
    s->frame_index = index0;
    s->timing = timing;
    s->graph_nd = rc == SICP_OK ? nd : -1;
    s->graph_epoch = rc == SICP_OK ? s->tuning_epoch : ~0ull;
    return rc;
}

int sicp_system_stage(sicp_system* s, const sicp_image* depth_u16, const sicp_image* rgb_u8x3,
                      const sicp_image* labels_u8, void* stream) {
    if (!s) return invalid("system_stage: null system");
    StageArgs a{};
    uint32_t wgs = 0;
    const int rc = stage_frame(s, depth_u16, rgb_u8x3, labels_u8, a, wgs);
    if (rc) return rc;
    return stage_launch(a, wgs, (hipStream_t)stream);
}

int sicp_system_grab_staged(sicp_system* s, int num_detection, int use_graph, void* stream) {
    if (!s) return invalid("system_grab_staged: null system");
    if (!s->in_depth.ptr) return invalid("system_grab_staged: call sicp_system_staging first");
    const int nd = num_detection > 0 ? num_detection : 0;
    const sicp_image lab = nd > 0 ? s->in_labels.img() : sicp_image{nullptr, 0, 0, 0};
    // The first frame (tracking initialisation) and timed frames run eagerly.
    if (!use_graph || s->frame_index == 0 || s->timing)
        return sicp_system_grab_image(s, s->in_depth.img(), s->in_rgb.img(), lab, nd, stream);
    if (!s->graph[0] || s->graph_nd != nd || s->graph_epoch != s->tuning_epoch) {
        const int rc = capture_graphs(s, nd);
        if (rc) return rc;
    }
    const int p = (s->next == &s->fa) ? 0 : 1;
    SICP_HIP(hipGraphLaunch(s->graph[p], (hipStream_t)stream));
    std::swap(s->next, s->last);  // what the recorded frame's host code does
    s->frame_index++;
    return SICP_OK;
}

int sicp_system_grab_batch(sicp_system* const* systems, int32_t n, const sicp_image* depth_u16,
                           const sicp_image* rgb_u8x3, const sicp_image* labels_u8,
                           const int32_t* num_detection, void* stream) {
    int rc = batch_ok(systems, n);
    if (rc) return rc;
    if (!depth_u16 || !rgb_u8x3) return invalid("system_grab_batch: null image array");
    return grab_batch(systems, n, depth_u16, rgb_u8x3, labels_u8, num_detection,
                      (hipStream_t)stream);
}

}  // extern "C"

// A batch of systems advanced together, with the graph form of a round: the
// frame-buffer parities of the systems flip together, so a round has two
// parity patterns and one instantiated graph each.
struct sicp_batch {
    std::vector<sicp_system*> sys;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    uint64_t pattern[2] = {0, 0};
    std::vector<int32_t> nd;  // num_detection the graphs were captured with
    std::vector<uint64_t> epochs;  // the systems' tuning epochs at capture
    hipStream_t capture_stream = nullptr;
    bool shared = true;  // tracking iterations in shared launches (else per-system chains)
};

namespace sicp {
static std::vector<uint64_t> batch_epochs(const sicp_batch* b) {
    std::vector<uint64_t> e;
    for (const sicp_system* s : b->sys) e.push_back(s->tuning_epoch);
    return e;
}
}  // namespace sicp

namespace sicp {

// Bit q: system q's `next` frame buffer is its first one.
static uint64_t batch_pattern(const sicp_batch* b) {
    uint64_t p = 0;
    for (size_t q = 0; q < b->sys.size(); ++q)
        if (b->sys[q]->next == &b->sys[q]->fa) p |= 1ull << q;
    return p;
}

static void batch_drop_graphs(sicp_batch* b) {
    for (hipGraphExec_t& g : b->graph)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
}

// Records a round of the batch from the systems' staged inputs for the
// current parity pattern and its complement, as capture_graphs does for one
// system; the host state of every system is restored afterwards.
static int batch_capture(sicp_batch* b, const int32_t* nd) {
    if (!b->capture_stream)
        SICP_HIP(hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking));
    batch_drop_graphs(b);
    const int n = (int)b->sys.size();
    std::vector<FrameBuf*> next0(n), last0(n);
    std::vector<int> index0(n);
    std::vector<bool> timing0(n);
    std::vector<sicp_image> depth(n), rgb(n), lab(n);
    for (int q = 0; q < n; ++q) {
        sicp_system* s = b->sys[q];
        next0[q] = s->next;
        last0[q] = s->last;
        index0[q] = s->frame_index;
        timing0[q] = s->timing;
        s->timing = false;
        depth[q] = s->in_depth.img();
        rgb[q] = s->in_rgb.img();
        lab[q] = nd[q] > 0 ? s->in_labels.img() : sicp_image{nullptr, 0, 0, 0};
    }
    const uint64_t cur = batch_pattern(b);
    int rc = SICP_OK;
    for (int p = 0; p < 2 && rc == SICP_OK; ++p) {
        for (int q = 0; q < n; ++q) {
            sicp_system* s = b->sys[q];
            const bool first_is_next = ((cur >> q) & 1) != (uint64_t)p;
            s->next = first_is_next ? &s->fa : &s->fb;
            s->last = first_is_next ? &s->fb : &s->fa;
            s->frame_index = index0[q] > 0 ? index0[q] : 1;
        }
        b->pattern[p] = batch_pattern(b);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamBeginCapture(b->capture_stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) {
            rc = hip_fail(e, "hipStreamBeginCapture");
            break;
        }
        rc = grab_batch(b->sys.data(), n, depth.data(), rgb.data(), lab.data(), nd,
                        b->capture_stream, b->shared);
        e = hipStreamEndCapture(b->capture_stream, &g);
        if (rc == SICP_OK && e != hipSuccess) rc = hip_fail(e, "hipStreamEndCapture");
        if (rc == SICP_OK) {
            e = hipGraphInstantiate(&b->graph[p], g, nullptr, nullptr, 0);
            if (e != hipSuccess) rc = hip_fail(e, "hipGraphInstantiate");
        }
        if (g) (void)hipGraphDestroy(g);
    }
    for (int q = 0; q < n; ++q) {
        sicp_system* s = b->sys[q];
        s->next = next0[q];
        s->last = last0[q];
        s->frame_index = index0[q];
        s->timing = timing0[q];
    }
    if (rc == SICP_OK) {
        b->nd.assign(nd, nd + n);
        b->epochs = batch_epochs(b);
    } else {
        batch_drop_graphs(b);
    }
    return rc;
}

}  // namespace sicp

extern "C" {

int sicp_batch_create(sicp_system* const* systems, int32_t n, sicp_batch** out) {
    if (!out) return invalid("batch_create: null argument");
    int rc = batch_ok(systems, n);
    if (rc) return rc;
    for (int q = 0; q < n; ++q)
        if ((rc = sicp_system_staging(systems[q], nullptr, nullptr, nullptr))) return rc;
    sicp_batch* b = new sicp_batch();
    b->sys.assign(systems, systems + n);
    *out = b;
    return SICP_OK;
}

int sicp_batch_set_shared_tracking(sicp_batch* b, int shared) {
    if (!b) return invalid("batch_set_shared_tracking: null batch");
    if ((shared != 0) != b->shared) {
        b->shared = shared != 0;
        batch_drop_graphs(b);
    }
    return SICP_OK;
}

int sicp_batch_destroy(sicp_batch* b) {
    if (!b) return SICP_OK;
    batch_drop_graphs(b);
    if (b->capture_stream) (void)hipStreamDestroy(b->capture_stream);
    delete b;
    return SICP_OK;
}

int sicp_batch_stage(sicp_batch* b, const sicp_image* depth_u16, const sicp_image* rgb_u8x3,
                     const sicp_image* labels_u8, void* stream) {
    if (!b) return invalid("batch_stage: null batch");
    if (!depth_u16 || !rgb_u8x3) return invalid("batch_stage: null image array");
    StageArgs a{};
    uint32_t wgs = 0;
    for (size_t q = 0; q < b->sys.size(); ++q) {
        const int rc = stage_frame(b->sys[q], depth_u16 + q, rgb_u8x3 + q,
                                   labels_u8 ? labels_u8 + q : nullptr, a, wgs);
        if (rc) return rc;
    }
    return stage_launch(a, wgs, (hipStream_t)stream);
}

int sicp_batch_grab_staged(sicp_batch* b, const int32_t* num_detection, int use_graph,
                           void* stream) {
    if (!b) return invalid("batch_grab_staged: null batch");
    const int n = (int)b->sys.size();
    std::vector<int32_t> nd(n, 0);
    std::vector<sicp_image> depth(n), rgb(n), lab(n);
    bool eager = !use_graph;
    for (int q = 0; q < n; ++q) {
        sicp_system* s = b->sys[q];
        nd[q] = num_detection && num_detection[q] > 0 ? num_detection[q] : 0;
        depth[q] = s->in_depth.img();
        rgb[q] = s->in_rgb.img();
        lab[q] = nd[q] > 0 ? s->in_labels.img() : sicp_image{nullptr, 0, 0, 0};
        // first frames (tracking initialisation) and timed frames run eagerly
        if (s->frame_index == 0 || s->timing) eager = true;
    }
    if (eager)
        return grab_batch(b->sys.data(), n, depth.data(), rgb.data(), lab.data(), nd.data(),
                          (hipStream_t)stream, b->shared);
    const uint64_t cur = batch_pattern(b);
    int p = (b->graph[0] && b->pattern[0] == cur) ? 0 : ((b->graph[1] && b->pattern[1] == cur) ? 1 : -1);
    if (p < 0 || b->nd != nd || b->epochs != batch_epochs(b)) {
        const int rc = batch_capture(b, nd.data());
        if (rc) return rc;
        p = b->pattern[0] == cur ? 0 : 1;
    }
    SICP_HIP(hipGraphLaunch(b->graph[p], (hipStream_t)stream));
    for (sicp_system* s : b->sys) {  // what the recorded host code does
        std::swap(s->next, s->last);
        s->frame_index++;
    }
    return SICP_OK;
}

int sicp_system_frame_image(sicp_system* s, int which, int level, sicp_image* out) {
    if (!s || !out || level < 0 || level > 2) return invalid("system_frame_image: bad argument");
    FrameBuf& f = (which & 16) ? *s->next : *s->last;
    switch (which & 15) {
        case 0: *out = f.vmap[level].img(); break;
        case 1: *out = f.nmap[level].img(); break;
        case 2: *out = f.range.img(); break;
        case 3: *out = f.depth[level].img(); break;
        case 4: *out = f.image[level].img(); break;
        case 5: *out = f.dx[level].img(); break;
        case 6: *out = f.dy[level].img(); break;
        default: return invalid("system_frame_image: unknown image");
    }
    return SICP_OK;
}

}  // extern "C"

// ---- raycast launch tuning -------------------------------------------------
// The raycast's two launch-shape parameters leave its output unchanged, so
// they are chosen by measurement. Each sicp_system_autotune call re-runs the
// last frame's raycast (same map, zRange images and pose, into scratch maps)
// under every candidate, interleaved over `repeats` rounds, and adds each
// candidate's median time to the system's running sums; the candidate with
// the lowest sum so far becomes the tuning. Calls on several frames of a
// sequence therefore choose for the sequence, not for one view.

namespace sicp {

// Candidate 0 is the default shape.
static const RayTuning kTuneCandidates[] = {
    {48, 1024},  {16, 256},  {16, 512},  {16, 1024}, {16, 2048},  {24, 256},  {24, 512},
    {24, 1024},  {24, 2048}, {32, 256},  {32, 512},  {32, 1024},  {32, 2048}, {48, 256},
    {48, 512},   {48, 2048}, {64, 256},  {64, 512},  {64, 1024},  {64, 2048}, {96, 256},
    {96, 512},   {96, 1024}, {96, 2048}, {128, 256}, {128, 512},  {128, 1024}, {128, 2048}};
constexpr int kNumTuneCandidates = (int)(sizeof(kTuneCandidates) / sizeof(kTuneCandidates[0]));

struct TuneScratch {
    DevImage<float4> vmap, nmap;
    std::vector<hipEvent_t> events;
    ~TuneScratch() {
        vmap.release();
        nmap.release();
        for (hipEvent_t e : events)
            if (e) (void)hipEventDestroy(e);
    }
};

static bool tuning_ok(const sicp_tuning& t) {
    return t.ray_first_pass_samples >= 1 && t.ray_first_pass_samples <= kRayTuneMaxSamples &&
           t.ray_second_pass_workgroups >= 1 &&
           t.ray_second_pass_workgroups <= kRayTuneMaxWorkgroups &&
           (t.ray_lds_block_table == 0 || t.ray_lds_block_table == 1);
}

// Candidate c with the system's own block-table setting.
static RayTuning tune_candidate(const sicp_system* s, int c) {
    RayTuning t = kTuneCandidates[c];
    t.lds_block_table = s->tuning.lds_block_table;
    return t;
}

static void set_tuning(sicp_system* s, const RayTuning& t) {
    if (t.first_pass_samples == s->tuning.first_pass_samples &&
        t.second_pass_workgroups == s->tuning.second_pass_workgroups &&
        t.lds_block_table == s->tuning.lds_block_table)
        return;
    s->tuning = t;
    s->tuning_epoch++;
}

}  // namespace sicp

extern "C" {

int sicp_system_get_tuning(const sicp_system* s, sicp_tuning* out) {
    if (!s || !out) return invalid("system_get_tuning: null argument");
    out->ray_first_pass_samples = s->tuning.first_pass_samples;
    out->ray_second_pass_workgroups = s->tuning.second_pass_workgroups;
    out->ray_lds_block_table = s->tuning.lds_block_table;
    return SICP_OK;
}

int sicp_system_set_tuning(sicp_system* s, const sicp_tuning* t) {
    if (!s || !t) return invalid("system_set_tuning: null argument");
    if (!tuning_ok(*t))
        return invalid("system_set_tuning: samples outside 1..4096, workgroups outside "
                       "1..65536 or block table not 0/1");
    // The autotuner's sums were measured with the block-table setting in
    // effect; a new setting starts them again.
    if (t->ray_lds_block_table != s->tuning.lds_block_table) (void)sicp_system_autotune_reset(s);
    set_tuning(s, RayTuning{t->ray_first_pass_samples, t->ray_second_pass_workgroups,
                            t->ray_lds_block_table});
    return SICP_OK;
}

int sicp_system_autotune_reset(sicp_system* s) {
    if (!s) return invalid("system_autotune_reset: null system");
    s->tune_sum.clear();
    s->tune_calls = 0;
    return SICP_OK;
}

int sicp_system_autotune_results(const sicp_system* s, sicp_tuning* shapes, float* mean_us,
                                 int32_t cap, int32_t* count) {
    if (!s || !count) return invalid("system_autotune_results: null argument");
    *count = s->tune_calls ? kNumTuneCandidates : 0;
    for (int c = 0; c < *count && c < cap; ++c) {
        if (shapes) {
            const RayTuning t = tune_candidate(s, c);
            shapes[c] = sicp_tuning{t.first_pass_samples, t.second_pass_workgroups,
                                    t.lds_block_table};
        }
        if (mean_us) mean_us[c] = (float)(1000.0 * s->tune_sum[c] / s->tune_calls);
    }
    return SICP_OK;
}

int sicp_system_autotune(sicp_system* s, int32_t repeats, sicp_tuning* chosen, float* chosen_us,
                         float* default_us, void* stream) {
    if (!s) return invalid("system_autotune: null system");
    if (repeats < 1 || repeats > 64) return invalid("system_autotune: repeats outside 1..64");
    if (s->frame_index == 0) return invalid("system_autotune: no frame processed yet");
    hipStream_t st = (hipStream_t)stream;
    sicp_map* m = s->map;
    TrackState t;
    uint32_t tiles = 0;
    SICP_HIP(hipMemcpyAsync(&t, s->st, sizeof(t), hipMemcpyDeviceToHost, st));
    SICP_HIP(hipMemcpyAsync(&tiles, &m->ctr->render_tiles, sizeof(tiles), hipMemcpyDeviceToHost, st));
    SICP_HIP(hipStreamSynchronize(st));
    if (!t.tracked || tiles == 0)
        return invalid("system_autotune: the last frame raycast nothing (lost, or no visible block)");
    const int nc = kNumTuneCandidates;
    TuneScratch ts;
    int rc;
    if ((rc = ts.vmap.create(s->cfg.cols, s->cfg.rows)) ||
        (rc = ts.nmap.create(s->cfg.cols, s->cfg.rows)))
        return rc;
    ts.events.assign((size_t)2 * nc * repeats, nullptr);
    for (hipEvent_t& e : ts.events) SICP_HIP(hipEventCreate(&e));
    const FrameBuf& f = *s->last;  // the frame whose raycast is re-run
    const PoseArg pose = make_pose_arg(nullptr, &f.meta->pose_f);
    const float ifx = (float)(1.0 / (double)s->K[0].fx), ify = (float)(1.0 / (double)s->K[0].fy);
    auto run = [&](const RayTuning& c) {
        return launch_raycast_gated(m->dev, ts.vmap.v(), ts.nmap.v(), s->zmin.v(), s->zmax.v(),
                                    pose, ifx, ify, s->K[0].cx, s->K[0].cy, &s->st->tracked,
                                    &m->ctr->render_tiles, st, false, c);
    };
    if ((rc = run(s->tuning))) return rc;  // warm the caches
    for (int r = 0; r < repeats; ++r)
        for (int c = 0; c < nc; ++c) {
            hipEvent_t* e = &ts.events[(size_t)2 * (r * nc + c)];
            SICP_HIP(hipEventRecord(e[0], st));
            if ((rc = run(tune_candidate(s, c)))) return rc;
            SICP_HIP(hipEventRecord(e[1], st));
        }
    SICP_HIP(hipStreamSynchronize(st));
    std::vector<float> x(repeats);
    if (s->tune_sum.size() != (size_t)nc) s->tune_sum.assign(nc, 0.0);
    for (int c = 0; c < nc; ++c) {
        for (int r = 0; r < repeats; ++r) {
            const hipEvent_t* e = &ts.events[(size_t)2 * (r * nc + c)];
            SICP_HIP(hipEventElapsedTime(&x[r], e[0], e[1]));
        }
        std::sort(x.begin(), x.end());
        s->tune_sum[c] += x[repeats / 2];
    }
    s->tune_calls++;
    int best = 0;
    for (int c = 1; c < nc; ++c)
        if (s->tune_sum[c] < s->tune_sum[best]) best = c;
    // Leave the hand-over queue count as the chosen launch shape leaves it.
    if ((rc = run(tune_candidate(s, best)))) return rc;
    SICP_HIP(hipStreamSynchronize(st));
    set_tuning(s, tune_candidate(s, best));
    if (chosen) (void)sicp_system_get_tuning(s, chosen);
    if (chosen_us) *chosen_us = (float)(1000.0 * s->tune_sum[best] / s->tune_calls);
    if (default_us) *default_us = (float)(1000.0 * s->tune_sum[0] / s->tune_calls);
    return SICP_OK;
}

}  // extern "C"
