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

static int grab_frame(sicp_system* s, sicp_image depth_u16, sicp_image rgb_u8x3,
                      sicp_image labels_u8, int num_detection, void* stream) {
    const int c = s->cfg.cols, r = s->cfg.rows;
    if (!image_ok(depth_u16, 2, c, r) || !image_ok(rgb_u8x3, 3, c, r))
        return invalid("system_grab_image: depth / rgb shape mismatch");
    const bool labels = num_detection > 0 && labels_u8.ptr;
    if (labels && !image_ok(labels_u8, 1, c, r))
        return invalid("system_grab_image: label image shape mismatch");
    hipStream_t st = (hipStream_t)stream;
    int rc;
    size_t ev;
    // Tracker::GrabFrame: model pyramid of the last frame, then the new frame.
    ev = stage_begin(s, st);
    if ((rc = frame_pyramid(s, *s->next, *s->last, depth_u16, rgb_u8x3, st))) return rc;
    stage_end(s, kStagePyramid, ev, st);
    FrameBuf& nf = *s->next;
    FrameBuf& lf = *s->last;
    ev = stage_begin(s, st);
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
        const TrackParams tp{s->cfg.icp_thresh, s->cfg.rgb_weight, rgb ? 1 : 0,
                             s->cfg.icp_sums_fp32};
        IcpLevel prev{};
        int k = 0;
        for (int l = 2; l >= 0; --l) {
            IcpLevel cur{nf.vmap[l].v(), lf.vmap[l].v(), nf.nmap[l].v(), lf.nmap[l].v(),
                         s->K[l], s->cfg.icp_min_count[l], RgbLevel{}};
            if (rgb) cur.rgb = RgbLevel{nf.key[l].v(), nf.dx[l].v(), nf.dy[l].v(), lf.image[l].v()};
            for (int j = 0; j < s->cfg.icp_iters[l]; ++j, ++k) {
                const size_t ek = stage_begin(s, st);
                rc = launch_icp_iteration(cur, k ? &prev : nullptr, lf.meta, tp,
                                          iter + ((k + 1) & 1), iter + (k & 1),
                                          s->partials + psz * ((k + 1) & 1),
                                          s->partials + psz * (k & 1), rgb ? &s->rgb : nullptr,
                                          st);
                if (rc) return rc;
                if (rgb && (rc = launch_rgb_rows(cur, iter + (k & 1), s->rgb,
                                                 tp.sums_k32 != 0, st)))
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
    std::swap(s->next, s->last);  // SwapFrame (Tracking.cc:318-320)
    s->frame_index++;
    FrameBuf& f = *s->last;
    const int* gate = &s->st->tracked;
    PoseArg pose = make_pose_arg(nullptr, &f.meta->pose_f);
    sicp_map* m = s->map;
    // Mapping::FuseColor (Mapping.cc:76-97) -> FuseMapColor (FuseMap.cu:1039-1162)
    ev = stage_begin(s, st);
    if ((rc = launch_allocate(m->dev, m->ctr, m->alloc, f.range.v(), pose, s->K[0], gate, st)))
        return rc;
    stage_end(s, kStageAlloc, ev, st);
    ev = stage_begin(s, st);
    if ((rc = launch_visibility(m->dev, m->ctr, m->alloc, m->visible, pose, s->K[0], c, r,
                               m->cfg.depth_max, m->cfg.depth_min, gate, st)))
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
                              &m->ctr->render_tiles, st, true);
    stage_end(s, kStageRaycast, ev, st);
    return rc;
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
    if (!s->in_depth.ptr) {
        int rc;
        const int c = s->cfg.cols, r = s->cfg.rows;
        if ((rc = s->in_depth.create(c, r)) || (rc = s->in_rgb.create(c, r, 3)) ||
            (rc = s->in_labels.create(c, r)))
            return rc;
    }
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
    return rc;
}

int sicp_system_grab_staged(sicp_system* s, int num_detection, int use_graph, void* stream) {
    if (!s) return invalid("system_grab_staged: null system");
    if (!s->in_depth.ptr) return invalid("system_grab_staged: call sicp_system_staging first");
    const int nd = num_detection > 0 ? num_detection : 0;
    const sicp_image lab = nd > 0 ? s->in_labels.img() : sicp_image{nullptr, 0, 0, 0};
    // The first frame (tracking initialisation) and timed frames run eagerly.
    if (!use_graph || s->frame_index == 0 || s->timing)
        return sicp_system_grab_image(s, s->in_depth.img(), s->in_rgb.img(), lab, nd, stream);
    if (!s->graph[0] || s->graph_nd != nd) {
        const int rc = capture_graphs(s, nd);
        if (rc) return rc;
    }
    const int p = (s->next == &s->fa) ? 0 : 1;
    SICP_HIP(hipGraphLaunch(s->graph[p], (hipStream_t)stream));
    std::swap(s->next, s->last);  // what the recorded frame's host code does
    s->frame_index++;
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
