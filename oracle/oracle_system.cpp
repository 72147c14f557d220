// oracle_system.cpp — CPU restatement of the per-frame sequence of
// System::GrabImage (Core/System.cc:131-172): Frame::ResizeImages /
// FillImages (Core/Frame.cc:68-98), Tracker::Track / TrackFrame /
// ComputeSE3 (Tracking/Tracking.cc:65-119,157-173,
// 369-517), Mapping::FuseColor and Mapping::RayTrace (Mapping/Mapping.cc:
// 76-97,142-161). TEST INFRASTRUCTURE ONLY (see oracle.h); also the CPU
// baseline that bench.py times.
#include <omp.h>

#include <type_traits>

#include "oracle_impl.hpp"

namespace oracle {

template <class R> struct OFrame {
    int cols[3], rows[3];
    std::vector<R> range;
    std::vector<uint8_t> color;
    std::vector<R> depth[3], vmap[3], nmap[3];
    std::vector<uint8_t> image[3];
    std::vector<int16_t> dx[3], dy[3];
    double pose[16];

    // Frame::Create (Core/Frame.cc:33-62): level i is cols/2^i x rows/2^i.
    void create(int c, int r) {
        for (int i = 0; i < 3; ++i) {
            cols[i] = c / (1 << i);
            rows[i] = r / (1 << i);
            size_t n = (size_t)cols[i] * rows[i];
            depth[i].assign(n, 0);
            vmap[i].assign(n * 4, 0);
            nmap[i].assign(n * 4, 0);
            image[i].assign(n, 0);
            dx[i].assign(n, 0);
            dy[i].assign(n, 0);
        }
        range.assign((size_t)c * r, 0);
        color.assign((size_t)c * r * 3, 0);
        for (int i = 0; i < 16; ++i) pose[i] = (i % 5 == 0) ? 1.0 : 0.0;
    }
};

template <class R> struct OSystem {
    sicp_system_config cfg;
    OMap<R> map;
    OFrame<R> fa, fb;
    OFrame<R>* next = &fa;
    OFrame<R>* last = &fb;
    int state = 1;  // 1: initialise on the next frame (Tracking.cc:74-78)
    int frame_index = 0;
    int zc, zr;
    std::vector<R> zmin, zmax;
    Intr<R> K[3];
    // Work of the path summed over the frames grabbed (inputs of the
    // algorithmic byte counts): raycasts, rays, march steps, interpolated
    // reads, raycast block lookups, allocation probes, voxels fused, visible
    // blocks fused.
    uint64_t work[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

template <class R> static void fill_images(OSystem<R>& s, OFrame<R>& f, const uint16_t* depth,
                                           const uint8_t* rgb) {
    std::memcpy(f.color.data(), rgb, f.color.size());
    R scale_inv = (R)(1.0 / (double)s.cfg.depth_scale);
    R cutoff = (R)s.cfg.depth_cutoff;
    filter_depth<R>(depth, f.cols[0], f.rows[0], scale_inv, cutoff, f.range.data(),
                    f.depth[0].data());
    image_to_intensity(rgb, f.cols[0], f.rows[0], f.image[0].data());
    for (int i = 1; i < 3; ++i) {
        pyrdown<R, R, R>(f.depth[i - 1].data(), f.cols[i - 1], f.rows[i - 1], f.depth[i].data(),
                         f.cols[i], f.rows[i]);
        pyrdown<uint8_t, uint8_t, float>(f.image[i - 1].data(), f.cols[i - 1], f.rows[i - 1],
                                         f.image[i].data(), f.cols[i], f.rows[i]);
    }
    for (int i = 0; i < 3; ++i) {
        const Intr<R>& k = s.K[i];
        compute_vmap<R>(f.depth[i].data(), f.cols[i], f.rows[i], (R)(1.0 / (double)k.fx),
                        (R)(1.0 / (double)k.fy), k.cx, k.cy, cutoff, f.vmap[i].data());
        compute_nmap<R>(f.vmap[i].data(), f.cols[i], f.rows[i], f.nmap[i].data());
        derivative_image(f.image[i].data(), f.cols[i], f.rows[i], f.dx[i].data(),
                         f.dy[i].data());
    }
}

template <class R> static void resize_images(OFrame<R>& f) {
    for (int i = 1; i < 3; ++i)
        resize_map<R>(f.vmap[i - 1].data(), f.nmap[i - 1].data(), f.cols[i - 1], f.rows[i - 1],
                      f.vmap[i].data(), f.nmap[i].data(), f.cols[i], f.rows[i]);
}

// ICPStep's host side: the 29 sums are float on the device (Reduction.cu:
// 377-384), so the FP32 restatement rounds its FP64 accumulation to float.
template <class R> static void round_sums(double* s);
template <> void round_sums<float>(double* s) {
    for (int i = 0; i < 29; ++i) s[i] = (double)(float)s[i];
}
template <> void round_sums<double>(double*) {}

static void compensated_sums(const std::vector<float>& rows, int n, bool spread, double* sums) {
    float f[29];
    sums_k32(rows.data(), n, true, spread, f);
    for (int j = 0; j < 29; ++j) sums[j] = (double)f[j];
}

// Tracker::ComputeSE3 (Tracking.cc:369-517): the point-to-plane term alone
// (icpOnly, :488-490) or with the weighted photometric term (:400-433,
// :484-487), as the configuration says.
template <class R>
static bool compute_se3(const sicp_system_config& cfg, const Intr<R>* K, OFrame<R>& nf,
                        OFrame<R>& lf, sicp_frame_result* res) {
    double last_pose[16], pose0[16], delta[16], candidate[16];
    std::memcpy(last_pose, lf.pose, sizeof(last_pose));
    std::memcpy(pose0, nf.pose, sizeof(pose0));
    for (int i = 0; i < 16; ++i) delta[i] = (i % 5 == 0) ? 1.0 : 0.0;
    std::memcpy(candidate, nf.pose, sizeof(candidate));
    R icp_error = 0, rgb_error = 0;
    int icp_count = 0, applied = 0;
    Se3Control<R> ctl;
    bool broke = false;
    const bool icp_only = cfg.icp_only != 0;
    const float w = cfg.rgb_weight;
    // The precision form's FP32 compensated sums (the FP64 restatement keeps
    // its own arithmetic).
    const bool k32 = std::is_same<R, float>::value && cfg.icp_sums_fp32 != 0;
    std::vector<float> rows(k32 ? (size_t)nf.cols[0] * nf.rows[0] * 29 : 0);
    for (int i = 2; i >= 0 && !broke; --i) {
        for (int j = 0; j < cfg.icp_iters[i]; ++j) {
            Pose<R> pc = pose_from_d<R>(nf.pose);
            Pose<R> pl = pose_from_d<R>(lf.pose);
            double A_rgb[36], b_rgb[6];
            if (!icp_only) {
                // RGBStep and its failure test (Tracking.cc:402-433)
                double rs[29];
                int res2[2];
                rgb_sums<R>(nf.image[i].data(), lf.image[i].data(), nf.vmap[i].data(),
                            lf.vmap[i].data(), nf.dx[i].data(), nf.dy[i].data(), nf.cols[i],
                            nf.rows[i], pc, pl, K[i], rs, res2, nullptr,
                            k32 ? rows.data() : nullptr);
                if (k32)
                    compensated_sums(rows, nf.cols[i] * nf.rows[i], cfg.icp_spread != 0, rs);
                else
                    round_sums<R>(rs);
                create_matrix(rs, A_rgb, b_rgb);
                R q0 = (R)rs[27], q1 = (R)rs[28];
                rgb_error = std::sqrt(q0) / q1;
                const int rgb_count = (int)q1;
                res->rgb_error = (float)rgb_error;
                res->rgb_count = rgb_count;
                if (isnan_(rgb_error) || rgb_count < cfg.icp_min_count[i]) {
                    std::memcpy(nf.pose, last_pose, sizeof(last_pose));
                    res->icp_error = (float)icp_error;  // the previous iteration's
                    res->iterations = applied;
                    return false;
                }
            }
            double sums[29], A[36], b[6];
            icp_sums<R>(nf.vmap[i].data(), lf.vmap[i].data(), nf.nmap[i].data(),
                        lf.nmap[i].data(), nf.cols[i], nf.rows[i], pc, pl, K[i], sums,
                        k32 ? rows.data() : nullptr);
            if (k32)
                compensated_sums(rows, nf.cols[i] * nf.rows[i], cfg.icp_spread != 0, sums);
            else
                round_sums<R>(sums);
            create_matrix(sums, A, b);
            R r0 = (R)sums[27], r1 = (R)sums[28];
            icp_error = std::sqrt(r0) / r1;
            icp_count = (int)r1;
            res->icp_count = icp_count;
            const int act = ctl.step(icp_error, icp_count, rgb_error, (R)w, cfg.icp_min_count[i]);
            if (act == kSe3Fail) {
                std::memcpy(nf.pose, last_pose, sizeof(last_pose));
                res->icp_error = (float)icp_error;
                res->iterations = applied;
                return false;
            }
            if (ctl.capture) std::memcpy(candidate, nf.pose, sizeof(candidate));
            if (act == kSe3Restore) {
                std::memcpy(nf.pose, candidate, sizeof(candidate));
                icp_error = ctl.candidate_icp;
                broke = true;
                break;
            }
            if (!icp_only)  // matA = matA_rgb * w + matA_icp (Tracking.cc:484-487)
                for (int k = 0; k < 36; ++k) {
                    A[k] = A_rgb[k] * (double)w + A[k];
                    if (k < 6) b[k] = b_rgb[k] * (double)w + b[k];
                }
            double x[6], dT[16], inv[16];
            cholesky_solve6(A, b, x);
            se3_exp(x, dT);
            mat4_mul(dT, delta, delta);
            rigid_inverse(delta, inv);
            mat4_mul(last_pose, inv, nf.pose);
            applied++;
        }
    }
    res->icp_error = (float)icp_error;
    res->iterations = applied;
    if (se3_gate<R>(icp_error, (R)cfg.icp_thresh, pose0, nf.pose)) return true;
    std::memcpy(nf.pose, last_pose, sizeof(last_pose));
    return false;
}

template <class R> static void system_init(OSystem<R>& s, const sicp_system_config& cfg) {
    s.cfg = cfg;
    map_init(s.map, cfg.map);
    s.fa.create(cfg.cols, cfg.rows);
    s.fb.create(cfg.cols, cfg.rows);
    for (int i = 0; i < 3; ++i) {
        // Frame::SetK (Core/Frame.cc:580-588): K / 2^i in float.
        float d = (float)(1 << i);
        s.K[i] = Intr<R>{(R)(cfg.K.fx / d), (R)(cfg.K.fy / d), (R)(cfg.K.cx / d),
                         (R)(cfg.K.cy / d)};
    }
    s.zc = (cfg.cols + 7) / 8;
    s.zr = (cfg.rows + 7) / 8;
    s.zmin.assign((size_t)s.zc * s.zr, 0);
    s.zmax.assign((size_t)s.zc * s.zr, 0);
}

template <class R>
static int system_grab(OSystem<R>& s, const uint16_t* depth, const uint8_t* rgb,
                       const uint8_t* labels, int num_detection, sicp_frame_result* res) {
    *res = sicp_frame_result{};
    res->frame_index = s.frame_index++;
    // Tracker::GrabFrame (Tracking.cc:246-280)
    resize_images(*s.last);
    fill_images(s, *s.next, depth, rgb);
    bool tracked;
    if (s.state == 1) {
        // InitTracking: both poses identity, then SwapFrame (Tracking.cc:74-78,233-240)
        for (int i = 0; i < 16; ++i) s.next->pose[i] = s.last->pose[i] = (i % 5 == 0) ? 1 : 0;
        std::swap(s.next, s.last);
        s.state = 0;
        tracked = true;
    } else {
        std::memcpy(s.next->pose, s.last->pose, sizeof(s.next->pose));  // TrackFrame (:166)
        tracked = compute_se3(s.cfg, s.K, *s.next, *s.last, res);
        if (tracked) std::swap(s.next, s.last);
    }
    res->tracked = tracked ? 1 : 0;
    if (tracked) {
        OFrame<R>& f = *s.last;
        Pose<R> p = pose_from_d<R>(f.pose);
        uint32_t nv = fuse_map_color(s.map, f.range.data(), f.color.data(), f.nmap[0].data(),
                                     labels, f.cols[0], f.rows[0], labels ? num_detection : 0, p,
                                     s.K[0]);
        res->visible_blocks = nv;
        if (nv > 0) {
            int any = create_rendering_blocks(s.map, s.zmin.data(), s.zmax.data(), s.zc, s.zr,
                                              (R)s.cfg.map.depth_max, (R)s.cfg.map.depth_min, p,
                                              s.K[0]);
            if (any) {
                raycast(s.map, f.vmap[0].data(), f.nmap[0].data(), f.cols[0], f.rows[0],
                        s.zmin.data(), s.zmax.data(), s.zc, p, (R)(1.0 / (double)s.K[0].fx),
                        (R)(1.0 / (double)s.K[0].fy), s.K[0].cx, s.K[0].cy);
                s.work[0] += 1;
                s.work[1] += (uint64_t)f.cols[0] * f.rows[0];
                s.work[2] += s.map.ctr.steps;
                s.work[3] += s.map.ctr.interp_reads;
                s.work[4] += s.map.ctr.interp;
            }
        }
        s.work[5] += s.map.ctr.probes;
        s.work[6] += s.map.stats.updated_voxels;
        s.work[7] += nv;
    }
    res->allocated_blocks = s.map.heap_used;
    std::memcpy(res->pose, s.last->pose, sizeof(res->pose));
    return SICP_OK;
}

// One ComputeSE3 (Tracking.cc:369-517, point-to-plane only) on given
// pyramids: the next frame's measured vertex / normal maps and the last
// frame's model maps at three levels (float4, level l of cols/2^l x
// rows/2^l), from the last frame's pose; arithmetic as `fp64` says. Writes
// the resulting pose (the last pose when the frame is not tracked) and the
// frame result; returns the tracked flag.
template <class R>
static int track_frame(const sicp_system_config& cfg, const float* const* nv,
                       const float* const* nn, const float* const* lv, const float* const* ln,
                       const double* last_pose, double* pose_out, sicp_frame_result* res) {
    OFrame<R> nf, lf;
    nf.create(cfg.cols, cfg.rows);
    lf.create(cfg.cols, cfg.rows);
    Intr<R> K[3];
    for (int i = 0; i < 3; ++i) {
        const float d = (float)(1 << i);  // Frame::SetK (Core/Frame.cc:580-588)
        K[i] = Intr<R>{(R)(cfg.K.fx / d), (R)(cfg.K.fy / d), (R)(cfg.K.cx / d),
                       (R)(cfg.K.cy / d)};
        const size_t n = (size_t)nf.cols[i] * nf.rows[i] * 4;
        for (size_t k = 0; k < n; ++k) {
            nf.vmap[i][k] = (R)nv[i][k];
            nf.nmap[i][k] = (R)nn[i][k];
            lf.vmap[i][k] = (R)lv[i][k];
            lf.nmap[i][k] = (R)ln[i][k];
        }
    }
    std::memcpy(lf.pose, last_pose, sizeof(lf.pose));
    std::memcpy(nf.pose, last_pose, sizeof(nf.pose));  // TrackFrame (Tracking.cc:166)
    *res = sicp_frame_result{};
    const bool ok = compute_se3<R>(cfg, K, nf, lf, res);
    res->tracked = ok ? 1 : 0;
    std::memcpy(pose_out, nf.pose, sizeof(nf.pose));
    std::memcpy(res->pose, nf.pose, sizeof(nf.pose));
    return ok ? 1 : 0;
}

}  // namespace oracle

using namespace oracle;

struct OracleSystemHandle {
    int fp64;
    OSystem<float> f;
    OSystem<double> d;
};

extern "C" {

int oracle_set_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
}

int oracle_max_threads(void) { return omp_get_max_threads(); }

void* oracle_system_create(const sicp_system_config* cfg, int fp64) {
    OracleSystemHandle* h = new OracleSystemHandle();
    h->fp64 = fp64;
    if (fp64)
        system_init(h->d, *cfg);
    else
        system_init(h->f, *cfg);
    return h;
}

void oracle_system_destroy(void* sys) { delete (OracleSystemHandle*)sys; }

int oracle_system_grab(void* sys, const uint16_t* depth, const uint8_t* rgb,
                       const uint8_t* labels, int num_detection, sicp_frame_result* out) {
    OracleSystemHandle* h = (OracleSystemHandle*)sys;
    return h->fp64 ? system_grab(h->d, depth, rgb, labels, num_detection, out)
                   : system_grab(h->f, depth, rgb, labels, num_detection, out);
}

void oracle_system_work(void* sys, uint64_t* out8) {
    OracleSystemHandle* h = (OracleSystemHandle*)sys;
    std::memcpy(out8, h->fp64 ? h->d.work : h->f.work, 8 * sizeof(uint64_t));
}

void* oracle_system_map(void* sys) {
    // Only the FP32 system's map is exposed through the map handle API.
    (void)sys;
    return nullptr;
}

// ComputeSE3's loop (Tracking.cc:398-503) driven by a given stream of
// per-iteration results instead of images: the same controller as the
// system restatement above, with the pose replaced by a marker (0: the
// frame's initial pose; k + 1: the pose after the update of iteration k).
// out4 = {final marker, failed, updates applied, iterations run}; the
// final ICP error (the candidate's after a break-out) into out_error.
// Returns 0, or -1 when the stream runs out before the loop ends.
int oracle_se3_control_trace(const int* iters3, const int* min_count3, const float* icp_error,
                             const int* icp_count, const float* rgb_error, const int* rgb_count,
                             float w, int n, int* out4, float* out_error) {
    Se3Control<float> ctl;
    int marker = 0, candidate = 0, applied = 0, k = 0;
    float err = 0.f, rgb = 0.f;
    bool broke = false;
    for (int l = 2; l >= 0 && !broke; --l) {
        for (int j = 0; j < iters3[l]; ++j, ++k) {
            if (k >= n) return -1;
            if (rgb_error) {
                rgb = rgb_error[k];
                if (isnan_(rgb) || rgb_count[k] < min_count3[l]) {
                    out4[0] = 0;
                    out4[1] = 1;
                    out4[2] = applied;
                    out4[3] = k + 1;
                    *out_error = err;
                    return 0;
                }
            }
            err = icp_error[k];
            const int act = ctl.step(err, icp_count[k], rgb, w, min_count3[l]);
            if (act == kSe3Fail) {
                out4[0] = 0;
                out4[1] = 1;
                out4[2] = applied;
                out4[3] = k + 1;
                *out_error = err;
                return 0;
            }
            if (ctl.capture) candidate = marker;
            if (act == kSe3Restore) {
                marker = candidate;
                err = ctl.candidate_icp;
                broke = true;
                ++k;
                break;
            }
            marker = k + 1;
            ++applied;
        }
    }
    out4[0] = marker;
    out4[1] = 0;
    out4[2] = applied;
    out4[3] = k;
    *out_error = err;
    return 0;
}

int oracle_track_frame(const sicp_system_config* cfg, int fp64, const float* const* next_vmap,
                       const float* const* next_nmap, const float* const* last_vmap,
                       const float* const* last_nmap, const double* last_pose, double* pose_out,
                       sicp_frame_result* res) {
    sicp_system_config c = *cfg;
    c.icp_only = 1;
    return fp64 ? track_frame<double>(c, next_vmap, next_nmap, last_vmap, last_nmap, last_pose,
                                      pose_out, res)
                : track_frame<float>(c, next_vmap, next_nmap, last_vmap, last_nmap, last_pose,
                                     pose_out, res);
}

// The final gate of ComputeSE3 (Tracking.cc:505-516) on given poses.
int oracle_se3_gate(float icp_error, float thresh, const double* pose0, const double* pose) {
    return se3_gate<float>(icp_error, thresh, pose0, pose) ? 1 : 0;
}

void oracle_system_model(void* sys, int level, float* vmap, float* nmap) {
    OracleSystemHandle* h = (OracleSystemHandle*)sys;
    auto go = [&](auto& s) {
        auto& f = *s.last;
        size_t n = f.vmap[level].size();
        for (size_t i = 0; i < n; ++i) {
            vmap[i] = (float)f.vmap[level][i];
            nmap[i] = (float)f.nmap[level][i];
        }
    };
    if (h->fp64)
        go(h->d);
    else
        go(h->f);
}

}  // extern "C"
