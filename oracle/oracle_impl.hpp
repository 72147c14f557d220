// oracle_impl.hpp — internal declarations shared by the oracle's translation
// units. TEST INFRASTRUCTURE ONLY (see oracle.h).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../semantic_icpslam_amd/csrc/sicp_expf.hpp"
#include "oracle.h"
#include "oracle_math.hpp"

namespace oracle {

// The FP32 restatement uses the project exponential (identical on the GPU);
// the FP64 transcription uses the exact double exponential.
template <class R> inline R rexp(R x);
template <> inline float rexp<float>(float x) { return sicp::expf_cw(x); }
template <> inline double rexp<double>(double x) { return std::exp(x); }

template <class U, class R> inline U cast_out(R v) { return (U)v; }
template <> inline uint8_t cast_out<uint8_t, float>(float v) { return (uint8_t)(int)v; }

template <class R>
void filter_depth(const uint16_t* depth, int cols, int rows, R scale_inv, R cutoff, R* raw,
                  R* filtered);
template <class R>
void compute_vmap(const R* depth, int cols, int rows, R invfx, R invfy, R cx, R cy, R cutoff,
                  R* vmap);
template <class R> void compute_nmap(const R* vmap, int cols, int rows, R* nmap);
template <class T, class U, class R>
void pyrdown(const T* src, int scols, int srows, U* dst, int dcols, int drows);
void image_to_intensity(const uint8_t* rgb, int cols, int rows, uint8_t* gray);
void derivative_image(const uint8_t* img, int cols, int rows, int16_t* dx, int16_t* dy);
template <class R>
void resize_map(const R* vsrc, const R* nsrc, int scols, int srows, R* vdst, R* ndst, int dcols,
                int drows);

// ---- ICP ----
template <class R> struct Intr { R fx, fy, cx, cy; };
// Sums of the 29 per-pixel values (Reduction.cu:292-322), accumulated in
// double; per-pixel values computed in R.
template <class R>
void icp_sums(const R* vcurr, const R* vlast, const R* ncurr, const R* nlast, int cols, int rows,
              const Pose<R>& curr, const Pose<R>& last, Intr<R> K, double* sums, float* rows_out);
template <class R>
void rgb_sums(const uint8_t* nimg, const uint8_t* limg, const R* nvmap, const R* lvmap,
              const int16_t* dx, const int16_t* dy, int cols, int rows, const Pose<R>& curr,
              const Pose<R>& last, Intr<R> K, double* sums, int* res2, float* sigma_out,
              float* rows_out = nullptr);
// The precision form's FP32 compensated sums of per-pixel rows (29 floats a
// pixel, slot 28 = 1 where the pixel contributes) in the GPU reduction's
// order; `compensate` = false: plain FP32 in the same order.
void sums_k32(const float* rows29, int n, bool compensate, bool spread, float* sums);
void create_matrix(const double* sums29, double* A, double* b);
bool cholesky_solve6(const double* A, const double* b, double* x);
void se3_exp(const double* xi, double* T);
void mat4_mul(const double* A, const double* B, double* C);
void rigid_inverse(const double* T, double* Ti);
void euler_sin_norm(const double* T, double* rot_norm, double* trans_norm);

// The decisions of ComputeSE3's inner loop (Tracking.cc:449-481), one call
// per Gauss-Newton iteration after its ICP sums (and, with the photometric
// term, after RGBStep's failure test, :425-432): the count / NaN failure,
// the error-increase counter with its candidate pose, the break-out after a
// third increase in a row. `capture` is set when the caller must store its
// current pose as the candidate (the first increase of a run). The
// reference's `candidateICP = icpError` sits outside the braces of
// `if (breakCount == 1)`, so it is taken on every increase (Q11). The state
// carries across pyramid levels, as the reference's locals do.
enum { kSe3Apply = 0, kSe3Fail = 1, kSe3Restore = 2 };
template <class R> struct Se3Control {
    R last_error = (R)100.f;  // Tracking.cc:391
    R candidate_icp = (R)0;
    int break_count = 0;
    bool capture = false;
    int step(R icp_error, int icp_count, R rgb_error, R w, int min_count) {
        capture = false;
        if (isnan_(icp_error) || icp_count < min_count) return kSe3Fail;  // :452-456
        const R curr = rgb_error * w + icp_error;                          // :458
        if (last_error < curr) {
            ++break_count;
            if (break_count == 1) capture = true;
            candidate_icp = icp_error;
        } else {
            break_count = 0;
        }
        if (break_count > 2) return kSe3Restore;  // :469-478
        last_error = curr;
        return kSe3Apply;
    }
};
// The final gate of ComputeSE3 (Tracking.cc:505-516): the frame is tracked
// when the last ICP error is below the threshold and the motion from the
// frame's initial pose `pose0` to `pose` is small (sine norm of the XYZ Euler
// angles and translation norm at most 0.1).
template <class R> bool se3_gate(R icp_error, R thresh, const double* pose0, const double* pose) {
    double inv0[16], p[16], rn, tn;
    rigid_inverse(pose0, inv0);
    mat4_mul(inv0, pose, p);
    euler_sin_norm(p, &rn, &tn);
    return icp_error < thresh && rn <= 0.1 && tn <= 0.1;
}

// ---- map ----
struct Voxel {
    double sdf;
    uint8_t w, r, g, b;
    uint8_t label, label2;
    uint16_t count, count2;
    uint8_t r2, g2, b2;
};

struct MapCounters {
    uint64_t probes = 0, interp = 0, steps = 0;
    // raycast outcomes: no depth range, back-face exit, no zero crossing,
    // normal rejected, surface found
    uint64_t ray_norange = 0, ray_back = 0, ray_nocross = 0, ray_nonormal = 0, ray_hit = 0;
    uint64_t interp_reads = 0;  // readSdfInterped calls of the last raycast
};

template <class R> struct OMap {
    sicp_map_config cfg;
    R voxel, voxel_inv, trunc, step_scale, block_width, dmin, dmax;
    uint32_t nb;
    std::vector<sicp_hash_entry> table;   // 2 * nb
    std::vector<sicp_hash_entry> excess;  // num_blocks, indexed by heap slot
    std::vector<sicp_block> blocks;       // num_blocks
    std::vector<R> sdf;                   // num_blocks * 512
    std::vector<sicp_voxel_geom> geom;    // colour + weight (sdf lives in `sdf`)
    std::vector<sicp_voxel_sem> sem;
    uint32_t heap_used = 0;
    std::vector<sicp_block> visible;
    sicp_map_stats stats{};
    MapCounters ctr;
};

uint64_t pack_key(int x, int y, int z);
bool key_in_range(int x, int y, int z);
uint32_t hash_bucket(int x, int y, int z, uint32_t nb);

template <class R> void map_init(OMap<R>& m, const sicp_map_config& cfg);
template <class R> void map_reset(OMap<R>& m);
template <class R> int map_lookup(const OMap<R>& m, int x, int y, int z);
template <class R> void unify_neighbor(OMap<R>& m);
template <class R>
uint32_t fuse_map_color(OMap<R>& m, const R* depth, const uint8_t* rgb, const R* nmap,
                        const uint8_t* mask, int cols, int rows, int num_detection,
                        const Pose<R>& pose, Intr<R> K);
template <class R>
uint32_t check_visibility(OMap<R>& m, const Pose<R>& pose, int cols, int rows, Intr<R> K,
                          R depth_max, R depth_min);
template <class R>
int create_rendering_blocks(OMap<R>& m, R* zmin, R* zmax, int zcols, int zrows, R depth_max,
                            R depth_min, const Pose<R>& pose, Intr<R> K);
template <class R>
void raycast(OMap<R>& m, R* vmap, R* nmap, int cols, int rows, const R* zmin, const R* zmax,
             int zcols, const Pose<R>& pose, R invfx, R invfy, R cx, R cy);

}  // namespace oracle
