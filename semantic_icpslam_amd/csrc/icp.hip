// icp.hip — point-to-plane ICP reduction and the device-resident
// Gauss-Newton loop of Tracker::ComputeSE3, for gfx950.
//
// One launch per iteration. Each thread accumulates the 29 per-pixel values
// (27 upper A|b entries, r^2, valid) of its grid-stride pixels in registers;
// each 64-lane wavefront folds them with cross-lane shuffles; the four
// wavefronts of a workgroup combine through LDS into one partial row; the
// last workgroup to arrive (release/acquire counter) sums the partial rows in
// workgroup order — a fixed order, so results are bitwise reproducible; the
// accumulation is FP64 throughout —
// and, in the tracking mode, solves the 6x6 system in FP64 and updates the
// pose on the device.
#include "se3_math.hpp"
#include "sicp_kernels.hpp"

namespace sicp {

namespace {

struct IcpGeom {
    int cols, rows, n;
    float fx, fy, cx, cy;
};

// ICPReduce::searchPoint + getRow (Tracking/Reduction.cu:228-290) for one
// pixel, in the reference's FP32 order with its FP64 promotions (Q9).
SICP_DEV bool icp_row(const View<float4>& vcurr, const View<float4>& vlast,
                      const View<float4>& ncurr, const View<float4>& nlast, int x, int y,
                      const PoseF& cur, const PoseF& last, const IcpGeom& g, float row[7]) {
    float4 vc4 = vcurr.at(y, x);
    v3 vcurr_c = xyz(vc4);
    if (isnan_(vcurr_c.x) || (double)vcurr_c.z < 1e-3) return false;
    v3 vcurr_g = add(mv(cur.R, vcurr_c), cur.t);
    v3 vcurr_p = mv(last.Rinv, sub(vcurr_g, last.t));
    float invz = (float)(1.0 / (double)vcurr_p.z);
    int u = (int)((double)(vcurr_p.x * invz * g.fx + g.cx) + 0.5);
    int v = (int)((double)(vcurr_p.y * invz * g.fy + g.cy) + 0.5);
    if (u < 0 || v < 0 || u >= g.cols || v >= g.rows) return false;
    v3 vlast_c = xyz(vlast.at(v, u));
    v3 vlast_g = add(mv(last.R, vlast_c), last.t);
    v3 ncurr_c = xyz(ncurr.at(y, x));
    v3 ncurr_g = mv(cur.R, ncurr_c);
    v3 nlast_c = xyz(nlast.at(v, u));
    v3 nlast_g = mv(last.R, nlast_c);
    float dist = norm(sub(vlast_g, vcurr_g));
    float sine = norm(cross(ncurr_g, nlast_g));
    if (!(sine < 0.6f && dist <= 0.1f && !isnan_(ncurr_c.x) && !isnan_(nlast_c.x))) return false;
    v3 nl = mv(last.Rinv, nlast_g);
    v3 vcl = mv(last.Rinv, sub(vcurr_g, last.t));
    v3 vll = mv(last.Rinv, sub(vlast_g, last.t));
    v3 mn = neg(nl);
    v3 c = cross(nl, vll);
    row[0] = mn.x;
    row[1] = mn.y;
    row[2] = mn.z;
    row[3] = c.x;
    row[4] = c.y;
    row[5] = c.z;
    row[6] = dot(mn, sub(vcl, vll));
    return true;
}

// Per-thread accumulation + workgroup reduction into partials[blockIdx.x].
// Returns true in the one workgroup that arrives last; its thread j < 29 then
// holds the total of sum j in `total`.
SICP_DEV bool icp_accumulate(const View<float4>& vcurr, const View<float4>& vlast,
                             const View<float4>& ncurr, const View<float4>& nlast,
                             const PoseF& cur, const PoseF& last, const IcpGeom& g,
                             double* partials, unsigned* counter, double& total) {
    // The per-pixel products are FP32 (as in the reference); their sums are
    // carried in FP64, so the totals equal the exact sum of the FP32 products
    // to double rounding, whatever the reduction order.
    double acc[kIcpSums];
#pragma unroll
    for (int j = 0; j < kIcpSums; ++j) acc[j] = 0.0;
    const int stride = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < g.n; i += stride) {
        int y = i / g.cols;
        int x = i - y * g.cols;
        float row[7];
        if (icp_row(vcurr, vlast, ncurr, nlast, x, y, cur, last, g, row)) {
            int k = 0;
#pragma unroll
            for (int a = 0; a < 7; ++a)
#pragma unroll
                for (int b = a; b < 7; ++b) acc[k++] += (double)(row[a] * row[b]);
            acc[28] += 1.0;
        }
    }
    // Wavefront fold: 64 -> 1 lanes.
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int j = 0; j < kIcpSums; ++j) acc[j] += __shfl_xor(acc[j], off, 64);
    }
    __shared__ double wave_sums[kIcpThreads / 64][kIcpSums];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int j = 0; j < kIcpSums; ++j) wave_sums[wave][j] = acc[j];
    }
    __syncthreads();
    if (threadIdx.x < kIcpSums) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kIcpThreads / 64; ++w) s += wave_sums[w][threadIdx.x];
        // column-major: sum j of every workgroup is contiguous
        partials[threadIdx.x * kIcpMaxBlocks + blockIdx.x] = s;
        __threadfence();
    }
    __shared__ int is_last;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned prev = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL,
                                               __HIP_MEMORY_SCOPE_AGENT);
        is_last = (prev == gridDim.x - 1) ? 1 : 0;
    }
    __syncthreads();
    if (!is_last) return false;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    // Wavefront w reduces sums w, w+4, ...: lane-strided partial sums over the
    // workgroups, then a butterfly — a fixed order for a given grid size.
    __shared__ double totals[kIcpSums];
    const int nb = (int)gridDim.x;
    for (int j = wave; j < kIcpSums; j += kIcpThreads / 64) {
        double s = 0.0;
        for (int b = lane; b < nb; b += 64) s += partials[j * kIcpMaxBlocks + b];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        if (lane == 0) totals[j] = s;
    }
    __syncthreads();
    if (threadIdx.x < kIcpSums) total = totals[threadIdx.x];
    if (threadIdx.x == 0) *counter = 0u;  // ready for the next launch on the stream
    return true;
}

__global__ __launch_bounds__(kIcpThreads) void icp_sums_kernel(
    View<float4> vcurr, View<float4> vlast, View<float4> ncurr, View<float4> nlast, PoseArg pc,
    PoseArg pl, IcpGeom g, double* partials, unsigned* counter, float* sums_out) {
    const PoseF cur = load_pose(pc);
    const PoseF last = load_pose(pl);
    double total = 0.0;
    if (!icp_accumulate(vcurr, vlast, ncurr, nlast, cur, last, g, partials, counter, total))
        return;
    if (threadIdx.x < kIcpSums) sums_out[threadIdx.x] = (float)total;
}

// One iteration of the inner loop of ComputeSE3 (Tracking.cc:433-514,
// icpOnly branch :488-490) executed by the last workgroup.
SICP_DEV void gauss_newton_step(TrackState* st, int min_count) {
    const float r0 = st->sums[27], r1 = st->sums[28];
    const float icp_error = sqrtf(r0) / r1;
    const int icp_count = (int)r1;
    st->icp_count = icp_count;
    if (isnan_(icp_error) || icp_count < min_count) {
        // "track icp failed": roll back (Tracking.cc:452-456)
        for (int i = 0; i < 16; ++i) st->next_pose[i] = st->last_pose[i];
        st->icp_error = icp_error;
        st->failed = 1;
        st->done = 1;
        pose_to_float(st->next_pose, &st->curr_f);
        return;
    }
    const float curr_error = icp_error;
    if (st->last_error < curr_error) {
        st->break_count++;
        if (st->break_count == 1)
            for (int i = 0; i < 16; ++i) st->candidate[i] = st->next_pose[i];
        st->candidate_icp = icp_error;  // assigned on every increase (Q11)
    } else {
        st->break_count = 0;
    }
    if (st->break_count > 2) {
        for (int i = 0; i < 16; ++i) st->next_pose[i] = st->candidate[i];
        st->icp_error = st->candidate_icp;
        st->done = 1;
        pose_to_float(st->next_pose, &st->curr_f);
        return;
    }
    st->last_error = curr_error;
    st->icp_error = icp_error;
    create_matrix(st->sums, st->A, st->b);
    ldlt_solve6(st->A, st->b, st->x);
    double dT[16], inv[16];
    se3_exp(st->x, dT);
    mat4_mul(dT, st->delta, st->delta);
    rigid_inverse(st->delta, inv);
    mat4_mul(st->last_pose, inv, st->next_pose);
    st->applied++;
    pose_to_float(st->next_pose, &st->curr_f);
}

__global__ __launch_bounds__(kIcpThreads) void icp_iterate_kernel(
    View<float4> vcurr, View<float4> vlast, View<float4> ncurr, View<float4> nlast, IcpGeom g,
    int min_count, TrackState* st, double* partials) {
    if (__hip_atomic_load(&st->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    const PoseF cur = to_pose(st->curr_f);
    const PoseF last = to_pose(st->last_f);
    double total = 0.0;
    if (!icp_accumulate(vcurr, vlast, ncurr, nlast, cur, last, g, partials, &st->counter, total))
        return;
    if (threadIdx.x < kIcpSums) st->sums[threadIdx.x] = (float)total;
    __syncthreads();
    if (threadIdx.x == 0) gauss_newton_step(st, min_count);
}

__global__ void track_init_kernel(FrameMeta* next, FrameMeta* last, TrackState* st) {
    set_identity(next->pose);
    set_identity(last->pose);
    pose_to_float(next->pose, &next->pose_f);
    pose_to_float(last->pose, &last->pose_f);
    st->tracked = 1;
    st->failed = 0;
    st->done = 1;
    st->applied = 0;
    st->icp_count = 0;
    st->icp_error = 0.f;
}

// Tracker::TrackFrame (Tracking.cc:157-173) up to the loop of ComputeSE3.
__global__ void track_begin_kernel(FrameMeta* next, FrameMeta* last, TrackState* st,
                                   float thresh) {
    for (int i = 0; i < 16; ++i) {
        st->last_pose[i] = last->pose[i];
        st->next_pose[i] = last->pose[i];
        st->pose0[i] = last->pose[i];
        st->candidate[i] = last->pose[i];
    }
    set_identity(st->delta);
    st->icp_error = 0.f;
    st->last_error = 100.f;
    st->candidate_icp = 0.f;
    st->thresh = thresh;
    st->icp_count = 0;
    st->break_count = 0;
    st->applied = 0;
    st->done = 0;
    st->failed = 0;
    st->tracked = 0;
    st->counter = 0u;
    pose_to_float(st->next_pose, &st->curr_f);
    pose_to_float(st->last_pose, &st->last_f);
}

// Final motion / error gate of ComputeSE3 (Tracking.cc:505-516).
__global__ void track_end_kernel(FrameMeta* next, TrackState* st) {
    int ok = 0;
    if (!st->failed) {
        double inv0[16], p[16], rn, tn;
        rigid_inverse(st->pose0, inv0);
        mat4_mul(inv0, st->next_pose, p);
        euler_sin_norm(p, &rn, &tn);
        ok = (st->icp_error < st->thresh && rn <= 0.1 && tn <= 0.1) ? 1 : 0;
        if (!ok)
            for (int i = 0; i < 16; ++i) st->next_pose[i] = st->last_pose[i];
    }
    st->tracked = ok;
    for (int i = 0; i < 16; ++i) next->pose[i] = st->next_pose[i];
    pose_to_float(next->pose, &next->pose_f);
}

__global__ __launch_bounds__(256) void rollback_maps_kernel(View<float4> vlast, View<float4> nlast,
                                                            View<float4> vnext, View<float4> nnext,
                                                            const TrackState* st) {
    if (st->tracked) return;
    const int n = vnext.cols * vnext.rows;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int y = i / vnext.cols, x = i - y * vnext.cols;
        vnext.at(y, x) = vlast.at(y, x);
        nnext.at(y, x) = nlast.at(y, x);
    }
}

// About four pixels per thread, at most one workgroup per CU.
inline int icp_blocks(int n) {
    int b = (int)div_up(n, kIcpThreads * 4);
    return b < 1 ? 1 : (b > kIcpMaxBlocks ? kIcpMaxBlocks : b);
}

}  // namespace

int launch_icp_sums(View<float4> vcurr, View<float4> vlast, View<float4> ncurr,
                    View<float4> nlast, PoseArg curr, PoseArg last, sicp_intrinsics K,
                    IcpScratch scratch, float* sums_out, hipStream_t s) {
    IcpGeom g{vcurr.cols, vcurr.rows, vcurr.cols * vcurr.rows, K.fx, K.fy, K.cx, K.cy};
    icp_sums_kernel<<<icp_blocks(g.n), kIcpThreads, 0, s>>>(vcurr, vlast, ncurr, nlast, curr,
                                                              last, g, scratch.partials,
                                                              scratch.counter, sums_out);
    SICP_LAUNCH_CHECK("icp_sums_kernel");
    return SICP_OK;
}

int launch_icp_iterate(View<float4> vcurr, View<float4> vlast, View<float4> ncurr,
                       View<float4> nlast, sicp_intrinsics K, int min_count, TrackState* st,
                       double* partials, hipStream_t s) {
    IcpGeom g{vcurr.cols, vcurr.rows, vcurr.cols * vcurr.rows, K.fx, K.fy, K.cx, K.cy};
    icp_iterate_kernel<<<icp_blocks(g.n), kIcpThreads, 0, s>>>(vcurr, vlast, ncurr, nlast, g,
                                                                 min_count, st, partials);
    SICP_LAUNCH_CHECK("icp_iterate_kernel");
    return SICP_OK;
}

int launch_track_init(FrameMeta* next, FrameMeta* last, TrackState* st, hipStream_t s) {
    track_init_kernel<<<1, 1, 0, s>>>(next, last, st);
    SICP_LAUNCH_CHECK("track_init_kernel");
    return SICP_OK;
}

int launch_track_begin(FrameMeta* next, FrameMeta* last, TrackState* st, float thresh,
                       hipStream_t s) {
    track_begin_kernel<<<1, 1, 0, s>>>(next, last, st, thresh);
    SICP_LAUNCH_CHECK("track_begin_kernel");
    return SICP_OK;
}

int launch_track_end(FrameMeta* next, TrackState* st, hipStream_t s) {
    track_end_kernel<<<1, 1, 0, s>>>(next, st);
    SICP_LAUNCH_CHECK("track_end_kernel");
    return SICP_OK;
}

int launch_rollback_maps(View<float4> vlast, View<float4> nlast, View<float4> vnext,
                         View<float4> nnext, const TrackState* st, hipStream_t s) {
    int n = vnext.cols * vnext.rows;
    int blocks = (int)div_up(n, 256 * 4);
    rollback_maps_kernel<<<blocks, 256, 0, s>>>(vlast, nlast, vnext, nnext, st);
    SICP_LAUNCH_CHECK("rollback_maps_kernel");
    return SICP_OK;
}

void se3_update_host(const double* A, const double* b, double* delta, const double* last_pose,
                     double* next_pose, double* x_out) {
    double x[6], dT[16], inv[16];
    ldlt_solve6(A, b, x);
    se3_exp(x, dT);
    mat4_mul(dT, delta, delta);
    rigid_inverse(delta, inv);
    mat4_mul(last_pose, inv, next_pose);
    if (x_out)
        for (int i = 0; i < 6; ++i) x_out[i] = x[i];
}

}  // namespace sicp
