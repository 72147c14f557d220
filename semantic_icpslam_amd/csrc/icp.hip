// icp.hip — point-to-plane ICP reduction and the device-resident
// Gauss-Newton loop of Tracker::ComputeSE3, for gfx950.
//
// One launch per Gauss-Newton iteration. Every workgroup first redoes the
// previous iteration's update: it sums the previous launch's partial rows in
// a fixed order (results are bitwise reproducible) and runs the FP64 6x6
// solve and the SE(3) update on one wavefront; workgroup 0 publishes the
// state. Then its pixel pass gives each thread four pixels with every memory
// access of a phase in flight together: the current frame's vertex and
// normal (coalesced, issued before the update), then the projective gathers
// from the last frame. The 29 per-pixel values (27 upper A|b entries, r^2,
// valid) are FP32 products, as in the reference, accumulated in FP64 (or
// compensated FP32); a workgroup reduces them by transposition through LDS
// into one partial row. No workgroup waits for another: on MI355X the kernel
// boundary is cheaper than an in-kernel hand-off, whose agent-scope fences
// write back and invalidate the per-XCD L2 caches.
#include "se3_math.hpp"
#include "sicp_kernels.hpp"

#include <type_traits>

namespace sicp {

#ifdef SICP_ICP_PHASES
// Profiling build: stamps of launch k of the tracking loop, workgroup w.
constexpr int kPhaseLaunches = 24, kPhaseStamps = 8;
__device__ unsigned long long g_icp_phase[kPhaseLaunches][kIcpMaxBlocks][kPhaseStamps];
#endif

namespace {

constexpr int kPixPerThread = 4;
// Images of at most this many pixels take one pixel per thread instead (four
// times the workgroups, each lane a quarter of the arithmetic).
constexpr int kIcpSpreadPixels = 320 * kIcpThreads;
// Every kernel here runs kIcpThreads threads per workgroup; the constant is
// used instead of blockDim, whose value the kernel would otherwise load from
// the dispatch packet (a memory round trip before the first address).

struct IcpGeom {
    int cols, rows, n;
    float fx, fy, cx, cy;
};

// A workgroup's place among the workgroups that share one image: the launch
// grid itself, or, in a launch serving several sequences, the sequence's own
// range of workgroups.
struct Wg {
    int id, n;
};
SICP_DEV Wg grid_wg() { return Wg{(int)blockIdx.x, (int)gridDim.x}; }

// ---- phase stamps (profiling build only) -----------------------------------
// Built with -DSICP_ICP_PHASES (`make phases`), every workgroup of the
// tracking loop's launches records eight wall-clock stamps (s_memrealtime,
// 100 MHz): 0 start, 1 partial rows folded, 6 6x6 system solved, 7 SE(3)
// exponential done, 2 Gauss-Newton done, 3 state ready, 4 pixel pass done,
// 5 end. The product build compiles none of it.
#ifdef SICP_ICP_PHASES
SICP_DEV long long* phase_lds() {
    __shared__ long long v[kPhaseStamps];
    return v;
}
SICP_DEV void phase_mark(int slot) {
    if (threadIdx.x == 0) phase_lds()[slot] = wall_clock64();
}
#else
SICP_DEV void phase_mark(int) {}
#endif

// Kernel arguments of the launches that serve several sequences.
struct IcpSeqArgs {
    View<float4> vcurr, vlast, ncurr, nlast;
    IcpGeom g;
    TrackParams tp;
    const TrackState* st_in;
    const FrameMeta* last;
    TrackState* st_out;
    const double* partials_prev;
    double* partials;
    int nb_prev, min_count_prev;
};
struct IcpBatchArgs {
    int n;
    int wg0[kMaxBatch + 1];  // sequence q: workgroups wg0[q] .. wg0[q + 1] - 1
    IcpSeqArgs seq[kMaxBatch];
};
struct FinishSeqArgs {
    TrackParams tp;
    const TrackState* st_in;
    const FrameMeta* last;
    const double* partials_prev;
    TrackState* st_final;
    FrameMeta* next;
    View<float4> vlast, nlast, vnext, nnext;
    int nb_prev, min_count_prev;
};
struct FinishBatchArgs {
    int n;
    int wg0[kMaxBatch + 1];
    FinishSeqArgs seq[kMaxBatch];
};

// ICPReduce::searchPoint up to the gather (Tracking/Reduction.cu:231-243):
// current vertex to world, into the last camera, projected with the
// reference's FP64 promotions (invz = 1.0 / z in double stored to float,
// +0.5 added in double before the truncating cast; Q9).
struct IcpProj {
    v3 vcurr_g, vcurr_p;  // current vertex in world and in the last camera
    int u, v;
    bool ok;
};
SICP_DEV IcpProj icp_project(float4 vc4, const PoseF& cur, const PoseF& last, const IcpGeom& g) {
    IcpProj p;
    v3 vcurr_c = xyz(vc4);
    p.ok = !(isnan_(vcurr_c.x) || (double)vcurr_c.z < 1e-3);
    p.vcurr_g = add(mv(cur.R, vcurr_c), cur.t);
    const v3 vcurr_p = mv(last.Rinv, sub(p.vcurr_g, last.t));
    p.vcurr_p = vcurr_p;
    float invz = (float)(1.0 / (double)vcurr_p.z);
    p.u = (int)((double)(vcurr_p.x * invz * g.fx + g.cx) + 0.5);
    p.v = (int)((double)(vcurr_p.y * invz * g.fy + g.cy) + 0.5);
    p.ok = p.ok && !(p.u < 0 || p.v < 0 || p.u >= g.cols || p.v >= g.rows);
    return p;
}

// The rest of searchPoint (:244-257) and getRow (:260-290): correspondence
// test, then J = [-n, n x v_last], r = -n . (v_curr - v_last) in the last
// camera's frame.
SICP_DEV bool icp_finish(const IcpProj& p, float4 nc4, float4 vl4, float4 nl4, const PoseF& cur,
                         const PoseF& last, float row[7]) {
    v3 vlast_g = add(mv(last.R, xyz(vl4)), last.t);
    v3 ncurr_c = xyz(nc4);
    v3 ncurr_g = mv(cur.R, ncurr_c);
    v3 nlast_c = xyz(nl4);
    v3 nlast_g = mv(last.R, nlast_c);
    float dist = norm(sub(vlast_g, p.vcurr_g));
    float sine = norm(cross(ncurr_g, nlast_g));
    if (!(sine < 0.6f && dist <= 0.1f && !isnan_(ncurr_c.x) && !isnan_(nlast_c.x))) return false;
    v3 nl = mv(last.Rinv, nlast_g);
    v3 vcl = mv(last.Rinv, sub(p.vcurr_g, last.t));
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

// RGBReduction::FindCorresp's projection into the last frame
// (Reduction.cu:436-439): __float2int_rd(fx * x / z + cx + 0.5), the sum
// taken in double and narrowed to float by the intrinsic's argument. `p` is
// the current vertex in the last camera (the same value as ICP's).
SICP_DEV bool rgb_project(v3 p, const IcpGeom& g, int& u, int& v) {
    const float tu = (float)((double)(g.fx * p.x / p.z + g.cx) + 0.5);
    const float tv = (float)((double)(g.fy * p.y / p.z + g.cy) + 0.5);
    u = f2i(floorf(tu));
    v = f2i(floorf(tv));
    return u >= 0 && v >= 0 && u < g.cols && v < g.rows;
}

// The rest of FindCorresp (:441-452): the last vertex valid and within 0.1 of
// the current one (float4 distance, w included), the last intensity
// non-zero. The record is {last pixel index, intensity difference}, or
// {-1, 0} without a correspondence.
SICP_DEV int2 rgb_record(int key, float4 vc, float4 vl, int li, int u, int v, int cols) {
    const float d0 = vl.x - vc.x, d1 = vl.y - vc.y, d2 = vl.z - vc.z, d3 = vl.w - vc.w;
    const float dist = sqrtf(d0 * d0 + d1 * d1 + d2 * d2 + d3 * d3);
    if (isnan_(vl.x) || !(dist < 0.1f) || li == 0) return make_int2(-1, 0);
    return make_int2(v * cols + u, key - li);
}

// RGBReduction::GetRow (:482-521) of one correspondence: the robust weight
// 1 / (sigma + |diff|), J = w [left, point x left], r = -w diff, with the
// last vertex in the current camera.
SICP_DEV void rgb_row(int diff, float4 vl, int gx, int gy, float sigma, const PoseF& cur,
                      const PoseF& last, const IcpGeom& g, float row[7]) {
    float w = sigma + (float)abs(diff);
    w = (double)w < 1e-3 ? 1.0f : 1.0f / w;
    if ((double)sigma < 1e-6) w = 1.0f;
    const v3 point = mv(cur.Rinv, sub(add(mv(last.R, xyz(vl)), last.t), cur.t));
    const float gxf = (float)((double)gx / 9.0), gyf = (float)((double)gy / 9.0);
    const float invz = 1.0f / point.z;
    v3 left;
    left.x = gxf * g.fx * invz;
    left.y = gyf * g.fy * invz;
    left.z = -(left.x * point.x + left.y * point.y) * invz;
    const v3 c = cross(point, left);
    row[0] = w * left.x;
    row[1] = w * left.y;
    row[2] = w * left.z;
    row[3] = w * c.x;
    row[4] = w * c.y;
    row[5] = w * c.z;
    row[6] = -w * (float)diff;
}

// The photometric views and outputs of the correspondence pass.
struct RgbPass {
    View<int16_t> key;
    View<uint8_t> ilast;
    int2* rec;
    uint32_t* ipart;
};

// The current frame's vertices, normals and photometric keys of one thread's
// pixels: they do not depend on the pose, so their loads are issued before
// the pose is known.
struct IcpBatch {
    float4 vc[kPixPerThread], nc[kPixPerThread];
    int key[kPixPerThread];
};
template <bool kRgb, int NP>
SICP_DEV void icp_load_batch(IcpBatch& bt, int base, int stride, const View<float4>& vcurr,
                             const View<float4>& ncurr, const RgbPass& rp, const IcpGeom& g) {
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const int i = base + k * stride;
        const int ii = i < g.n ? i : 0;
        const int y = ii / g.cols, x = ii - y * g.cols;
        bt.vc[k] = vcurr.at(y, x);
        bt.nc[k] = ncurr.at(y, x);
        if (kRgb) bt.key[k] = rp.key.at(y, x);
        if (i >= g.n) {
            bt.vc[k].x = qnanf();
            bt.key[k] = -1;
        }
    }
}

// Sum of two u32 values over the workgroup (wrapping, so any order gives the
// same result, as the reference's int sums). Every thread gets the totals.
SICP_DEV uint2 block_sum_u32x2(uint32_t a, uint32_t b) {
    __shared__ uint32_t red[2][kIcpThreads / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = a;
        red[1][threadIdx.x >> 6] = b;
    }
    __syncthreads();
    uint2 r = make_uint2(0u, 0u);
#pragma unroll
    for (int w = 0; w < kIcpThreads / 64; ++w) {
        r.x += red[0][w];
        r.y += red[1][w];
    }
    return r;
}

// The sums of the reduction, one 8-byte word each at every level (thread,
// workgroup partial row, total).
// SumF64 (the plain form): the FP32 products accumulated in FP64.
// SumK32 (the precision form): FP32 with compensation, no FP64 arithmetic.
// Every addition of a sum — a thread adding a product, two partial sums
// meeting in the workgroup's transposed rows, the lane butterflies and the
// partial rows of the workgroups — is Knuth's TwoSum, whose rounding error
// r is exact; the error terms are added in FP32 and the value s + e is
// rounded once, at the end (the cascaded summation of Ogita, Rump and
// Oishi). Kahan's cheaper update, e = y - ((s + y) - s), is not exact when a
// product outweighs the running sum, as happens with few pixels per thread.
struct SumF64 {
    double v;
    SICP_DEV void zero() { v = 0.0; }
    SICP_DEV void add(float p) { v += (double)p; }
    SICP_DEV void merge(const SumF64& o) { v += o.v; }
    SICP_DEV SumF64 shfl_xor(int m) const { return SumF64{__shfl_xor(v, m, 64)}; }
    SICP_DEV double value() const { return v; }
};
struct SumK32 {
    float s, e;  // running sum and its error term
    SICP_DEV void zero() {
        s = 0.f;
        e = 0.f;
    }
    SICP_DEV void add(float p) {
        const float t = s + p;
        const float z = t - s;
        const float r = (s - (t - z)) + (p - z);
        e = e + r;
        s = t;
    }
    SICP_DEV void merge(const SumK32& o) {
        const float t = s + o.s;
        const float z = t - s;
        const float r = (s - (t - z)) + (o.s - z);
        e = (e + o.e) + r;
        s = t;
    }
    SICP_DEV SumK32 shfl_xor(int m) const {
        return SumK32{__shfl_xor(s, m, 64), __shfl_xor(e, m, 64)};
    }
    SICP_DEV double value() const { return (double)(s + e); }
};
static_assert(sizeof(SumF64) == 8 && sizeof(SumK32) == 8, "one 8-byte partial per sum");

// Workgroup reduction of every thread's NS sums into this workgroup's
// partial row (column-major: partials[j * kIcpMaxBlocks + workgroup]), by
// transposition through LDS: every thread stores its sums as one row; thread
// (j, k) then adds rows k, k+8, ... of column j and eight lanes fold with
// three butterflies. Fixed order; much cheaper than folding 29 doubles
// across 64 lanes with cross-lane permutes.
template <int NS, class Acc>
SICP_DEV void block_rows_to_partials(const Acc (&acc)[NS], double* partials, Wg wg) {
    __shared__ Acc tile[kIcpThreads][NS + 1];
#pragma unroll
    for (int j = 0; j < NS; ++j) tile[threadIdx.x][j] = acc[j];
    __syncthreads();
    const int t = threadIdx.x;
    if (t < NS * 8) {
        const int j = t >> 3, k = t & 7;
        Acc s;
        s.zero();
#pragma unroll 8
        for (int r = k; r < kIcpThreads; r += 8) s.merge(tile[r][j]);
        s.merge(s.shfl_xor(1));
        s.merge(s.shfl_xor(2));
        s.merge(s.shfl_xor(4));
        // column-major: sum j of every workgroup is contiguous
        if (k == 0) reinterpret_cast<Acc*>(partials)[j * kIcpMaxBlocks + wg.id] = s;
    }
}

template <class Acc>
SICP_DEV void add_row(Acc acc[kIcpSums], const float row[7]) {
    int q = 0;
#pragma unroll
    for (int a = 0; a < 7; ++a)
#pragma unroll
        for (int b = a; b < 7; ++b) acc[q++].add(row[a] * row[b]);
    acc[28].add(1.0f);
}

// Per-thread accumulation + workgroup reduction into this workgroup's
// partial row. `get_pose(cur, last)` is called by every thread once the first
// pixels' loads are in flight; it returns false when there is nothing to do.
// kRgb: the same pass also runs RGBStep's correspondence test
// (FindCorresp, Reduction.cu:409-461) on every pixel, writing its record and
// the workgroup's int count and sum of squared differences.
template <bool kRgb, class Acc, class PoseFn>
SICP_DEV void icp_accumulate(const View<float4>& vcurr, const View<float4>& vlast,
                             const View<float4>& ncurr, const View<float4>& nlast,
                             const IcpGeom& g, double* partials, const RgbPass& rp, Wg wg,
                             PoseFn get_pose) {
    const int stride = wg.n * kIcpThreads;
    int base = wg.id * kIcpThreads + threadIdx.x;
    // A launch of one pixel per thread runs a one-slot copy of the pass (a
    // uniform choice), so neither copy carries per-slot branches that would
    // split the slots' loads and arithmetic into separate blocks.
    const bool single = stride >= g.n;
    IcpBatch bt;
    if (single)
        icp_load_batch<kRgb, 1>(bt, base, stride, vcurr, ncurr, rp, g);
    else
        icp_load_batch<kRgb, kPixPerThread>(bt, base, stride, vcurr, ncurr, rp, g);
    PoseF cur, last;
    if (!get_pose(cur, last)) {
        phase_mark(4);
        return;
    }
    Acc acc[kIcpSums];
#pragma unroll
    for (int j = 0; j < kIcpSums; ++j) acc[j].zero();
    uint32_t cnt = 0, sq = 0;
    auto pass = [&](auto slots) {
        constexpr int NP = decltype(slots)::value;
        while (base < g.n) {
            float4 vl[NP], nl[NP], rvl[NP];
            IcpProj pr[NP];
            int ru[NP], rv[NP], rli[NP];
            bool rok[NP];
#pragma unroll
            for (int k = 0; k < NP; ++k) {
                pr[k] = icp_project(bt.vc[k], cur, last, g);
                const int u = pr[k].ok ? pr[k].u : 0, v = pr[k].ok ? pr[k].v : 0;
                vl[k] = vlast.at(v, u);
                nl[k] = nlast.at(v, u);
                if (kRgb) {
                    rok[k] = rgb_project(pr[k].vcurr_p, g, ru[k], rv[k]) && bt.key[k] >= 0;
                    const int uu = rok[k] ? ru[k] : 0, vv = rok[k] ? rv[k] : 0;
                    rvl[k] = vlast.at(vv, uu);
                    rli[k] = rp.ilast.at(vv, uu);
                }
            }
#pragma unroll
            for (int k = 0; k < NP; ++k) {
                float row[7];
                if (pr[k].ok && icp_finish(pr[k], bt.nc[k], vl[k], nl[k], cur, last, row))
                    add_row(acc, row);
                if (kRgb) {
                    const int i = base + k * stride;
                    if (i < g.n) {
                        const int2 rc = rok[k] ? rgb_record(bt.key[k], bt.vc[k], rvl[k], rli[k],
                                                            ru[k], rv[k], g.cols)
                                               : make_int2(-1, 0);
                        rp.rec[i] = rc;
                        if (rc.x >= 0) {
                            cnt += 1u;
                            sq += (uint32_t)(rc.y * rc.y);
                        }
                    }
                }
            }
            base += stride * NP;
            if (base < g.n) icp_load_batch<kRgb, NP>(bt, base, stride, vcurr, ncurr, rp, g);
        }
    };
    if (single)
        pass(std::integral_constant<int, 1>{});
    else
        pass(std::integral_constant<int, kPixPerThread>{});
    phase_mark(4);
    if (kRgb) {
        const uint2 r = block_sum_u32x2(cnt, sq);
        if (threadIdx.x == 0) {
            rp.ipart[wg.id] = r.x;
            rp.ipart[kIcpMaxBlocks + wg.id] = r.y;
        }
    }
    block_rows_to_partials(acc, partials, wg);
}

// RGBStep's row pass over the records of the correspondence pass of
// `nb_match` workgroups: sigma = sqrt(sum diff^2 / max(count, 1)) from the
// int sums (RGBStep, Reduction.cu:618-620; every workgroup forms the same
// value), then GetRow of every record into this workgroup's partial row.
// Workgroup 0 also stores the count, the sum and sigma in `rs.totals`.
template <class Acc, class PoseFn>
SICP_DEV void rgb_rows(const View<float4>& vlast, const View<int16_t>& dx,
                       const View<int16_t>& dy, const IcpGeom& g, const RgbScratch& rs,
                       int nb_match, PoseFn get_pose) {
    const Wg wg = grid_wg();
    const int stride = wg.n * kIcpThreads;
    int base = wg.id * kIcpThreads + threadIdx.x;
    // One-slot or four-slot copy of the pass, as in icp_accumulate.
    const bool single = stride >= g.n;
    int2 rc[kPixPerThread];
    auto load_records = [&](auto slots) {
#pragma unroll
        for (int k = 0; k < decltype(slots)::value; ++k) {
            const int i = base + k * stride;
            rc[k] = i < g.n ? rs.rec[i] : make_int2(-1, 0);
        }
    };
    if (single)
        load_records(std::integral_constant<int, 1>{});
    else
        load_records(std::integral_constant<int, kPixPerThread>{});
    uint32_t c = 0, q = 0;
    for (int b = threadIdx.x; b < nb_match; b += kIcpThreads) {
        c += rs.ipart[b];
        q += rs.ipart[kIcpMaxBlocks + b];
    }
    PoseF cur, last;
    if (!get_pose(cur, last)) return;
    const uint2 r = block_sum_u32x2(c, q);
    const int count = (int)r.x, sum_sq = (int)r.y;
    const float sigma = sqrtf((float)sum_sq / (float)(count == 0 ? 1 : count));
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        rs.totals[0] = count;
        rs.totals[1] = sum_sq;
        rs.totals[2] = __float_as_int(sigma);
    }
    Acc acc[kIcpSums];
#pragma unroll
    for (int j = 0; j < kIcpSums; ++j) acc[j].zero();
    auto pass = [&](auto slots) {
        constexpr int NP = decltype(slots)::value;
        while (base < g.n) {
            float4 vl[NP];
            int gx[NP], gy[NP];
#pragma unroll
            for (int k = 0; k < NP; ++k) {
                const int li = rc[k].x >= 0 ? rc[k].x : 0;
                const int ly = li / g.cols, lx = li - ly * g.cols;
                vl[k] = vlast.at(ly, lx);
                const int i = rc[k].x >= 0 ? base + k * stride : 0;
                const int y = i / g.cols, x = i - y * g.cols;
                gx[k] = dx.at(y, x);
                gy[k] = dy.at(y, x);
            }
#pragma unroll
            for (int k = 0; k < NP; ++k) {
                if (rc[k].x >= 0) {
                    float row[7];
                    rgb_row(rc[k].y, vl[k], gx[k], gy[k], sigma, cur, last, g, row);
                    add_row(acc, row);
                }
            }
            base += stride * NP;
            if (base < g.n) load_records(slots);
        }
    };
    if (single)
        pass(std::integral_constant<int, 1>{});
    else
        pass(std::integral_constant<int, kPixPerThread>{});
    block_rows_to_partials(acc, rs.partials, wg);
}

// Sum of the partial rows of `nb` workgroups into totals (LDS), in a fixed
// order: thread (j, k) adds rows k, k+8, ... (below nb) of column j, then
// eight lanes fold. NS sums per row. Above kFoldPairRows rows the thread
// adds the row pairs (2p, 2p + 1), p = k, k+8, ... instead, a pair being one
// 16-byte load (column j starts 16-byte aligned). Measured per launch at
// 640x480 (launch span, single rows against pairs): 75 rows 6.9-8.0 against
// 7.6-8.7 us, 300 rows 11.9 against 11.6 us.
constexpr int kFoldPairRows = 128;
template <int NS = kIcpSums, class Acc = SumF64>
SICP_DEV void icp_totals_issue(const double* partials, int nb, Acc& s) {
    const int t = threadIdx.x;
    s.zero();
    if (t < NS * 8 && nb <= kFoldPairRows) {
        const int j = t >> 3, k = t & 7;
        const Acc* col = reinterpret_cast<const Acc*>(partials) + j * kIcpMaxBlocks;
        // Rows below nb only: sixteen loads in flight per step while sixteen
        // complete rows remain, then eight with predicates. A fixed 64-row
        // unroll with predicated loads cost every launch the full sequence
        // of loads and merges (3.2 against 1.5 us at 19 rows).
        int r0 = 0;
        for (; k + 8 * (r0 + 15) < nb; r0 += 16) {
            Acc v[16];
#pragma unroll
            for (int r = 0; r < 16; ++r) v[r] = col[k + 8 * (r0 + r)];
#pragma unroll
            for (int r = 0; r < 16; ++r) s.merge(v[r]);
        }
        for (; k + 8 * r0 < nb; r0 += 8) {
            Acc v[8];
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                const int bi = k + 8 * (r0 + r);
                if (bi < nb)
                    v[r] = col[bi];
                else
                    v[r].zero();
            }
#pragma unroll
            for (int r = 0; r < 8; ++r) s.merge(v[r]);
        }
    } else if (t < NS * 8) {
        const int j = t >> 3, k = t & 7;
        struct alignas(16) Pair {
            Acc a, b;
        };
        const Pair* col = reinterpret_cast<const Pair*>(partials + j * kIcpMaxBlocks);
        const int full = nb >> 1, pairs = (nb + 1) >> 1;
        // As above by pairs; the second row of a last, odd pair counts as zero.
        int r0 = 0;
        for (; k + 8 * (r0 + 15) < full; r0 += 16) {
            Pair v[16];
#pragma unroll
            for (int r = 0; r < 16; ++r) v[r] = col[k + 8 * (r0 + r)];
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                s.merge(v[r].a);
                s.merge(v[r].b);
            }
        }
        for (; k + 8 * r0 < pairs; r0 += 8) {
            Pair v[8];
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                const int p = k + 8 * (r0 + r);
                if (p < pairs) {
                    v[r] = col[p];
                    if (p >= full) v[r].b.zero();
                } else {
                    v[r].a.zero();
                    v[r].b.zero();
                }
            }
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                s.merge(v[r].a);
                s.merge(v[r].b);
            }
        }
    }
}
template <int NS = kIcpSums, class Acc = SumF64>
SICP_DEV void icp_totals_fold(Acc s, double* totals) {
    const int t = threadIdx.x;
    s.merge(s.shfl_xor(1));
    s.merge(s.shfl_xor(2));
    s.merge(s.shfl_xor(4));
    if (t < NS * 8 && (t & 7) == 0) totals[t >> 3] = s.value();
}
template <int NS = kIcpSums, class Acc = SumF64>
SICP_DEV void icp_totals(const double* partials, int nb, double* totals) {
    Acc s;
    icp_totals_issue<NS, Acc>(partials, nb, s);
    icp_totals_fold<NS, Acc>(s, totals);
    __syncthreads();
}

template <class Acc>
__global__ __launch_bounds__(kIcpThreads) void icp_partials_kernel(
    View<float4> vcurr, View<float4> vlast, View<float4> ncurr, View<float4> nlast, PoseArg pc,
    PoseArg pl, IcpGeom g, double* partials) {
    icp_accumulate<false, Acc>(vcurr, vlast, ncurr, nlast, g, partials, RgbPass{}, grid_wg(),
                          [&](PoseF& cur, PoseF& last) {
                              cur = load_pose(pc);
                              last = load_pose(pl);
                              return true;
                          });
}

template <int NS = kIcpSums, class Acc = SumF64>
__global__ __launch_bounds__(kIcpThreads) void icp_sums_finalize_kernel(const double* partials,
                                                                        int nb,
                                                                        float* sums_out) {
    __shared__ double totals[NS];
    icp_totals<NS, Acc>(partials, nb, totals);
    if (threadIdx.x < NS) sums_out[threadIdx.x] = (float)totals[threadIdx.x];
}

// SO3Reduce::GetRow (Tracking/Reduction.cu:96-131): the last frame's pixel
// bearing rotated into the current camera, its rounded projection
// (__float2int_rn), and the rotation-only photometric row
// J = -(left x p), r = -(I_next(u, v) - I_last(x, y)) with the next frame's
// gradient at (u, v); 10 products and the match flag.
constexpr int kSo3Sums = 11;
__global__ __launch_bounds__(kIcpThreads) void so3_partials_kernel(
    View<uint8_t> inext, View<uint8_t> ilast, View<int16_t> dx, View<int16_t> dy, PoseArg pc,
    PoseArg pl, IcpGeom g, double* partials) {
    const PoseF cur = load_pose(pc), last = load_pose(pl);
    SumF64 acc[kSo3Sums];
#pragma unroll
    for (int j = 0; j < kSo3Sums; ++j) acc[j].zero();
    for (int i = blockIdx.x * kIcpThreads + threadIdx.x; i < g.n; i += gridDim.x * kIcpThreads) {
        const int y = i / g.cols, x = i - y * g.cols;
        const v3 vl = v3{((float)x - g.cx) / g.fx, ((float)y - g.cy) / g.fy, 1.0f};
        const v3 p = mv(cur.Rinv, mv(last.R, vl));
        const int u = f2i(rintf(g.fx * p.x / p.z + g.cx));
        const int v = f2i(rintf(g.fy * p.y / p.z + g.cy));
        const bool found = u >= 5 && v >= 5 && u < g.cols - 5 && v < g.rows - 5 && x >= 5 &&
                           y >= 5 && x < g.cols - 5 && y < g.rows - 5;
        if (!found) continue;
        const float gx = (float)dx.at(v, u) / 9.0f, gy = (float)dy.at(v, u) / 9.0f;
        const float invz = 1.0f / p.z;
        v3 left;
        left.x = gx * g.fx * invz;
        left.y = gy * g.fy * invz;
        left.z = -(p.x * left.x + p.y * left.y) * invz;
        const v3 c = cross(left, p);
        const float row[4] = {-c.x, -c.y, -c.z,
                              -((float)inext.at(v, u) - (float)ilast.at(y, x))};
        int q = 0;
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = a; b < 4; ++b) acc[q++].add(row[a] * row[b]);
        acc[10].add(1.0f);
    }
    block_rows_to_partials(acc, partials, grid_wg());
}

SICP_DEV void copy16(double* d, const double* s) {
#pragma unroll
    for (int i = 0; i < 16; ++i) d[i] = s[i];
}

// Lane j's value of v, in every lane (j uniform, every lane of the wave
// active).
SICP_DEV double lane_bcast(double v, int j) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, j);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), j);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// pose_to_float spread over lanes: lane e writes R[e] and Rinv[e] (e < 9)
// and t[e] (e < 3).
SICP_DEV void pose_to_float_lanes(const double* T, sicp_pose_f* out, int e) {
    if (e < 9) {
        const int r = e / 3, c = e - 3 * r;
        out->R[e] = (float)T[r * 4 + c];
        out->Rinv[e] = (float)T[c * 4 + r];
    }
    if (e < 3) out->t[e] = (float)T[e * 4 + 3];
}

// The update of one iteration of the inner loop of ComputeSE3
// (Tracking.cc:400-514), from the iteration's 29 ICP sums and, with the
// photometric term, its 29 RGB sums (`rgb_totals`, else null), run by the
// 64 lanes of one wavefront. The arithmetic is se3_math.hpp's —
// create_matrix, ldlt_solve6, se3_exp, mat4_mul, rigid_inverse — every
// value formed by the same sequence of operations, so the result is the
// same to the bit; only the work is spread: lane i < 6 owns row i of A and
// of L (column j of the factorisation is one step for every row), lane
// e < 16 owns entry e of the 4x4 products. The state lives in LDS; `tmp`
// is 32 doubles of LDS scratch. Every lane computes the uniform values
// (errors, pivots, the twist, its exponential); lane 0 writes the scalars.
SICP_DEV void gauss_newton_wave(TrackState* st, const double* totals, const double* rgb_totals,
                                int min_count, double* tmp) {
    const int lane = threadIdx.x & 63;
    float rgb_error = 0.f;
    if (rgb_totals) {
        // rgbError and "track rgb failed" (Tracking.cc:425-432), before ICPStep's
        const float q0 = (float)rgb_totals[27], q1 = (float)rgb_totals[28];
        rgb_error = sqrtf(q0) / q1;
        const int rgb_count = (int)q1;
        if (lane == 0) {
            st->rgb_error = rgb_error;
            st->rgb_count = rgb_count;
        }
        if (isnan_(rgb_error) || rgb_count < min_count) {
            if (lane < 16) st->next_pose[lane] = st->last_pose[lane];
            if (lane == 0) {
                st->failed = 1;
                st->done = 1;
            }
            pose_to_float_lanes(st->last_pose, &st->curr_f, lane);
            return;
        }
    }
    int break_count = st->break_count;
    const float last_error = st->last_error;
    const float r0 = (float)totals[27], r1 = (float)totals[28];  // the float download
    const float icp_error = sqrtf(r0) / r1;
    const int icp_count = (int)r1;
    if (lane == 0) st->icp_count = icp_count;
    if (isnan_(icp_error) || icp_count < min_count) {
        // "track icp failed": roll back (Tracking.cc:452-456)
        if (lane < 16) st->next_pose[lane] = st->last_pose[lane];
        if (lane == 0) {
            st->icp_error = icp_error;
            st->failed = 1;
            st->done = 1;
        }
        pose_to_float_lanes(st->last_pose, &st->curr_f, lane);
        return;
    }
    const float curr_error = rgb_error * st->rgb_weight + icp_error;  // Tracking.cc:458
    if (last_error < curr_error) {
        break_count++;
        if (break_count == 1 && lane < 16) st->candidate[lane] = st->next_pose[lane];
        if (lane == 0) st->candidate_icp = icp_error;  // assigned on every increase (Q11)
    } else {
        break_count = 0;
    }
    if (lane == 0) st->break_count = break_count;
    if (break_count > 2) {
        if (lane < 16) st->next_pose[lane] = st->candidate[lane];
        if (lane == 0) {
            st->icp_error = st->candidate_icp;
            st->done = 1;
        }
        pose_to_float_lanes(st->candidate, &st->curr_f, lane);
        return;
    }
    if (lane == 0) {
        st->last_error = curr_error;
        st->icp_error = icp_error;
    }
    // CreateMatrix<6,7>: row i of A and b_i from the float sums; entry (i, j),
    // j >= i, is sum 7i - i(i-1)/2 + (j - i), mirrored below the diagonal.
    // With the photometric term, A = A_rgb * w + A (Tracking.cc:484-487).
    const int i = lane < 6 ? lane : 5;
    double a[6], bi;
    {
        const int di = 7 * i - i * (i - 1) / 2;
        int q[7];
#pragma unroll
        for (int j = 0; j < 6; ++j) q[j] = j >= i ? di + (j - i) : 7 * j - j * (j - 1) / 2 + (i - j);
        q[6] = di + 6 - i;
#pragma unroll
        for (int j = 0; j < 6; ++j) a[j] = (double)(float)totals[q[j]];
        bi = (double)(float)totals[q[6]];
        if (rgb_totals) {
            const double w = (double)st->rgb_weight;
#pragma unroll
            for (int j = 0; j < 6; ++j) a[j] = (double)(float)rgb_totals[q[j]] * w + a[j];
            bi = (double)(float)rgb_totals[q[6]] * w + bi;
        }
    }
    // ldlt_solve6: column j of L and pivot j, every row at once.
    double L[6], D[6], Dinv[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        double s = a[j];
#pragma unroll
        for (int k = 0; k < j; ++k) s -= L[k] * lane_bcast(L[k], j) * D[k];
        const double d = lane_bcast(s, j);
        D[j] = d;
        Dinv[j] = (d != 0.0) ? 1.0 / d : 0.0;
        L[j] = i > j ? s * Dinv[j] : 1.0;
    }
    double zs = bi;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const double zk = lane_bcast(zs, k);
        if (i > k) zs -= L[k] * zk;
    }
    double dinv_i = Dinv[0];
#pragma unroll
    for (int k = 1; k < 6; ++k)
        if (i == k) dinv_i = Dinv[k];
    zs = zs * dinv_i;
    double x[6];
#pragma unroll
    for (int r = 5; r >= 0; --r) {
        double s = lane_bcast(zs, r);
#pragma unroll
        for (int k = r + 1; k < 6; ++k) s -= lane_bcast(L[r], k) * x[k];
        x[r] = s;
    }
    phase_mark(6);
    double dT[16];
    se3_exp(x, dT);
    phase_mark(7);
    // delta = exp(x) * delta; next = last * delta^-1; lane e owns entry e
    if (lane == 0) {
#pragma unroll
        for (int e = 0; e < 16; ++e) tmp[e] = dT[e];
    }
    const int e = lane & 15, r = e >> 2, c = e & 3;
    const double nd = tmp[r * 4 + 0] * st->delta[0 * 4 + c] + tmp[r * 4 + 1] * st->delta[1 * 4 + c] +
                      tmp[r * 4 + 2] * st->delta[2 * 4 + c] + tmp[r * 4 + 3] * st->delta[3 * 4 + c];
    if (lane < 16) st->delta[e] = nd;
    double inv;
    if (r < 3 && c < 3)
        inv = st->delta[c * 4 + r];
    else if (r < 3)
        inv = -(st->delta[0 * 4 + r] * st->delta[3] + st->delta[1 * 4 + r] * st->delta[7] +
                st->delta[2 * 4 + r] * st->delta[11]);
    else
        inv = c == 3 ? 1.0 : 0.0;
    if (lane < 16) tmp[16 + e] = inv;
    const double np = st->last_pose[r * 4 + 0] * tmp[16 + 0 * 4 + c] +
                      st->last_pose[r * 4 + 1] * tmp[16 + 1 * 4 + c] +
                      st->last_pose[r * 4 + 2] * tmp[16 + 2 * 4 + c] +
                      st->last_pose[r * 4 + 3] * tmp[16 + 3 * 4 + c];
    if (lane < 16) st->next_pose[e] = np;
    if (lane == 0) st->applied++;
    pose_to_float_lanes(st->next_pose, &st->curr_f, lane);
}

// ComputeSE3's set-up (Tracking.cc:369-431 with TrackFrame :157-173): the
// iteration state of a new frame, from the last frame's pose.
SICP_DEV void track_begin_state(TrackState* st, const FrameMeta* last, const TrackParams& tp) {
    double lp[16], I[16];
    copy16(lp, last->pose);
    set_identity(I);
    copy16(st->last_pose, lp);
    copy16(st->next_pose, lp);
    copy16(st->pose0, lp);
    copy16(st->candidate, lp);
    copy16(st->delta, I);
    st->icp_error = 0.f;
    st->last_error = 100.f;
    st->candidate_icp = 0.f;
    st->thresh = tp.thresh;
    st->rgb_error = 0.f;
    st->rgb_weight = tp.rgb_weight;
    st->rgb_count = 0;
    st->rgb_on = tp.rgb;
    st->icp_count = 0;
    st->break_count = 0;
    st->applied = 0;
    st->done = 0;
    st->failed = 0;
    st->tracked = 0;
    pose_to_float(lp, &st->curr_f);
    pose_to_float(lp, &st->last_f);
}

// Final motion / error gate of ComputeSE3 (Tracking.cc:505-516).
SICP_DEV void track_end_state(TrackState* st) {
    int ok = 0;
    if (!st->failed) {
        double p0[16], np_[16], inv0[16], p[16], rn, tn;
        copy16(p0, st->pose0);
        copy16(np_, st->next_pose);
        rigid_inverse(p0, inv0);
        mat4_mul(inv0, np_, p);
        euler_sin_norm(p, &rn, &tn);
        ok = (st->icp_error < st->thresh && rn <= 0.1 && tn <= 0.1) ? 1 : 0;
        if (!ok) copy16(st->next_pose, st->last_pose);
    }
    st->tracked = ok;
}

constexpr int kStateWords = (int)(sizeof(TrackState) / sizeof(unsigned long long));
static_assert(sizeof(TrackState) % sizeof(unsigned long long) == 0, "TrackState padding");
static_assert(kStateWords <= kIcpThreads, "TrackState copied one word per thread");

// Brings the iteration state into LDS. kFirst: built from the last frame's
// pose; otherwise the previous iteration's state, advanced by that
// iteration's Gauss-Newton update from its partial rows. Every workgroup of
// a launch computes the same state (same inputs, same order), so no
// workgroup waits for another; workgroup 0 publishes it to `st_out`.
// With `rgb_prev` (the previous iteration's photometric partial rows) the
// update folds those too; `totals` then holds 2 * kIcpSums values.
template <bool kFirst, class Acc>
SICP_DEV void icp_state(TrackState* S, double* totals, const TrackState* st_in,
                        const FrameMeta* last, const TrackParams& tp,
                        const double* partials_prev, const double* rgb_prev, int nb_prev,
                        int min_count_prev, TrackState* st_out, Wg wg) {
    if (kFirst) {
        phase_mark(1);
        if (threadIdx.x == 0) track_begin_state(S, last, tp);
        __syncthreads();
        phase_mark(2);
    } else {
        const unsigned long long* src = reinterpret_cast<const unsigned long long*>(st_in);
        unsigned long long w = threadIdx.x < kStateWords ? src[threadIdx.x] : 0ull;
        Acc s, sr;
        sr.zero();
        icp_totals_issue<kIcpSums, Acc>(partials_prev, nb_prev, s);
        if (rgb_prev) icp_totals_issue<kIcpSums, Acc>(rgb_prev, nb_prev, sr);
        if (threadIdx.x < kStateWords) reinterpret_cast<unsigned long long*>(S)[threadIdx.x] = w;
        icp_totals_fold<kIcpSums, Acc>(s, totals);
        if (rgb_prev) icp_totals_fold<kIcpSums, Acc>(sr, totals + kIcpSums);
        __syncthreads();
        phase_mark(1);
        __shared__ double gn_tmp[32];
        if (threadIdx.x < 64 && !S->done)
            gauss_newton_wave(S, totals, rgb_prev ? totals + kIcpSums : nullptr, min_count_prev,
                              gn_tmp);
        __syncthreads();
        phase_mark(2);
    }
    if (st_out && wg.id == 0 && threadIdx.x < kStateWords)
        reinterpret_cast<unsigned long long*>(st_out)[threadIdx.x] =
            reinterpret_cast<const unsigned long long*>(S)[threadIdx.x];
}

// One launch per Gauss-Newton iteration: the update of the previous
// iteration, then this iteration's pixel pass (ICPReduce at the state's pose).
// kRgb: the photometric term is on; its correspondence pass runs here and
// the previous iteration's photometric rows (`rgb_prev`) join the update.
// Acc: the arithmetic of the sums (SumF64, or SumK32 in the precision form).
template <bool kFirst, bool kRgb, class Acc>
SICP_DEV void icp_iteration_body(const View<float4>& vcurr, const View<float4>& vlast,
                                 const View<float4>& ncurr, const View<float4>& nlast,
                                 const IcpGeom& g, const TrackState* st_in, const FrameMeta* last,
                                 const TrackParams& tp, const double* partials_prev,
                                 const double* rgb_prev, int nb_prev, int min_count_prev,
                                 TrackState* st_out, double* partials, const RgbPass& rp, Wg wg) {
    __shared__ TrackState S;
    __shared__ double totals[2 * kIcpSums];
    phase_mark(0);
    icp_accumulate<kRgb, Acc>(vcurr, vlast, ncurr, nlast, g, partials, rp, wg,
                              [&](PoseF& cur, PoseF& lst) {
                                  icp_state<kFirst, Acc>(&S, totals, st_in, last, tp,
                                                         partials_prev, kRgb ? rgb_prev : nullptr,
                                                         nb_prev, min_count_prev, st_out, wg);
                                  phase_mark(3);
                                  if (S.done) return false;
                                  cur = to_pose(S.curr_f);
                                  lst = to_pose(S.last_f);
                                  return true;
                              });
    phase_mark(5);
#ifdef SICP_ICP_PHASES
    if (threadIdx.x == 0 && tp.launch >= 0 && tp.launch < kPhaseLaunches)
        for (int i = 0; i < kPhaseStamps; ++i)
            g_icp_phase[tp.launch][wg.id][i] = (unsigned long long)phase_lds()[i];
#endif
}

template <bool kFirst, bool kRgb, class Acc>
__global__ __launch_bounds__(kIcpThreads) void icp_iteration_kernel(
    View<float4> vcurr, View<float4> vlast, View<float4> ncurr, View<float4> nlast, IcpGeom g,
    const TrackState* st_in, const FrameMeta* last, TrackParams tp, const double* partials_prev,
    const double* rgb_prev, int nb_prev, int min_count_prev, TrackState* st_out,
    double* partials, RgbPass rp) {
    icp_iteration_body<kFirst, kRgb, Acc>(vcurr, vlast, ncurr, nlast, g, st_in, last, tp,
                                          partials_prev, rgb_prev, nb_prev, min_count_prev,
                                          st_out, partials, rp, grid_wg());
}

// One Gauss-Newton iteration of several sequences in one launch: sequence q
// owns workgroups wg0[q] .. wg0[q + 1] - 1 and runs exactly the arithmetic of
// its own icp_iteration_kernel launch on them (the point-to-plane term).
template <bool kFirst, class Acc>
__global__ __launch_bounds__(kIcpThreads) void icp_iteration_batch_kernel(IcpBatchArgs a) {
    int q = 0;
    while (q + 1 < a.n && (int)blockIdx.x >= a.wg0[q + 1]) ++q;
    const IcpSeqArgs& e = a.seq[q];
    const Wg wg{(int)blockIdx.x - a.wg0[q], a.wg0[q + 1] - a.wg0[q]};
    icp_iteration_body<kFirst, false, Acc>(e.vcurr, e.vlast, e.ncurr, e.nlast, e.g, e.st_in,
                                           e.last, e.tp, e.partials_prev, nullptr, e.nb_prev,
                                           e.min_count_prev, e.st_out, e.partials, RgbPass{}, wg);
}

// RGBStep's row pass. The pose is the iteration state's (`st`, published by
// the iteration's pixel pass) or, without a state, `pc` / `pl`.
template <class Acc>
__global__ __launch_bounds__(kIcpThreads) void rgb_rows_kernel(
    View<float4> vlast, View<int16_t> dx, View<int16_t> dy, IcpGeom g, RgbScratch rs,
    int nb_match, const TrackState* st, PoseArg pc, PoseArg pl) {
    rgb_rows<Acc>(vlast, dx, dy, g, rs, nb_match, [&](PoseF& cur, PoseF& lst) {
        if (st) {
            if (st->done) return false;
            cur = to_pose(st->curr_f);
            lst = to_pose(st->last_f);
        } else {
            cur = load_pose(pc);
            lst = load_pose(pl);
        }
        return true;
    });
}

// RGBStep's correspondence pass on its own (the pixel pass of the tracking
// loop runs the same test fused with ICP's).
__global__ __launch_bounds__(kIcpThreads) void rgb_match_kernel(View<float4> vcurr,
                                                                View<float4> vlast, IcpGeom g,
                                                                PoseArg pc, PoseArg pl,
                                                                RgbPass rp) {
    const PoseF cur = load_pose(pc), last = load_pose(pl);
    uint32_t cnt = 0, sq = 0;
    for (int i = blockIdx.x * kIcpThreads + threadIdx.x; i < g.n; i += gridDim.x * kIcpThreads) {
        const int y = i / g.cols, x = i - y * g.cols;
        const int key = rp.key.at(y, x);
        const float4 vc = vcurr.at(y, x);
        int2 rc = make_int2(-1, 0);
        int u, v;
        const v3 p = mv(last.Rinv, sub(add(mv(cur.R, xyz(vc)), cur.t), last.t));
        if (rgb_project(p, g, u, v) && key >= 0)
            rc = rgb_record(key, vc, vlast.at(v, u), rp.ilast.at(v, u), u, v, g.cols);
        rp.rec[i] = rc;
        if (rc.x >= 0) {
            cnt += 1u;
            sq += (uint32_t)(rc.y * rc.y);
        }
    }
    const uint2 r = block_sum_u32x2(cnt, sq);
    if (threadIdx.x == 0) {
        rp.ipart[blockIdx.x] = r.x;
        rp.ipart[kIcpMaxBlocks + blockIdx.x] = r.y;
    }
}

// The finishing step of a frame's tracking on one workgroup range (the body
// of track_finish_kernel below).
template <bool kFirst, class Acc>
SICP_DEV void track_finish_body(const TrackState* st_in, const FrameMeta* last,
                                const TrackParams& tp, const double* partials_prev,
                                const double* rgb_prev, int nb_prev, int min_count_prev,
                                TrackState* st_final, FrameMeta* next, const View<float4>& vlast,
                                const View<float4>& nlast, const View<float4>& vnext,
                                const View<float4>& nnext, const FrameImages& images, Wg wg) {
    __shared__ TrackState S;
    __shared__ double totals[2 * kIcpSums];
    icp_state<kFirst, Acc>(&S, totals, st_in, last, tp, partials_prev, rgb_prev, nb_prev,
                           min_count_prev, nullptr, wg);
    if (threadIdx.x == 0) track_end_state(&S);
    __syncthreads();
    if (wg.id == 0) {
        if (threadIdx.x < kStateWords)
            reinterpret_cast<unsigned long long*>(st_final)[threadIdx.x] =
                reinterpret_cast<const unsigned long long*>(&S)[threadIdx.x];
        if (threadIdx.x < 16) next->pose[threadIdx.x] = S.next_pose[threadIdx.x];
        if (threadIdx.x == 0) pose_to_float(S.next_pose, &next->pose_f);
    }
    if (S.tracked) return;
    const int n = vnext.cols * vnext.rows;
    for (int i = wg.id * kIcpThreads + threadIdx.x; i < n; i += wg.n * kIcpThreads) {
        int y = i / vnext.cols, x = i - y * vnext.cols;
        vnext.at(y, x) = vlast.at(y, x);
        nnext.at(y, x) = nlast.at(y, x);
    }
    // The photometric term reads the last frame's intensity pyramid.
    for (int l = 0; l < 3; ++l) {
        const View<uint8_t>& src = images.last[l];
        const View<uint8_t>& dst = images.next[l];
        if (!dst.ptr) continue;
        const int m = dst.cols * dst.rows;
        for (int i = wg.id * kIcpThreads + threadIdx.x; i < m; i += wg.n * kIcpThreads) {
            int y = i / dst.cols, x = i - y * dst.cols;
            dst.at(y, x) = src.at(y, x);
        }
    }
}

// The end of ComputeSE3 and TrackFrame: the last iteration's update, the
// motion / error gate, the frame pose; when the frame is not tracked, the
// last frame's level-0 model maps are copied into the new frame so the frame
// swap keeps the last model (Tracking.cc:98-117 keeps lastFrame on failure).
template <bool kFirst, class Acc>
__global__ __launch_bounds__(kIcpThreads) void track_finish_kernel(
    const TrackState* st_in, const FrameMeta* last, TrackParams tp, const double* partials_prev,
    const double* rgb_prev, int nb_prev, int min_count_prev, TrackState* st_final,
    FrameMeta* next, View<float4> vlast, View<float4> nlast, View<float4> vnext,
    View<float4> nnext, FrameImages images) {
    track_finish_body<kFirst, Acc>(st_in, last, tp, partials_prev, rgb_prev, nb_prev,
                                   min_count_prev, st_final, next, vlast, nlast, vnext, nnext,
                                   images, grid_wg());
}

// The tracking finish of several sequences in one launch (point-to-plane
// term), sequence q on workgroups wg0[q] .. wg0[q + 1] - 1.
template <class Acc>
__global__ __launch_bounds__(kIcpThreads) void track_finish_batch_kernel(FinishBatchArgs a) {
    int q = 0;
    while (q + 1 < a.n && (int)blockIdx.x >= a.wg0[q + 1]) ++q;
    const FinishSeqArgs& e = a.seq[q];
    const Wg wg{(int)blockIdx.x - a.wg0[q], a.wg0[q + 1] - a.wg0[q]};
    track_finish_body<false, Acc>(e.st_in, e.last, e.tp, e.partials_prev, nullptr, e.nb_prev,
                                  e.min_count_prev, e.st_final, e.next, e.vlast, e.nlast, e.vnext,
                                  e.nnext, FrameImages{}, wg);
}

__global__ void track_init_kernel(FrameMeta* next, FrameMeta* last, TrackState* st) {
    double I[16];
    set_identity(I);
    copy16(next->pose, I);
    copy16(last->pose, I);
    pose_to_float(I, &next->pose_f);
    pose_to_float(I, &last->pose_f);
    st->tracked = 1;
    st->failed = 0;
    st->done = 1;
    st->applied = 0;
    st->icp_count = 0;
    st->icp_error = 0.f;
    st->rgb_count = 0;
    st->rgb_error = 0.f;
}

// Workgroups of the tracking finish.
constexpr int kFinishBlocks = 64;

// Four pixels per thread (one up to kIcpSpreadPixels when `spread`), at
// most kIcpMaxBlocks workgroups. The pixel pass of a lane is a chain of about 360
// arithmetic instructions per pixel after one dependent gather; at 640x480's
// levels 1 and 2 (76.8 k and 19.2 k pixels) four pixels per thread keep 75
// and 19 of the 256 compute units busy, and one per thread cuts the launch by
// more than the fold of four times the partial rows adds to the next one.
// Systems run together fill the GPU already and leave it off: there the
// extra workgroups only queue (batched eight sequences: 3390 against 3568
// frames/s).
inline int icp_blocks(int n, bool spread) {
    const int per_thread = spread && n <= kIcpSpreadPixels ? 1 : kPixPerThread;
    int b = (int)div_up(n, kIcpThreads * per_thread);
    return b < 1 ? 1 : (b > kIcpMaxBlocks ? kIcpMaxBlocks : b);
}

}  // namespace

int launch_icp_sums(View<float4> vcurr, View<float4> vlast, View<float4> ncurr,
                    View<float4> nlast, PoseArg curr, PoseArg last, sicp_intrinsics K,
                    IcpScratch scratch, float* sums_out, bool k32, hipStream_t s) {
    IcpGeom g{vcurr.cols, vcurr.rows, vcurr.cols * vcurr.rows, K.fx, K.fy, K.cx, K.cy};
    const int nb = icp_blocks(g.n, true);
    if (k32) {
        icp_partials_kernel<SumK32><<<nb, kIcpThreads, 0, s>>>(vcurr, vlast, ncurr, nlast, curr,
                                                               last, g, scratch.partials);
        SICP_LAUNCH_CHECK("icp_partials_kernel");
        icp_sums_finalize_kernel<kIcpSums, SumK32><<<1, kIcpThreads, 0, s>>>(scratch.partials,
                                                                              nb, sums_out);
    } else {
        icp_partials_kernel<SumF64><<<nb, kIcpThreads, 0, s>>>(vcurr, vlast, ncurr, nlast, curr,
                                                               last, g, scratch.partials);
        SICP_LAUNCH_CHECK("icp_partials_kernel");
        icp_sums_finalize_kernel<><<<1, kIcpThreads, 0, s>>>(scratch.partials, nb, sums_out);
    }
    SICP_LAUNCH_CHECK("icp_sums_finalize_kernel");
    return SICP_OK;
}

namespace {
IcpGeom geom_of(const IcpLevel& l) {
    return IcpGeom{l.vcurr.cols, l.vcurr.rows, l.vcurr.cols * l.vcurr.rows,
                   l.K.fx,       l.K.fy,       l.K.cx,
                   l.K.cy};
}
}  // namespace

template <class Acc>
static void icp_iteration_launch(const IcpLevel& cur, const IcpLevel* prev,
                                 const FrameMeta* last, TrackParams tp, const TrackState* st_in,
                                 TrackState* st_out, const double* partials_prev,
                                 double* partials, const RgbScratch* rgb, hipStream_t s) {
    const IcpGeom g = geom_of(cur);
    const int nb = icp_blocks(g.n, tp.spread);
    const int nb_prev = prev ? icp_blocks(prev->vcurr.cols * prev->vcurr.rows, tp.spread) : 0;
    const int min_prev = prev ? prev->min_count : 0;
    RgbPass rp{};
    const double* rgb_prev = nullptr;
    if (rgb) {
        rp = RgbPass{cur.rgb.key, cur.rgb.ilast, rgb->rec, rgb->ipart};
        rgb_prev = rgb->partials;
    }
    if (prev && rgb)
        icp_iteration_kernel<false, true, Acc><<<nb, kIcpThreads, 0, s>>>(
            cur.vcurr, cur.vlast, cur.ncurr, cur.nlast, g, st_in, last, tp, partials_prev,
            rgb_prev, nb_prev, min_prev, st_out, partials, rp);
    else if (prev)
        icp_iteration_kernel<false, false, Acc><<<nb, kIcpThreads, 0, s>>>(
            cur.vcurr, cur.vlast, cur.ncurr, cur.nlast, g, st_in, last, tp, partials_prev,
            nullptr, nb_prev, min_prev, st_out, partials, rp);
    else if (rgb)
        icp_iteration_kernel<true, true, Acc><<<nb, kIcpThreads, 0, s>>>(
            cur.vcurr, cur.vlast, cur.ncurr, cur.nlast, g, nullptr, last, tp, nullptr, nullptr,
            0, 0, st_out, partials, rp);
    else
        icp_iteration_kernel<true, false, Acc><<<nb, kIcpThreads, 0, s>>>(
            cur.vcurr, cur.vlast, cur.ncurr, cur.nlast, g, nullptr, last, tp, nullptr, nullptr,
            0, 0, st_out, partials, rp);
}

int launch_icp_iteration(const IcpLevel& cur, const IcpLevel* prev, const FrameMeta* last,
                         TrackParams tp, const TrackState* st_in, TrackState* st_out,
                         const double* partials_prev, double* partials, const RgbScratch* rgb,
                         hipStream_t s) {
    if (tp.sums_k32)
        icp_iteration_launch<SumK32>(cur, prev, last, tp, st_in, st_out, partials_prev, partials,
                                     rgb, s);
    else
        icp_iteration_launch<SumF64>(cur, prev, last, tp, st_in, st_out, partials_prev, partials,
                                     rgb, s);
    SICP_LAUNCH_CHECK("icp_iteration_kernel");
    return SICP_OK;
}

int launch_icp_iteration_batch(const IcpBatchEntry* e, int n, hipStream_t s) {
    if (n < 1 || n > kMaxBatch) return invalid("icp batch: 1 to 8 sequences");
    IcpBatchArgs a{};
    a.n = n;
    int wg = 0;
    const bool first = e[0].prev == nullptr;
    const int k32 = e[0].tp.sums_k32;
    for (int q = 0; q < n; ++q) {
        if ((e[q].prev == nullptr) != first || e[q].tp.sums_k32 != k32 || e[q].tp.rgb)
            return invalid("icp batch: sequences at different iterations or forms");
        const IcpGeom g = geom_of(e[q].cur);
        a.wg0[q] = wg;
        wg += icp_blocks(g.n, e[q].tp.spread);
        const IcpLevel* p = e[q].prev;
        a.seq[q] = IcpSeqArgs{e[q].cur.vcurr,
                              e[q].cur.vlast,
                              e[q].cur.ncurr,
                              e[q].cur.nlast,
                              g,
                              e[q].tp,
                              e[q].st_in,
                              e[q].last,
                              e[q].st_out,
                              e[q].partials_prev,
                              e[q].partials,
                              p ? icp_blocks(p->vcurr.cols * p->vcurr.rows, e[q].tp.spread) : 0,
                              p ? p->min_count : 0};
    }
    a.wg0[n] = wg;
    if (first && k32)
        icp_iteration_batch_kernel<true, SumK32><<<wg, kIcpThreads, 0, s>>>(a);
    else if (first)
        icp_iteration_batch_kernel<true, SumF64><<<wg, kIcpThreads, 0, s>>>(a);
    else if (k32)
        icp_iteration_batch_kernel<false, SumK32><<<wg, kIcpThreads, 0, s>>>(a);
    else
        icp_iteration_batch_kernel<false, SumF64><<<wg, kIcpThreads, 0, s>>>(a);
    SICP_LAUNCH_CHECK("icp_iteration_batch_kernel");
    return SICP_OK;
}

int launch_track_finish_batch(const FinishBatchEntry* e, int n, hipStream_t s) {
    if (n < 1 || n > kMaxBatch) return invalid("finish batch: 1 to 8 sequences");
    FinishBatchArgs a{};
    a.n = n;
    int wg = 0;
    const int k32 = e[0].tp.sums_k32;
    for (int q = 0; q < n; ++q) {
        if (!e[q].prev || e[q].tp.sums_k32 != k32 || e[q].tp.rgb)
            return invalid("finish batch: sequences in different forms");
        const int npx = e[q].vnext.cols * e[q].vnext.rows;
        a.wg0[q] = wg;
        wg += std::min((int)div_up(npx, kIcpThreads * 4), kFinishBlocks);
        const IcpLevel* p = e[q].prev;
        a.seq[q] = FinishSeqArgs{e[q].tp,       e[q].st_in,  e[q].last,  e[q].partials_prev,
                                 e[q].st_final, e[q].next,   e[q].vlast, e[q].nlast,
                                 e[q].vnext,    e[q].nnext,
                                 icp_blocks(p->vcurr.cols * p->vcurr.rows, e[q].tp.spread),
                                 p->min_count};
    }
    a.wg0[n] = wg;
    if (k32)
        track_finish_batch_kernel<SumK32><<<wg, kIcpThreads, 0, s>>>(a);
    else
        track_finish_batch_kernel<SumF64><<<wg, kIcpThreads, 0, s>>>(a);
    SICP_LAUNCH_CHECK("track_finish_batch_kernel");
    return SICP_OK;
}

int launch_rgb_rows(const IcpLevel& cur, const TrackState* st, RgbScratch rgb, bool k32,
                    bool spread, hipStream_t s) {
    const IcpGeom g = geom_of(cur);
    const int nb = icp_blocks(g.n, spread);
    if (k32)
        rgb_rows_kernel<SumK32><<<nb, kIcpThreads, 0, s>>>(cur.vlast, cur.rgb.dx, cur.rgb.dy, g,
                                                           rgb, nb, st, PoseArg{}, PoseArg{});
    else
        rgb_rows_kernel<SumF64><<<nb, kIcpThreads, 0, s>>>(cur.vlast, cur.rgb.dx, cur.rgb.dy, g,
                                                           rgb, nb, st, PoseArg{}, PoseArg{});
    SICP_LAUNCH_CHECK("rgb_rows_kernel");
    return SICP_OK;
}

int launch_track_finish(const IcpLevel* prev, const FrameMeta* last, TrackParams tp,
                        const TrackState* st_in, const double* partials_prev,
                        const RgbScratch* rgb, TrackState* st_final, FrameMeta* next,
                        View<float4> vlast, View<float4> nlast, View<float4> vnext,
                        View<float4> nnext, const FrameImages* images, hipStream_t s) {
    // Every workgroup redoes the final update (reading all partial rows), so
    // the grid is kept small; the copy of a lost frame's maps is grid-stride.
    const int n = vnext.cols * vnext.rows;
    const int blocks = std::min((int)div_up(n, kIcpThreads * 4), kFinishBlocks);
    FrameImages im{};
    if (images) im = *images;
    if (prev) {
        const int nb_prev = icp_blocks(prev->vcurr.cols * prev->vcurr.rows, tp.spread);
        if (tp.sums_k32)
            track_finish_kernel<false, SumK32><<<blocks, kIcpThreads, 0, s>>>(
                st_in, last, tp, partials_prev, rgb ? rgb->partials : nullptr, nb_prev,
                prev->min_count, st_final, next, vlast, nlast, vnext, nnext, im);
        else
            track_finish_kernel<false, SumF64><<<blocks, kIcpThreads, 0, s>>>(
                st_in, last, tp, partials_prev, rgb ? rgb->partials : nullptr, nb_prev,
                prev->min_count, st_final, next, vlast, nlast, vnext, nnext, im);
    } else {
        track_finish_kernel<true, SumF64><<<blocks, kIcpThreads, 0, s>>>(
            nullptr, last, tp, nullptr, nullptr, 0, 0, st_final, next, vlast, nlast, vnext,
            nnext, im);
    }
    SICP_LAUNCH_CHECK("track_finish_kernel");
    return SICP_OK;
}

int launch_rgb_sums(View<uint8_t> inext, View<uint8_t> ilast, View<float4> vnext,
                    View<float4> vlast, View<int16_t> dx, View<int16_t> dy, PoseArg curr,
                    PoseArg last, sicp_intrinsics K, int16_t* key, RgbScratch rgb,
                    float* sums_out, hipStream_t s) {
    const int cols = inext.cols, rows = inext.rows;
    View<int16_t> kv{(char*)key, (size_t)cols * sizeof(int16_t), cols, rows};
    int rc = launch_rgb_keys(inext, dx, dy, vnext, kv, s);
    if (rc) return rc;
    IcpLevel lv{};
    lv.vcurr = vnext;
    lv.vlast = vlast;
    lv.K = K;
    lv.rgb = RgbLevel{kv, dx, dy, ilast};
    const IcpGeom g = geom_of(lv);
    const int nb = icp_blocks(g.n, true);
    rgb_match_kernel<<<nb, kIcpThreads, 0, s>>>(vnext, vlast, g, curr, last,
                                                 RgbPass{kv, ilast, rgb.rec, rgb.ipart});
    SICP_LAUNCH_CHECK("rgb_match_kernel");
    rgb_rows_kernel<SumF64><<<nb, kIcpThreads, 0, s>>>(vlast, dx, dy, g, rgb, nb, nullptr, curr,
                                                       last);
    SICP_LAUNCH_CHECK("rgb_rows_kernel");
    icp_sums_finalize_kernel<><<<1, kIcpThreads, 0, s>>>(rgb.partials, nb, sums_out);
    SICP_LAUNCH_CHECK("icp_sums_finalize_kernel");
    return SICP_OK;
}

int launch_so3_sums(View<uint8_t> inext, View<uint8_t> ilast, View<int16_t> dx,
                    View<int16_t> dy, PoseArg curr, PoseArg last, sicp_intrinsics K,
                    IcpScratch scratch, float* sums_out, hipStream_t s) {
    const IcpGeom g{inext.cols, inext.rows, inext.cols * inext.rows, K.fx, K.fy, K.cx, K.cy};
    const int nb = icp_blocks(g.n, true);
    so3_partials_kernel<<<nb, kIcpThreads, 0, s>>>(inext, ilast, dx, dy, curr, last, g,
                                                     scratch.partials);
    SICP_LAUNCH_CHECK("so3_partials_kernel");
    icp_sums_finalize_kernel<kSo3Sums><<<1, kIcpThreads, 0, s>>>(scratch.partials, nb, sums_out);
    SICP_LAUNCH_CHECK("icp_sums_finalize_kernel");
    return SICP_OK;
}

int launch_track_init(FrameMeta* next, FrameMeta* last, TrackState* st, hipStream_t s) {
    track_init_kernel<<<1, 1, 0, s>>>(next, last, st);
    SICP_LAUNCH_CHECK("track_init_kernel");
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

#ifdef SICP_ICP_PHASES
// Profiling build: copies the stamps (kPhaseLaunches x kIcpMaxBlocks x 8
// words, 100 MHz) to `host` after synchronising the device.
extern "C" int sicp_profile_icp_phases(unsigned long long* host) {
    if (!host) return sicp::invalid("profile_icp_phases: null argument");
    SICP_HIP(hipDeviceSynchronize());
    SICP_HIP(hipMemcpyFromSymbol(host, HIP_SYMBOL(sicp::g_icp_phase), sizeof(sicp::g_icp_phase),
                                 0, hipMemcpyDeviceToHost));
    return SICP_OK;
}
#endif
