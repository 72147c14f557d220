// oracle_icp.cpp — CPU restatement of the point-to-plane ICP reduction
// (Tracking/Reduction.cu:6-16,213-385) and of the Gauss-Newton pose update
// of Tracker::ComputeSE3 (Tracking/Tracking.cc:369-517).
// TEST INFRASTRUCTURE ONLY (see oracle.h).
#include "oracle_impl.hpp"

namespace oracle {

// ICPReduce::searchPoint (Reduction.cu:228-258). The projection keeps the
// reference's FP64 promotions: invz = 1.0 / z in double stored to float, and
// the +0.5 rounding offset added in double before the truncating cast (Q9).
template <class R>
static inline bool search_point(const R* vcurr, const R* vlast, const R* ncurr, const R* nlast,
                                int cols, int rows, int x, int y, const Pose<R>& curr,
                                const Pose<R>& last, Intr<R> K, V3<R>& vcurr_g, V3<R>& vlast_g,
                                V3<R>& nlast_g) {
    const R* vc = vcurr + ((size_t)y * cols + x) * 4;
    V3<R> vcurr_c{vc[0], vc[1], vc[2]};
    if (isnan_(vcurr_c.x) || (double)vcurr_c.z < 1e-3) return false;
    vcurr_g = add(mv(curr.R_, vcurr_c), curr.t);
    V3<R> vcurr_p = mv(last.Rinv, sub(vcurr_g, last.t));
    R invz = (R)(1.0 / (double)vcurr_p.z);
    int u = f2i((double)(vcurr_p.x * invz * K.fx + K.cx) + 0.5);
    int v = f2i((double)(vcurr_p.y * invz * K.fy + K.cy) + 0.5);
    if (u < 0 || v < 0 || u >= cols || v >= rows) return false;
    const R* vl = vlast + ((size_t)v * cols + u) * 4;
    V3<R> vlast_c{vl[0], vl[1], vl[2]};
    vlast_g = add(mv(last.R_, vlast_c), last.t);
    const R* nc = ncurr + ((size_t)y * cols + x) * 4;
    V3<R> ncurr_c{nc[0], nc[1], nc[2]};
    V3<R> ncurr_g = mv(curr.R_, ncurr_c);
    const R* nl = nlast + ((size_t)v * cols + u) * 4;
    V3<R> nlast_c{nl[0], nl[1], nl[2]};
    nlast_g = mv(last.R_, nlast_c);
    R dist = norm(sub(vlast_g, vcurr_g));
    R sine = norm(cross(ncurr_g, nlast_g));
    const R angle_thresh = (R)0.6, dist_thresh = (R)0.1;  // Reduction.cu:361-362
    return sine < angle_thresh && dist <= dist_thresh && !isnan_(ncurr_c.x) &&
           !isnan_(nlast_c.x);
}

// ICPReduce::getRow (Reduction.cu:260-290): J = [-n, n x v_last],
// r = -n . (v_curr - v_last), all in the last camera's frame; then the 28
// products row_i * row_j (i <= j) and the valid flag.
template <class R>
static inline void get_row(const R* vcurr, const R* vlast, const R* ncurr, const R* nlast,
                           int cols, int rows, int i, const Pose<R>& curr, const Pose<R>& last,
                           Intr<R> K, R* out29) {
    int y = i / cols;
    int x = i - y * cols;
    V3<R> vc, vl, nl;
    bool found = search_point(vcurr, vlast, ncurr, nlast, cols, rows, x, y, curr, last, K, vc,
                              vl, nl);
    R row[7] = {0, 0, 0, 0, 0, 0, 0};
    if (found) {
        nl = mv(last.Rinv, nl);
        vc = mv(last.Rinv, sub(vc, last.t));
        vl = mv(last.Rinv, sub(vl, last.t));
        V3<R> mn = neg(nl);
        row[0] = mn.x;
        row[1] = mn.y;
        row[2] = mn.z;
        V3<R> c = cross(nl, vl);
        row[3] = c.x;
        row[4] = c.y;
        row[5] = c.z;
        row[6] = dot(mn, sub(vc, vl));
    }
    int count = 0;
    for (int a = 0; a < 7; ++a)
        for (int b = a; b < 7; ++b) out29[count++] = row[a] * row[b];
    out29[count] = (R)(found ? 1 : 0);
}

template <class R>
void icp_sums(const R* vcurr, const R* vlast, const R* ncurr, const R* nlast, int cols, int rows,
              const Pose<R>& curr, const Pose<R>& last, Intr<R> K, double* sums,
              float* rows_out) {
    const int N = cols * rows;
    for (int j = 0; j < 29; ++j) sums[j] = 0.0;
#pragma omp parallel
    {
        double loc[29] = {0};
#pragma omp for schedule(static)
        for (int i = 0; i < N; ++i) {
            R v[29];
            get_row(vcurr, vlast, ncurr, nlast, cols, rows, i, curr, last, K, v);
            for (int j = 0; j < 29; ++j) loc[j] += (double)v[j];
            if (rows_out)
                for (int j = 0; j < 29; ++j) rows_out[(size_t)i * 29 + j] = (float)v[j];
        }
#pragma omp critical
        for (int j = 0; j < 29; ++j) sums[j] += loc[j];
    }
}

template void icp_sums<float>(const float*, const float*, const float*, const float*, int, int,
                              const Pose<float>&, const Pose<float>&, Intr<float>, double*,
                              float*);
template void icp_sums<double>(const double*, const double*, const double*, const double*, int,
                               int, const Pose<double>&, const Pose<double>&, Intr<double>,
                               double*, float*);

// ---- photometric term: RGBReduction / RGBStep (Reduction.cu:387-642) ------
// One pixel's correspondence of RGBReduction::FindCorresp (:413-463): inside
// a 5-pixel border, a 4x4 window of non-zero intensity, a non-zero gradient,
// a valid current vertex projected into the last frame (float argument,
// +0.5 added in double, rounded down as __float2int_rd), a valid last vertex
// within 0.1 of the current one (float4 distance, compared unmoved) and a
// non-zero last intensity. Returns false or the last pixel and the
// intensity difference.
template <class R>
static inline bool rgb_corresp(const uint8_t* nimg, const uint8_t* limg, const R* nvmap,
                               const R* lvmap, const int16_t* dx, const int16_t* dy, int cols,
                               int rows, int x, int y, const Pose<R>& curr, const Pose<R>& last,
                               Intr<R> K, int& u, int& v, int& diff) {
    if (!(x >= 5 && x < cols - 5 && y >= 5 && y < rows - 5)) return false;
    for (int a = std::max(y - 2, 0); a < std::min(y + 2, rows); ++a)
        for (int b = std::max(x - 2, 0); b < std::min(x + 2, cols); ++b)
            if (!(nimg[(size_t)a * cols + b] > 0)) return false;
    const int gx = dx[(size_t)y * cols + x], gy = dy[(size_t)y * cols + x];
    if (!(std::sqrt((float)(gx * gx + gy * gy)) > 0)) return false;
    const R* vc = nvmap + ((size_t)y * cols + x) * 4;
    if (isnan_(vc[0])) return false;
    V3<R> vcurr_g = add(mv(curr.R_, V3<R>{vc[0], vc[1], vc[2]}), curr.t);
    V3<R> p = mv(last.Rinv, sub(vcurr_g, last.t));
    R tu = (R)((double)(K.fx * p.x / p.z + K.cx) + 0.5);
    R tv = (R)((double)(K.fy * p.y / p.z + K.cy) + 0.5);
    u = f2i(std::floor(tu));
    v = f2i(std::floor(tv));
    if (!(u >= 0 && v >= 0 && u < cols && v < rows)) return false;
    const R* vl = lvmap + ((size_t)v * cols + u) * 4;
    if (isnan_(vl[0])) return false;
    R d0 = vl[0] - vc[0], d1 = vl[1] - vc[1], d2 = vl[2] - vc[2], d3 = vl[3] - vc[3];
    R dist = std::sqrt(d0 * d0 + d1 * d1 + d2 * d2 + d3 * d3);
    if (!(dist < (R)0.1 && limg[(size_t)v * cols + u] != 0)) return false;
    diff = (int)nimg[(size_t)y * cols + x] - (int)limg[(size_t)v * cols + u];
    return true;
}

// RGBStep: the correspondence pass with its int sums (count, sum of squared
// differences; 32-bit, wrapping as the reference's int accumulation), sigma
// = sqrt(sum / max(count, 1)) in float, then RGBReduction::GetRow
// (:482-521) with the robust weight 1 / (sigma + |diff|) and the 29 sums.
template <class R>
void rgb_sums(const uint8_t* nimg, const uint8_t* limg, const R* nvmap, const R* lvmap,
              const int16_t* dx, const int16_t* dy, int cols, int rows, const Pose<R>& curr,
              const Pose<R>& last, Intr<R> K, double* sums, int* res2, float* sigma_out,
              float* rows_out) {
    const int N = cols * rows;
    uint32_t cnt = 0, sq = 0;
#pragma omp parallel for schedule(static) reduction(+ : cnt, sq)
    for (int i = 0; i < N; ++i) {
        int y = i / cols, x = i - y * cols, u, v, diff;
        if (rgb_corresp(nimg, limg, nvmap, lvmap, dx, dy, cols, rows, x, y, curr, last, K, u, v,
                        diff)) {
            cnt += 1u;
            sq += (uint32_t)(diff * diff);
        }
    }
    res2[0] = (int)cnt;
    res2[1] = (int)sq;
    const float sigma = std::sqrt((float)res2[1] / (float)(res2[0] == 0 ? 1 : res2[0]));
    if (sigma_out) *sigma_out = sigma;
    for (int j = 0; j < 29; ++j) sums[j] = 0.0;
#pragma omp parallel
    {
        double loc[29] = {0};
#pragma omp for schedule(static)
        for (int i = 0; i < N; ++i) {
            int y = i / cols, x = i - y * cols, u, v, diff;
            R row[7] = {0, 0, 0, 0, 0, 0, 0};
            const bool ok = rgb_corresp(nimg, limg, nvmap, lvmap, dx, dy, cols, rows, x, y, curr,
                                        last, K, u, v, diff);
            if (ok) {
                R w = (R)sigma + (R)std::abs(diff);
                w = (double)w < 1e-3 ? (R)1 : (R)1 / w;
                if ((double)sigma < 1e-6) w = (R)1;
                const R* vl = lvmap + ((size_t)v * cols + u) * 4;
                V3<R> q = add(mv(last.R_, V3<R>{vl[0], vl[1], vl[2]}), last.t);
                V3<R> point = mv(curr.Rinv, sub(q, curr.t));
                R gx = (R)((double)dx[(size_t)y * cols + x] / 9.0);
                R gy = (R)((double)dy[(size_t)y * cols + x] / 9.0);
                R invz = (R)1 / point.z;
                V3<R> left;
                left.x = gx * K.fx * invz;
                left.y = gy * K.fy * invz;
                left.z = -(left.x * point.x + left.y * point.y) * invz;
                V3<R> c = cross(point, left);
                row[0] = w * left.x;
                row[1] = w * left.y;
                row[2] = w * left.z;
                row[3] = w * c.x;
                row[4] = w * c.y;
                row[5] = w * c.z;
                row[6] = -w * (R)diff;
            }
            int k = 0;
            for (int a = 0; a < 7; ++a)
                for (int b = a; b < 7; ++b) {
                    const R p = row[a] * row[b];
                    loc[k] += (double)p;
                    if (rows_out) rows_out[(size_t)i * 29 + k] = (float)p;
                    ++k;
                }
            loc[28] += ok ? 1.0 : 0.0;
            if (rows_out) rows_out[(size_t)i * 29 + 28] = ok ? 1.f : 0.f;
        }
#pragma omp critical
        for (int j = 0; j < 29; ++j) sums[j] += loc[j];
    }
}

template void rgb_sums<float>(const uint8_t*, const uint8_t*, const float*, const float*,
                              const int16_t*, const int16_t*, int, int, const Pose<float>&,
                              const Pose<float>&, Intr<float>, double*, int*, float*, float*);
template void rgb_sums<double>(const uint8_t*, const uint8_t*, const double*, const double*,
                               const int16_t*, const int16_t*, int, int, const Pose<double>&,
                               const Pose<double>&, Intr<double>, double*, int*, float*, float*);

// ---- the precision form's sums: FP32 with compensation ---------------------
// The reference sums the per-pixel rows in plain FP32 (BlockReduce and
// Reduce<float, 29>, Reduction.cu:29-60, 315, 374). The precision form keeps
// FP32 but compensates: each partial sum is a pair (s, e) whose value is
// s + e. Every addition — a thread adding a product p, two pairs meeting —
// is Knuth's TwoSum of the sums (t = a + b, z = t - a,
// r = (a - (t - z)) + (b - z): r is the exact rounding error of t), the
// error terms added in FP32 (cascaded summation, Ogita-Rump-Oishi). The
// order is that of the GPU reduction this build runs (a decision of this
// build, not the reference's), so the result is bit-exact:
// - 256 threads per workgroup, nb = clamp(ceil(n / 1024), 1, 512)
//   workgroups (ceil(n / 256) for n <= 320 * 256 when the system spreads
//   small levels, icp_spread: one pixel per thread), stride = 256 nb; thread g adds the contributing pixels
//   g, g + stride, g + 2 stride, ... in that order (28 products, then 1 to
//   the count);
// - in each workgroup, for each sum, eight partial sums k = 0..7 each meet
//   the threads k, k + 8, ..., k + 248 in turn (from (0, 0)), then a
//   butterfly over the eight (distance 1, 2, 4);
// - over the workgroups likewise: partial k meets the workgroup rows
//   k, k + 8, ..., k + 504, or for nb > 128 by pairs of rows, 2k, 2k + 1,
//   2k + 16, 2k + 17, ..., 2k + 497 (rows past nb count as (0, 0)), then the
//   same butterfly; the sum is s + e rounded to FP32.
// `compensate` = false keeps every step but drops the error terms: plain
// FP32 in the same order (used by the tests to show what compensation buys).
struct K32 {
    float s, e;
};
static inline K32 k32_merge(K32 a, const K32& b, bool comp) {
    if (!comp) return K32{a.s + b.s, 0.f};
    const float t = a.s + b.s;
    const float z = t - a.s;
    const float r = (a.s - (t - z)) + (b.s - z);
    return K32{t, (a.e + b.e) + r};
}
static inline void k32_add(K32& a, float p, bool comp) {
    if (!comp) {
        a.s = a.s + p;
        return;
    }
    const float t = a.s + p;
    const float z = t - a.s;
    const float r = (a.s - (t - z)) + (p - z);
    a.e = a.e + r;
    a.s = t;
}
// Butterfly over eight lanes (distances 1, 2, 4); every lane ends with the
// same pair (TwoSum's error is exact, so a meeting is symmetric); lane 0's.
static inline K32 k32_fold8(K32 lane[8], bool comp) {
    for (int m = 1; m < 8; m <<= 1) {
        K32 nx[8];
        for (int k = 0; k < 8; ++k) nx[k] = k32_merge(lane[k], lane[k ^ m], comp);
        for (int k = 0; k < 8; ++k) lane[k] = nx[k];
    }
    return lane[0];
}

void sums_k32(const float* rows29, int n, bool compensate, bool spread, float* sums) {
    constexpr int kThreads = 256, kMaxWg = 512;
    const int kPix = spread && n <= 320 * kThreads ? 1 : 4;
    int nb = (n + kThreads * kPix - 1) / (kThreads * kPix);
    nb = nb < 1 ? 1 : (nb > kMaxWg ? kMaxWg : nb);
    const long stride = (long)nb * kThreads;
    std::vector<K32> part((size_t)nb * 29);
#pragma omp parallel for schedule(static)
    for (int w = 0; w < nb; ++w) {
        std::vector<K32> acc((size_t)kThreads * 29, K32{0.f, 0.f});
        for (int t = 0; t < kThreads; ++t) {
            K32* a = &acc[(size_t)t * 29];
            for (long i = (long)w * kThreads + t; i < n; i += stride) {
                const float* r = rows29 + (size_t)i * 29;
                if (r[28] == 0.f) continue;
                for (int j = 0; j < 28; ++j) k32_add(a[j], r[j], compensate);
                k32_add(a[28], 1.0f, compensate);
            }
        }
        for (int j = 0; j < 29; ++j) {
            K32 lane[8];
            for (int k = 0; k < 8; ++k) {
                K32 s{0.f, 0.f};
                for (int r = k; r < kThreads; r += 8)
                    s = k32_merge(s, acc[(size_t)r * 29 + j], compensate);
                lane[k] = s;
            }
            part[(size_t)w * 29 + j] = k32_fold8(lane, compensate);
        }
    }
    for (int j = 0; j < 29; ++j) {
        K32 lane[8];
        for (int k = 0; k < 8; ++k) {
            K32 s{0.f, 0.f};
            for (int r = 0; r < kMaxWg / 8; ++r) {
                // row r of thread k's sequence
                const int bi = nb > 128 ? 2 * (k + 8 * (r >> 1)) + (r & 1) : k + 8 * r;
                s = k32_merge(s, bi < nb ? part[(size_t)bi * 29 + j] : K32{0.f, 0.f},
                              compensate);
            }
            lane[k] = s;
        }
        const K32 v = k32_fold8(lane, compensate);
        sums[j] = v.s + v.e;
    }
}


// SO3Reduce (Tracking/Reduction.cu:67-160): the last frame's pixel bearing
// ((x - cx) / fx, (y - cy) / fy, 1) rotated by RcurrInv * Rlast, matched by
// the rounded projection (__float2int_rn: nearest, ties to even) with a
// 5-pixel border on both pixels; J = -(left x p) from the next frame's
// gradient at the match (float / 9.0f), r = -(I_next(u, v) - I_last(x, y)).
// 10 FP32 products accumulated in FP64, then the count.
void so3_sums(const uint8_t* nimg, const uint8_t* limg, const int16_t* dx, const int16_t* dy,
              int cols, int rows, const M3<float>& rcurr_inv, const M3<float>& rlast,
              Intr<float> K, double* sums) {
    const int N = cols * rows;
    for (int j = 0; j < 11; ++j) sums[j] = 0.0;
#pragma omp parallel
    {
        double loc[11] = {0};
#pragma omp for schedule(static)
        for (int k = 0; k < N; ++k) {
            const int y = k / cols, x = k - y * cols;
            const V3<float> vl{((float)x - K.cx) / K.fx, ((float)y - K.cy) / K.fy, 1.0f};
            const V3<float> p = mv(rcurr_inv, mv(rlast, vl));
            const int u = f2i(std::nearbyint(K.fx * p.x / p.z + K.cx));
            const int v = f2i(std::nearbyint(K.fy * p.y / p.z + K.cy));
            if (!(u >= 5 && v >= 5 && u < cols - 5 && v < rows - 5 && x >= 5 && y >= 5 &&
                  x < cols - 5 && y < rows - 5))
                continue;
            const float gx = (float)dx[(size_t)v * cols + u] / 9.0f;
            const float gy = (float)dy[(size_t)v * cols + u] / 9.0f;
            const float invz = 1.0f / p.z;
            V3<float> left;
            left.x = gx * K.fx * invz;
            left.y = gy * K.fy * invz;
            left.z = -(p.x * left.x + p.y * left.y) * invz;
            const V3<float> c = cross(left, p);
            const float row[4] = {-c.x, -c.y, -c.z,
                                  -((float)nimg[(size_t)v * cols + u] -
                                    (float)limg[(size_t)y * cols + x])};
            int q = 0;
            for (int a = 0; a < 4; ++a)
                for (int b = a; b < 4; ++b) loc[q++] += (double)(row[a] * row[b]);
            loc[10] += 1.0;
        }
#pragma omp critical
        for (int j = 0; j < 11; ++j) sums[j] += loc[j];
    }
}

// CreateMatrix<6,7> (Reduction.cu:6-16): upper-triangle sums -> mirrored A, b.
void create_matrix(const double* s, double* A, double* b) {
    int shift = 0;
    for (int i = 0; i < 6; ++i)
        for (int j = i; j < 7; ++j) {
            double value = s[shift++];
            if (j == 6)
                b[i] = value;
            else
                A[j * 6 + i] = A[i * 6 + j] = value;
        }
}

// x = A^-1 b by an LDL^T factorisation without pivoting; a zero pivot gives a
// zero solution component, as Eigen's LDLT::solve does (Tracking.cc:493).
// Equal to the Cholesky solution for symmetric positive definite A.
bool cholesky_solve6(const double* A, const double* b, double* x) {
    double L[36] = {0}, D[6];
    for (int j = 0; j < 6; ++j) {
        double d = A[j * 6 + j];
        for (int k = 0; k < j; ++k) d -= L[j * 6 + k] * L[j * 6 + k] * D[k];
        D[j] = d;
        L[j * 6 + j] = 1.0;
        for (int i = j + 1; i < 6; ++i) {
            double s = A[i * 6 + j];
            for (int k = 0; k < j; ++k) s -= L[i * 6 + k] * L[j * 6 + k] * D[k];
            L[i * 6 + j] = (d != 0.0) ? s / d : 0.0;
        }
    }
    double z[6];
    for (int i = 0; i < 6; ++i) {
        double s = b[i];
        for (int k = 0; k < i; ++k) s -= L[i * 6 + k] * z[k];
        z[i] = s;
    }
    bool ok = true;
    for (int i = 0; i < 6; ++i) {
        if (D[i] != 0.0)
            z[i] = z[i] / D[i];
        else {
            z[i] = 0.0;
            ok = false;
        }
    }
    for (int i = 5; i >= 0; --i) {
        double s = z[i];
        for (int k = i + 1; k < 6; ++k) s -= L[k * 6 + i] * x[k];
        x[i] = s;
    }
    return ok;
}

// SE(3) exponential with the tangent ordered (translation, rotation), the
// order of the Jacobian columns [-n, n x v] (Reduction.cu:276-277) that
// Sophus::SE3d::exp consumes (Tracking.cc:494). Rotation from the unit
// quaternion (cos(t/2), sin(t/2) w/t), translation V(w) v with
// V = I + (1-cos t)/t^2 W + (t - sin t)/t^3 W^2; both with the small-angle
// series below 1e-10.
void se3_exp(const double* xi, double* T) {
    const double* v = xi;
    const double* w = xi + 3;
    const double eps = 1e-10;
    double theta_sq = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    double theta = std::sqrt(theta_sq);
    double half = 0.5 * theta;
    double imag, real;
    if (theta_sq < eps * eps) {
        double t4 = theta_sq * theta_sq;
        imag = 0.5 - theta_sq / 48.0 + t4 / 3840.0;
        real = 1.0 - theta_sq / 8.0 + t4 / 384.0;
    } else {
        imag = std::sin(half) / theta;
        real = std::cos(half);
    }
    double qw = real, qx = imag * w[0], qy = imag * w[1], qz = imag * w[2];
    double tx = 2 * qx, ty = 2 * qy, tz = 2 * qz;
    double twx = tx * qw, twy = ty * qw, twz = tz * qw;
    double txx = tx * qx, txy = ty * qx, txz = tz * qx;
    double tyy = ty * qy, tyz = tz * qy, tzz = tz * qz;
    double Rm[9] = {1 - (tyy + tzz), txy - twz,       txz + twy,  //
                    txy + twz,       1 - (txx + tzz), tyz - twx,  //
                    txz - twy,       tyz + twx,       1 - (txx + tyy)};
    double V[9];
    if (theta < eps) {
        for (int i = 0; i < 9; ++i) V[i] = Rm[i];
    } else {
        double W[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
        double W2[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                W2[i * 3 + j] = W[i * 3 + 0] * W[0 * 3 + j] + W[i * 3 + 1] * W[1 * 3 + j] +
                                W[i * 3 + 2] * W[2 * 3 + j];
        double c1 = (1.0 - std::cos(theta)) / theta_sq;
        double c2 = (theta - std::sin(theta)) / (theta_sq * theta);
        for (int i = 0; i < 9; ++i) V[i] = (i % 4 == 0 ? 1.0 : 0.0) + c1 * W[i] + c2 * W2[i];
    }
    for (int i = 0; i < 3; ++i) {
        T[i * 4 + 0] = Rm[i * 3 + 0];
        T[i * 4 + 1] = Rm[i * 3 + 1];
        T[i * 4 + 2] = Rm[i * 3 + 2];
        T[i * 4 + 3] = V[i * 3 + 0] * v[0] + V[i * 3 + 1] * v[1] + V[i * 3 + 2] * v[2];
    }
    T[12] = T[13] = T[14] = 0.0;
    T[15] = 1.0;
}

void mat4_mul(const double* A, const double* B, double* C) {
    double t[16];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            t[i * 4 + j] = A[i * 4 + 0] * B[0 * 4 + j] + A[i * 4 + 1] * B[1 * 4 + j] +
                           A[i * 4 + 2] * B[2 * 4 + j] + A[i * 4 + 3] * B[3 * 4 + j];
    std::memcpy(C, t, sizeof(t));
}

// Inverse of a rigid transform [R t; 0 1] = [R^T, -R^T t; 0 1].
void rigid_inverse(const double* T, double* Ti) {
    double t[16];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) t[i * 4 + j] = T[j * 4 + i];
    for (int i = 0; i < 3; ++i)
        t[i * 4 + 3] = -(t[i * 4 + 0] * T[3] + t[i * 4 + 1] * T[7] + t[i * 4 + 2] * T[11]);
    t[12] = t[13] = t[14] = 0.0;
    t[15] = 1.0;
    std::memcpy(Ti, t, sizeof(t));
}

// Motion gate of ComputeSE3 (Tracking.cc:505-511): the norm of the sines of
// the XYZ Euler angles of the relative rotation and the norm of the relative
// translation. Both Euler solutions of a rotation give the same sine norm, so
// the textbook decomposition R = Rx(a) Ry(b) Rz(c) stands in for Eigen's
// eulerAngles(0, 1, 2).
void euler_sin_norm(const double* T, double* rot_norm, double* trans_norm) {
    double a = std::atan2(-T[6], T[10]);
    double b = std::atan2(T[2], std::sqrt(T[0] * T[0] + T[1] * T[1]));
    double c = std::atan2(-T[1], T[0]);
    double sa = std::sin(a), sb = std::sin(b), sc = std::sin(c);
    *rot_norm = std::sqrt(sa * sa + sb * sb + sc * sc);
    *trans_norm = std::sqrt(T[3] * T[3] + T[7] * T[7] + T[11] * T[11]);
}

}  // namespace oracle

using namespace oracle;

extern "C" {

void oracle_icp_sums(const float* vcurr, const float* vlast, const float* ncurr,
                     const float* nlast, int cols, int rows, const sicp_pose_f* curr,
                     const sicp_pose_f* last, sicp_intrinsics K, double* sums_fp64,
                     float* rows_out) {
    Pose<float> pc = pose_from_f<float>(curr->R, curr->Rinv, curr->t);
    Pose<float> pl = pose_from_f<float>(last->R, last->Rinv, last->t);
    icp_sums<float>(vcurr, vlast, ncurr, nlast, cols, rows, pc, pl,
                    Intr<float>{K.fx, K.fy, K.cx, K.cy}, sums_fp64, rows_out);
}

void oracle_sums_k32(const float* rows29, int n, int compensate, float* sums) {
    sums_k32(rows29, n, compensate != 0, true, sums);
}

void oracle_icp_sums_fp64(const double* vcurr, const double* vlast, const double* ncurr,
                          const double* nlast, int cols, int rows, const double* curr_pose,
                          const double* last_pose, const double* K4, double* sums) {
    Pose<double> pc = pose_from_d<double>(curr_pose);
    Pose<double> pl = pose_from_d<double>(last_pose);
    icp_sums<double>(vcurr, vlast, ncurr, nlast, cols, rows, pc, pl,
                     Intr<double>{K4[0], K4[1], K4[2], K4[3]}, sums, nullptr);
}

void oracle_rgb_sums(const uint8_t* nimg, const uint8_t* limg, const float* nvmap,
                     const float* lvmap, const int16_t* dx, const int16_t* dy, int cols, int rows,
                     const sicp_pose_f* curr, const sicp_pose_f* last, sicp_intrinsics K,
                     double* sums_fp64, int32_t* corresp, float* sigma) {
    Pose<float> pc = pose_from_f<float>(curr->R, curr->Rinv, curr->t);
    Pose<float> pl = pose_from_f<float>(last->R, last->Rinv, last->t);
    int res2[2];
    rgb_sums<float>(nimg, limg, nvmap, lvmap, dx, dy, cols, rows, pc, pl,
                    Intr<float>{K.fx, K.fy, K.cx, K.cy}, sums_fp64, res2, sigma, nullptr);
    if (corresp) {
        corresp[0] = res2[0];
        corresp[1] = res2[1];
    }
}

void oracle_so3_sums(const uint8_t* nimg, const uint8_t* limg, const int16_t* dx,
                     const int16_t* dy, int cols, int rows, const sicp_pose_f* curr,
                     const sicp_pose_f* last, sicp_intrinsics K, double* sums) {
    Pose<float> pc = pose_from_f<float>(curr->R, curr->Rinv, curr->t);
    Pose<float> pl = pose_from_f<float>(last->R, last->Rinv, last->t);
    so3_sums(nimg, limg, dx, dy, cols, rows, pc.Rinv, pl.R_, Intr<float>{K.fx, K.fy, K.cx, K.cy},
             sums);
}

void oracle_create_matrix(const float* sums29, double* A, double* b) {
    double s[29];
    for (int i = 0; i < 29; ++i) s[i] = (double)sums29[i];
    create_matrix(s, A, b);
}

int oracle_cholesky_solve6(const double* A, const double* b, double* x) {
    return cholesky_solve6(A, b, x) ? 1 : 0;
}

void oracle_se3_exp(const double* xi, double* T) { se3_exp(xi, T); }

// One Gauss-Newton update of ComputeSE3 (Tracking.cc:493-498).
int oracle_se3_update(const double* A, const double* b, double* delta, const double* last_pose,
                      double* next_pose, double* x_out) {
    double x[6];
    int ok = cholesky_solve6(A, b, x) ? 1 : 0;
    double dT[16], inv[16];
    se3_exp(x, dT);
    mat4_mul(dT, delta, delta);
    rigid_inverse(delta, inv);
    mat4_mul(last_pose, inv, next_pose);
    if (x_out)
        for (int i = 0; i < 6; ++i) x_out[i] = x[i];
    return ok;
}

void oracle_euler_sin_norm(const double* pose_a, const double* pose_b, double* rot_norm,
                           double* trans_norm) {
    double inv[16], p[16];
    rigid_inverse(pose_a, inv);
    mat4_mul(inv, pose_b, p);
    euler_sin_norm(p, rot_norm, trans_norm);
}

}  // extern "C"
