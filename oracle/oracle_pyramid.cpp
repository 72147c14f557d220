// oracle_pyramid.cpp — CPU restatement of the depth / intensity pyramid
// kernels of Tracking/Pyrdown.cu. TEST INFRASTRUCTURE ONLY (see oracle.h).
#include "oracle_impl.hpp"

namespace oracle {

// FilterDepthKernel (Tracking/Pyrdown.cu:72-108); constants sigSpace, sigRange
// (Pyrdown.cu:3-4). The window runs over rows [max(y-2,0), min(y+3, rows-1))
// and likewise for columns, so the last image row/column is never a tap;
// zero-depth taps are included (Q7).
template <class R>
void filter_depth(const uint16_t* depth, int cols, int rows, R scale_inv, R cutoff, R* raw,
                  R* filtered) {
    const R sig_space = (R)(0.5 / (4 * 4));
    const R sig_range = (R)(0.5 / (0.5 * 0.5));
    const R nanv = qnan<R>();
#pragma omp parallel for schedule(static)
    for (int y = 0; y < rows; ++y) {
        for (int x = 0; x < cols; ++x) {
            R center = (R)depth[(size_t)y * cols + x] * scale_inv;
            raw[(size_t)y * cols + x] = center < cutoff ? center : nanv;
            const int D = 5;
            int tx = std::min(x - D / 2 + D, cols - 1);
            int ty = std::min(y - D / 2 + D, rows - 1);
            R sum1 = 0, sum2 = 0;
            for (int cy = std::max(y - D / 2, 0); cy < ty; ++cy) {
                for (int cx = std::max(x - D / 2, 0); cx < tx; ++cx) {
                    R val = (R)depth[(size_t)cy * cols + cx] * scale_inv;
                    R space2 = (R)((x - cx) * (x - cx) + (y - cy) * (y - cy));
                    R color2 = (center - val) * (center - val);
                    R weight = rexp<R>(-(space2 * sig_space + color2 * sig_range));
                    sum1 += val * weight;
                    sum2 += weight;
                }
            }
            R fin = sum1 / sum2;
            filtered[(size_t)y * cols + x] = fin < cutoff ? fin : nanv;
        }
    }
}

// ComputeVMapKernel (Pyrdown.cu:123-145): invalid pixels get x = NaN (the
// reference writes only x and the depth into z; y and w are set to NaN here).
template <class R>
void compute_vmap(const R* depth, int cols, int rows, R invfx, R invfy, R cx, R cy, R cutoff,
                  R* vmap) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < rows; ++y) {
        for (int x = 0; x < cols; ++x) {
            R z = depth[(size_t)y * cols + x];
            R* v = vmap + ((size_t)y * cols + x) * 4;
            if (!isnan_(z) && (double)z > 0.1 && z < cutoff) {
                v[0] = z * ((R)x - cx) * invfx;
                v[1] = z * ((R)y - cy) * invfy;
                v[2] = z;
                v[3] = (R)1.0;
            } else {
                v[0] = qnan<R>();
                v[1] = qnan<R>();
                v[2] = z;
                v[3] = qnan<R>();
            }
        }
    }
}

// ComputeNMapKernel (Pyrdown.cu:160-181).
template <class R> void compute_nmap(const R* vmap, int cols, int rows, R* nmap) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < rows; ++y) {
        for (int x = 0; x < cols; ++x) {
            R* n = nmap + ((size_t)y * cols + x) * 4;
            if (x == cols - 1 || y == rows - 1) {
                n[0] = n[1] = n[2] = n[3] = qnan<R>();
                continue;
            }
            const R* c = vmap + ((size_t)y * cols + x) * 4;
            const R* r = vmap + ((size_t)y * cols + x + 1) * 4;
            const R* d = vmap + ((size_t)(y + 1) * cols + x) * 4;
            if (!isnan_(c[0]) && !isnan_(r[0]) && !isnan_(d[0])) {
                V3<R> a{r[0] - c[0], r[1] - c[1], r[2] - c[2]};
                V3<R> b{d[0] - c[0], d[1] - c[1], d[2] - c[2]};
                V3<R> nn = normalised(cross(a, b));
                n[0] = nn.x;
                n[1] = nn.y;
                n[2] = nn.z;
                n[3] = (R)1.0;
            } else {
                n[0] = n[1] = n[2] = n[3] = qnan<R>();
            }
        }
    }
}

// PyrDownGaussKernel (Pyrdown.cu:194-225). Taps over rows
// [max(0,2y-2), min(2y+3, srows-1)); the 5x5 binomial weight is indexed from
// the clipped window end, (ty-cy-1)*5 + (tx-cx-1) (Q7). NaN taps are skipped;
// the integer weight sum accumulates the float weights.
static const float kGauss[25] = {1, 4, 6, 4, 1, 4, 16, 24, 16, 4, 6, 24, 36,
                                 24, 6, 4, 16, 24, 16, 4, 1, 4, 6, 4, 1};
template <class T, class U, class R>
void pyrdown(const T* src, int scols, int srows, U* dst, int dcols, int drows) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < drows; ++y) {
        for (int x = 0; x < dcols; ++x) {
            const int D = 5;
            int tx = std::min(2 * x - D / 2 + D, scols - 1);
            int ty = std::min(2 * y - D / 2 + D, srows - 1);
            R sum = 0;
            int count = 0;
            for (int cy = std::max(0, 2 * y - D / 2); cy < ty; ++cy) {
                for (int cx = std::max(0, 2 * x - D / 2); cx < tx; ++cx) {
                    R s = (R)src[(size_t)cy * scols + cx];
                    if (!isnan_(s)) {
                        R g = (R)kGauss[(ty - cy - 1) * 5 + (tx - cx - 1)];
                        sum += s * g;
                        count = (int)((R)count + g);
                    }
                }
            }
            R q = sum / (R)count;
            dst[(size_t)y * dcols + x] = cast_out<U, R>(q);
        }
    }
}

// ImageToIntensityKernel (Pyrdown.cu:250-260): the literals are double and
// weight G, R, B in that order.
void image_to_intensity(const uint8_t* rgb, int cols, int rows, uint8_t* gray) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < cols * rows; ++i) {
        const uint8_t* p = rgb + (size_t)i * 3;
        int value = (int)(0.2125 * (double)p[1] + 0.7154 * (double)p[0] + 0.0721 * (double)p[2]);
        gray[i] = (uint8_t)value;
    }
}

// ComputeDerivativeImageKernel (Pyrdown.cu:274-304): the Sobel index counts
// down from 8 over the clipped window, so border windows use shifted taps.
void derivative_image(const uint8_t* img, int cols, int rows, int16_t* dx, int16_t* dy) {
    static const float sx[9] = {1, 0, -1, 2, 0, -2, 1, 0, -1};
    static const float sy[9] = {1, 2, 1, 0, 0, 0, -1, -2, -1};
#pragma omp parallel for schedule(static)
    for (int y = 0; y < rows; ++y) {
        for (int x = 0; x < cols; ++x) {
            float dxv = 0, dyv = 0;
            int k = 8;
            for (int j = std::max(y - 1, 0); j <= std::min(y + 1, rows - 1); ++j) {
                for (int i = std::max(x - 1, 0); i <= std::min(x + 1, cols - 1); ++i) {
                    float p = (float)img[(size_t)j * cols + i];
                    dxv += p * sx[k];
                    dyv += p * sy[k];
                    --k;
                }
            }
            dx[(size_t)y * cols + x] = (int16_t)(int)dxv;
            dy[(size_t)y * cols + x] = (int16_t)(int)dyv;
        }
    }
}

// ResizeMapKernel (Pyrdown.cu:318-346). The reference guards on the source
// dimensions; the destination guard is used here (same in-bounds results).
// The normal is the float4 mean normalised including w (Q6).
template <class R>
void resize_map(const R* vsrc, const R* nsrc, int scols, int srows, R* vdst, R* ndst,
                int dcols, int drows) {
    (void)srows;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < drows; ++y) {
        for (int x = 0; x < dcols; ++x) {
            const R* v00 = vsrc + ((size_t)(2 * y) * scols + 2 * x) * 4;
            const R* v01 = v00 + 4;
            const R* v10 = vsrc + ((size_t)(2 * y + 1) * scols + 2 * x) * 4;
            const R* v11 = v10 + 4;
            const R* n00 = nsrc + ((size_t)(2 * y) * scols + 2 * x) * 4;
            const R* n01 = n00 + 4;
            const R* n10 = nsrc + ((size_t)(2 * y + 1) * scols + 2 * x) * 4;
            const R* n11 = n10 + 4;
            R* vd = vdst + ((size_t)y * dcols + x) * 4;
            R* nd = ndst + ((size_t)y * dcols + x) * 4;
            if (isnan_(v00[0]) || isnan_(v01[0]) || isnan_(v10[0]) || isnan_(v11[0])) {
                vd[0] = vd[1] = vd[2] = vd[3] = qnan<R>();
            } else {
                for (int c = 0; c < 4; ++c) vd[c] = (((v00[c] + v01[c]) + v10[c]) + v11[c]) / (R)4;
            }
            if (isnan_(n00[0]) || isnan_(n01[0]) || isnan_(n10[0]) || isnan_(n11[0])) {
                nd[0] = nd[1] = nd[2] = nd[3] = qnan<R>();
            } else {
                V4<R> m;
                m.x = (((n00[0] + n01[0]) + n10[0]) + n11[0]) / (R)4;
                m.y = (((n00[1] + n01[1]) + n10[1]) + n11[1]) / (R)4;
                m.z = (((n00[2] + n01[2]) + n10[2]) + n11[2]) / (R)4;
                m.w = (((n00[3] + n01[3]) + n10[3]) + n11[3]) / (R)4;
                V4<R> q = normalised4(m);
                nd[0] = q.x;
                nd[1] = q.y;
                nd[2] = q.z;
                nd[3] = q.w;
            }
        }
    }
}

template void filter_depth<float>(const uint16_t*, int, int, float, float, float*, float*);
template void filter_depth<double>(const uint16_t*, int, int, double, double, double*, double*);
template void compute_vmap<float>(const float*, int, int, float, float, float, float, float,
                                  float*);
template void compute_vmap<double>(const double*, int, int, double, double, double, double,
                                   double, double*);
template void compute_nmap<float>(const float*, int, int, float*);
template void compute_nmap<double>(const double*, int, int, double*);
template void pyrdown<float, float, float>(const float*, int, int, float*, int, int);
template void pyrdown<double, double, double>(const double*, int, int, double*, int, int);
template void pyrdown<uint8_t, uint8_t, float>(const uint8_t*, int, int, uint8_t*, int, int);
template void resize_map<float>(const float*, const float*, int, int, float*, float*, int, int);
template void resize_map<double>(const double*, const double*, int, int, double*, double*, int,
                                 int);

}  // namespace oracle

using namespace oracle;

extern "C" {

float oracle_expf(float x) { return sicp::expf_cw(x); }

void oracle_filter_depth(const uint16_t* depth, int cols, int rows, float depth_scale,
                         float depth_cutoff, float* raw, float* filtered) {
    // FilterDepth passes 1.0 / depthScale, a double rounded to float (Pyrdown.cu:117).
    float inv = (float)(1.0 / (double)depth_scale);
    filter_depth<float>(depth, cols, rows, inv, depth_cutoff, raw, filtered);
}

void oracle_compute_vmap(const float* depth, int cols, int rows, float fx, float fy, float cx,
                         float cy, float depth_cutoff, float* vmap) {
    // ComputeVMap passes 1.0 / fx rounded to float (Pyrdown.cu:154).
    compute_vmap<float>(depth, cols, rows, (float)(1.0 / (double)fx), (float)(1.0 / (double)fy),
                        cx, cy, depth_cutoff, vmap);
}

void oracle_compute_nmap(const float* vmap, int cols, int rows, float* nmap) {
    compute_nmap<float>(vmap, cols, rows, nmap);
}

void oracle_pyrdown_f32(const float* src, int scols, int srows, float* dst, int dcols,
                        int drows) {
    pyrdown<float, float, float>(src, scols, srows, dst, dcols, drows);
}

void oracle_pyrdown_u8(const uint8_t* src, int scols, int srows, uint8_t* dst, int dcols,
                       int drows) {
    pyrdown<uint8_t, uint8_t, float>(src, scols, srows, dst, dcols, drows);
}

void oracle_image_to_intensity(const uint8_t* rgb, int cols, int rows, uint8_t* gray) {
    image_to_intensity(rgb, cols, rows, gray);
}

void oracle_derivative_image(const uint8_t* gray, int cols, int rows, int16_t* dx, int16_t* dy) {
    derivative_image(gray, cols, rows, dx, dy);
}

void oracle_resize_map(const float* vsrc, const float* nsrc, int scols, int srows, float* vdst,
                       float* ndst, int dcols, int drows) {
    resize_map<float>(vsrc, nsrc, scols, srows, vdst, ndst, dcols, drows);
}

}  // extern "C"
