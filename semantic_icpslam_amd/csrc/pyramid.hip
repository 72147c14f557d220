// pyramid.hip — depth / intensity pyramid kernels for gfx950.
//
// One thread per output pixel; 64x4 workgroups so each wavefront covers 64
// consecutive pixels of a row (coalesced row loads and stores). Every kernel
// is a bandwidth-bound stencil; the 5x5 windows are served from L1/L2.
#include "sicp_kernels.hpp"

namespace sicp {

namespace {

constexpr int kTX = 64, kTY = 4;

inline dim3 grid2d(int cols, int rows) { return dim3(div_up(cols, kTX), div_up(rows, kTY)); }

// FilterDepthKernel (Tracking/Pyrdown.cu:72-108). Window rows
// [max(y-2,0), min(y+3, rows-1)), columns likewise (the last row / column is
// never a tap); zero-depth taps included; sigSpace = 0.5/16, sigRange = 2
// (Pyrdown.cu:3-4).
__global__ __launch_bounds__(256) void filter_depth_kernel(View<uint16_t> depth, View<float> raw,
                                                           View<float> filtered, float scale_inv,
                                                           float cutoff) {
    const int x = blockIdx.x * kTX + threadIdx.x;
    const int y = blockIdx.y * kTY + threadIdx.y;
    if (x >= depth.cols || y >= depth.rows) return;
    const float sig_space = 0.03125f, sig_range = 2.0f;
    float center = (float)depth.at(y, x) * scale_inv;
    raw.at(y, x) = center < cutoff ? center : qnanf();
    const int tx = min(x + 3, depth.cols - 1);
    const int ty = min(y + 3, depth.rows - 1);
    float sum1 = 0.f, sum2 = 0.f;
    for (int cy = max(y - 2, 0); cy < ty; ++cy) {
        const uint16_t* row = &depth.at(cy, 0);
        for (int cx = max(x - 2, 0); cx < tx; ++cx) {
            float val = (float)row[cx] * scale_inv;
            float space2 = (float)((x - cx) * (x - cx) + (y - cy) * (y - cy));
            float color2 = (center - val) * (center - val);
            float weight = expf_cw(-(space2 * sig_space + color2 * sig_range));
            sum1 += val * weight;
            sum2 += weight;
        }
    }
    float fin = sum1 / sum2;
    filtered.at(y, x) = fin < cutoff ? fin : qnanf();
}

// ComputeVMapKernel (Pyrdown.cu:123-145). The depth test compares in double
// (z > 0.1); invalid pixels carry x = NaN.
__global__ __launch_bounds__(256) void vmap_kernel(View<float> depth, View<float4> vmap,
                                                   float invfx, float invfy, float cx, float cy,
                                                   float cutoff) {
    const int x = blockIdx.x * kTX + threadIdx.x;
    const int y = blockIdx.y * kTY + threadIdx.y;
    if (x >= depth.cols || y >= depth.rows) return;
    float z = depth.at(y, x);
    float4 v;
    if (!isnan_(z) && (double)z > 0.1 && z < cutoff) {
        v.x = z * ((float)x - cx) * invfx;
        v.y = z * ((float)y - cy) * invfy;
        v.z = z;
        v.w = 1.0f;
    } else {
        v.x = qnanf();
        v.y = qnanf();
        v.z = z;
        v.w = qnanf();
    }
    vmap.at(y, x) = v;
}

// ComputeNMapKernel (Pyrdown.cu:160-181).
__global__ __launch_bounds__(256) void nmap_kernel(View<float4> vmap, View<float4> nmap) {
    const int x = blockIdx.x * kTX + threadIdx.x;
    const int y = blockIdx.y * kTY + threadIdx.y;
    if (x >= vmap.cols || y >= vmap.rows) return;
    const float nan = qnanf();
    if (x == vmap.cols - 1 || y == vmap.rows - 1) {
        nmap.at(y, x) = make_float4(nan, nan, nan, nan);
        return;
    }
    float4 c = vmap.at(y, x);
    float4 r = vmap.at(y, x + 1);
    float4 d = vmap.at(y + 1, x);
    if (!isnan_(c.x) && !isnan_(r.x) && !isnan_(d.x)) {
        v3 n = normalised(cross(sub(xyz(r), xyz(c)), sub(xyz(d), xyz(c))));
        nmap.at(y, x) = make_float4(n.x, n.y, n.z, 1.0f);
    } else {
        nmap.at(y, x) = make_float4(nan, nan, nan, nan);
    }
}

// PyrDownGaussKernel (Pyrdown.cu:194-225): 5x5 binomial, weight index taken
// from the clipped window end, NaN taps skipped, integer weight count.
__constant__ float kGauss[25] = {1, 4, 6, 4, 1, 4, 16, 24, 16, 4, 6, 24, 36,
                                 24, 6, 4, 16, 24, 16, 4, 1, 4, 6, 4, 1};

template <class T> SICP_DEV T out_cast(float v);
template <> SICP_DEV float out_cast<float>(float v) { return v; }
template <> SICP_DEV uint8_t out_cast<uint8_t>(float v) { return (uint8_t)(int)v; }

template <class T>
__global__ __launch_bounds__(256) void pyrdown_kernel(View<T> src, View<T> dst) {
    const int x = blockIdx.x * kTX + threadIdx.x;
    const int y = blockIdx.y * kTY + threadIdx.y;
    if (x >= dst.cols || y >= dst.rows) return;
    const int tx = min(2 * x + 3, src.cols - 1);
    const int ty = min(2 * y + 3, src.rows - 1);
    float sum = 0.f;
    int count = 0;
    for (int cy = max(0, 2 * y - 2); cy < ty; ++cy) {
        const T* row = &src.at(cy, 0);
        for (int cx = max(0, 2 * x - 2); cx < tx; ++cx) {
            float s = (float)row[cx];
            if (!isnan_(s)) {
                float g = kGauss[(ty - cy - 1) * 5 + (tx - cx - 1)];
                sum += s * g;
                count = (int)((float)count + g);
            }
        }
    }
    dst.at(y, x) = out_cast<T>(sum / (float)count);
}

// ImageToIntensityKernel (Pyrdown.cu:250-260): double weights of G, R, B.
__global__ __launch_bounds__(256) void intensity_kernel(View<uint8_t> rgb, View<uint8_t> gray) {
    const int x = blockIdx.x * kTX + threadIdx.x;
    const int y = blockIdx.y * kTY + threadIdx.y;
    if (x >= gray.cols || y >= gray.rows) return;
    const uint8_t* p = &rgb.at(y, 3 * x);
    int value = (int)(0.2125 * (double)p[1] + 0.7154 * (double)p[0] + 0.0721 * (double)p[2]);
    gray.at(y, x) = (uint8_t)value;
}

// ComputeDerivativeImageKernel (Pyrdown.cu:280-304): the tap index counts
// down from 8 over the clipped 3x3 window.
__constant__ float kSobelX[9] = {1, 0, -1, 2, 0, -2, 1, 0, -1};
__constant__ float kSobelY[9] = {1, 2, 1, 0, 0, 0, -1, -2, -1};

__global__ __launch_bounds__(256) void derivative_kernel(View<uint8_t> img, View<int16_t> dx,
                                                         View<int16_t> dy) {
    const int x = blockIdx.x * kTX + threadIdx.x;
    const int y = blockIdx.y * kTY + threadIdx.y;
    if (x >= img.cols || y >= img.rows) return;
    float dxv = 0.f, dyv = 0.f;
    int k = 8;
    for (int j = max(y - 1, 0); j <= min(y + 1, img.rows - 1); ++j) {
        for (int i = max(x - 1, 0); i <= min(x + 1, img.cols - 1); ++i) {
            float p = (float)img.at(j, i);
            dxv += p * kSobelX[k];
            dyv += p * kSobelY[k];
            --k;
        }
    }
    dx.at(y, x) = (int16_t)(int)dxv;
    dy.at(y, x) = (int16_t)(int)dyv;
}

// ResizeMapKernel (Pyrdown.cu:318-346), guarded on the destination size.
__global__ __launch_bounds__(256) void resize_map_kernel(View<float4> vsrc, View<float4> nsrc,
                                                         View<float4> vdst, View<float4> ndst) {
    const int x = blockIdx.x * kTX + threadIdx.x;
    const int y = blockIdx.y * kTY + threadIdx.y;
    if (x >= vdst.cols || y >= vdst.rows) return;
    const float nan = qnanf();
    float4 v00 = vsrc.at(2 * y, 2 * x), v01 = vsrc.at(2 * y, 2 * x + 1);
    float4 v10 = vsrc.at(2 * y + 1, 2 * x), v11 = vsrc.at(2 * y + 1, 2 * x + 1);
    float4 n00 = nsrc.at(2 * y, 2 * x), n01 = nsrc.at(2 * y, 2 * x + 1);
    float4 n10 = nsrc.at(2 * y + 1, 2 * x), n11 = nsrc.at(2 * y + 1, 2 * x + 1);
    if (isnan_(v00.x) || isnan_(v01.x) || isnan_(v10.x) || isnan_(v11.x)) {
        vdst.at(y, x) = make_float4(nan, nan, nan, nan);
    } else {
        float4 o;
        o.x = (((v00.x + v01.x) + v10.x) + v11.x) / 4.0f;
        o.y = (((v00.y + v01.y) + v10.y) + v11.y) / 4.0f;
        o.z = (((v00.z + v01.z) + v10.z) + v11.z) / 4.0f;
        o.w = (((v00.w + v01.w) + v10.w) + v11.w) / 4.0f;
        vdst.at(y, x) = o;
    }
    if (isnan_(n00.x) || isnan_(n01.x) || isnan_(n10.x) || isnan_(n11.x)) {
        ndst.at(y, x) = make_float4(nan, nan, nan, nan);
    } else {
        float mx = (((n00.x + n01.x) + n10.x) + n11.x) / 4.0f;
        float my = (((n00.y + n01.y) + n10.y) + n11.y) / 4.0f;
        float mz = (((n00.z + n01.z) + n10.z) + n11.z) / 4.0f;
        float mw = (((n00.w + n01.w) + n10.w) + n11.w) / 4.0f;
        float nn = sqrtf(mx * mx + my * my + mz * mz + mw * mw);
        ndst.at(y, x) = make_float4(mx / nn, my / nn, mz / nn, mw / nn);
    }
}

}  // namespace

int launch_filter_depth(View<uint16_t> depth, View<float> raw, View<float> filtered,
                        float scale_inv, float cutoff, hipStream_t s) {
    filter_depth_kernel<<<grid2d(depth.cols, depth.rows), dim3(kTX, kTY), 0, s>>>(
        depth, raw, filtered, scale_inv, cutoff);
    SICP_LAUNCH_CHECK("filter_depth_kernel");
    return SICP_OK;
}

int launch_vmap(View<float> depth, View<float4> vmap, float invfx, float invfy, float cx,
                float cy, float cutoff, hipStream_t s) {
    vmap_kernel<<<grid2d(depth.cols, depth.rows), dim3(kTX, kTY), 0, s>>>(depth, vmap, invfx,
                                                                           invfy, cx, cy, cutoff);
    SICP_LAUNCH_CHECK("vmap_kernel");
    return SICP_OK;
}

int launch_nmap(View<float4> vmap, View<float4> nmap, hipStream_t s) {
    nmap_kernel<<<grid2d(vmap.cols, vmap.rows), dim3(kTX, kTY), 0, s>>>(vmap, nmap);
    SICP_LAUNCH_CHECK("nmap_kernel");
    return SICP_OK;
}

int launch_pyrdown_f32(View<float> src, View<float> dst, hipStream_t s) {
    pyrdown_kernel<float><<<grid2d(dst.cols, dst.rows), dim3(kTX, kTY), 0, s>>>(src, dst);
    SICP_LAUNCH_CHECK("pyrdown_kernel<float>");
    return SICP_OK;
}

int launch_pyrdown_u8(View<uint8_t> src, View<uint8_t> dst, hipStream_t s) {
    pyrdown_kernel<uint8_t><<<grid2d(dst.cols, dst.rows), dim3(kTX, kTY), 0, s>>>(src, dst);
    SICP_LAUNCH_CHECK("pyrdown_kernel<uint8_t>");
    return SICP_OK;
}

int launch_intensity(View<uint8_t> rgb, View<uint8_t> gray, hipStream_t s) {
    intensity_kernel<<<grid2d(gray.cols, gray.rows), dim3(kTX, kTY), 0, s>>>(rgb, gray);
    SICP_LAUNCH_CHECK("intensity_kernel");
    return SICP_OK;
}

int launch_derivative(View<uint8_t> img, View<int16_t> dx, View<int16_t> dy, hipStream_t s) {
    derivative_kernel<<<grid2d(img.cols, img.rows), dim3(kTX, kTY), 0, s>>>(img, dx, dy);
    SICP_LAUNCH_CHECK("derivative_kernel");
    return SICP_OK;
}

int launch_resize_map(View<float4> vsrc, View<float4> nsrc, View<float4> vdst, View<float4> ndst,
                      hipStream_t s) {
    resize_map_kernel<<<grid2d(vdst.cols, vdst.rows), dim3(kTX, kTY), 0, s>>>(vsrc, nsrc, vdst,
                                                                              ndst);
    SICP_LAUNCH_CHECK("resize_map_kernel");
    return SICP_OK;
}

}  // namespace sicp
