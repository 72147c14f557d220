// pyramid.hip — depth / intensity pyramid kernels for gfx950.
//
// One thread per output pixel; 64x4 workgroups so each wavefront covers 64
// consecutive pixels of a row (coalesced row loads and stores). Every kernel
// is a bandwidth-bound stencil; the 5x5 windows are served from L1/L2.
#include "sicp_kernels.hpp"

namespace sicp {

namespace {

constexpr int kTX = 64, kTY = 4;
constexpr int kModelTX = kTX;  // model pyramid tile: kModelTX x 8 level-1 pixels
static_assert(kTY == 4 && kModelTX == 64, "two level-1 rows per thread, 128 level-2 pixels");

inline dim3 grid2d(int cols, int rows) { return dim3(div_up(cols, kTX), div_up(rows, kTY)); }

// FilterDepthKernel (Tracking/Pyrdown.cu:72-108). Window rows
// [max(y-2,0), min(y+3, rows-1)), columns likewise (the last row / column is
// never a tap); zero-depth taps included; sigSpace = 0.5/16, sigRange = 2
// (Pyrdown.cu:3-4).
// The 5x5 window is unrolled: every tap's load is issued at a clamped
// address first, then the taps inside the clipped window are summed in the
// reference's row-major order.
SICP_DEV void filter_px(const View<uint16_t>& depth, const View<float>& raw,
                        const View<float>& filtered, float scale_inv, float cutoff, int x, int y) {
    const float sig_space = 0.03125f, sig_range = 2.0f;
    const int tx = min(x + 3, depth.cols - 1);
    const int ty = min(y + 3, depth.rows - 1);
    uint16_t tap[25];
#pragma unroll
    for (int dy = 0; dy < 5; ++dy) {
        const uint16_t* row = &depth.at(min(max(y - 2 + dy, 0), depth.rows - 1), 0);
#pragma unroll
        for (int dx = 0; dx < 5; ++dx) tap[dy * 5 + dx] = row[min(max(x - 2 + dx, 0), depth.cols - 1)];
    }
    float center = (float)tap[12] * scale_inv;
    raw.at(y, x) = center < cutoff ? center : qnanf();
    float sum1 = 0.f, sum2 = 0.f;
#pragma unroll
    for (int dy = 0; dy < 5; ++dy) {
        const int cy = y - 2 + dy;
#pragma unroll
        for (int dx = 0; dx < 5; ++dx) {
            const int cx = x - 2 + dx;
            // Every tap is computed; one outside the clipped window adds +0
            // through a select (the sums are never -0, so their bits do not
            // change). One straight block lets the compiler interleave the
            // 25 taps' exponentials; a branch per tap kept them apart.
            const bool in = cy >= 0 && cy < ty && cx >= 0 && cx < tx;
            float val = (float)tap[dy * 5 + dx] * scale_inv;
            float space2 = (float)((x - cx) * (x - cx) + (y - cy) * (y - cy));
            float color2 = (center - val) * (center - val);
            float weight = expf_cw(-(space2 * sig_space + color2 * sig_range));
            sum1 += in ? val * weight : 0.f;
            sum2 += in ? weight : 0.f;
        }
    }
    float fin = sum1 / sum2;
    filtered.at(y, x) = fin < cutoff ? fin : qnanf();
}

__global__ __launch_bounds__(256) void filter_depth_kernel(View<uint16_t> depth, View<float> raw,
                                                           View<float> filtered, float scale_inv,
                                                           float cutoff) {
    const int x = blockIdx.x * kTX + threadIdx.x;
    const int y = blockIdx.y * kTY + threadIdx.y;
    if (x >= depth.cols || y >= depth.rows) return;
    filter_px(depth, raw, filtered, scale_inv, cutoff, x, y);
}

// ComputeVMapKernel (Pyrdown.cu:123-145). The depth test compares in double
// (z > 0.1); invalid pixels carry x = NaN.
SICP_DEV float4 vmap_px(const View<float>& depth, int x, int y, float invfx, float invfy, float cx,
                        float cy, float cutoff) {
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
    return v;
}

__global__ __launch_bounds__(256) void vmap_kernel(View<float> depth, View<float4> vmap,
                                                   float invfx, float invfy, float cx, float cy,
                                                   float cutoff) {
    const int x = blockIdx.x * kTX + threadIdx.x;
    const int y = blockIdx.y * kTY + threadIdx.y;
    if (x >= depth.cols || y >= depth.rows) return;
    vmap.at(y, x) = vmap_px(depth, x, y, invfx, invfy, cx, cy, cutoff);
}

// ComputeNMapKernel (Pyrdown.cu:160-181).
SICP_DEV float4 nmap_of(float4 c, float4 r, float4 d) {
    const float nan = qnanf();
    if (!isnan_(c.x) && !isnan_(r.x) && !isnan_(d.x)) {
        v3 n = normalised(cross(sub(xyz(r), xyz(c)), sub(xyz(d), xyz(c))));
        return make_float4(n.x, n.y, n.z, 1.0f);
    }
    return make_float4(nan, nan, nan, nan);
}

__global__ __launch_bounds__(256) void nmap_kernel(View<float4> vmap, View<float4> nmap) {
    const int x = blockIdx.x * kTX + threadIdx.x;
    const int y = blockIdx.y * kTY + threadIdx.y;
    if (x >= vmap.cols || y >= vmap.rows) return;
    const float nan = qnanf();
    if (x == vmap.cols - 1 || y == vmap.rows - 1) {
        nmap.at(y, x) = make_float4(nan, nan, nan, nan);
        return;
    }
    nmap.at(y, x) = nmap_of(vmap.at(y, x), vmap.at(y, x + 1), vmap.at(y + 1, x));
}

// PyrDownGaussKernel (Pyrdown.cu:194-225): 5x5 binomial, weight index taken
// from the clipped window end, NaN taps skipped, integer weight count.
// 5x5 binomial {1,4,6,4,1} x {1,4,6,4,1} of the reference, as a product of
// its two 1-D taps (exact in float).
SICP_DEV float binom5(int i) { return i == 2 ? 6.f : (i == 1 || i == 3) ? 4.f : 1.f; }

template <class T> SICP_DEV T out_cast(float v);
template <> SICP_DEV float out_cast<float>(float v) { return v; }
template <> SICP_DEV uint8_t out_cast<uint8_t>(float v) { return (uint8_t)(int)v; }

// One output pixel from a source read through `src(y, x)` (global memory or
// an LDS tile) of size scols x srows.
template <class T, class Src>
SICP_DEV T pyrdown_px(const Src& src, int scols, int srows, int x, int y) {
    const int tx = min(2 * x + 3, scols - 1);
    const int ty = min(2 * y + 3, srows - 1);
    T tap[25];
#pragma unroll
    for (int dy = 0; dy < 5; ++dy) {
        const int cy = min(max(2 * y - 2 + dy, 0), srows - 1);
#pragma unroll
        for (int dx = 0; dx < 5; ++dx) tap[dy * 5 + dx] = src(cy, min(max(2 * x - 2 + dx, 0), scols - 1));
    }
    float sum = 0.f;
    int count = 0;
#pragma unroll
    for (int dy = 0; dy < 5; ++dy) {
        const int cy = 2 * y - 2 + dy;
#pragma unroll
        for (int dx = 0; dx < 5; ++dx) {
            const int cx = 2 * x - 2 + dx;
            // Selects instead of branches, as in filter_px.
            const float s = (float)tap[dy * 5 + dx];
            const bool use = cy >= 0 && cy < ty && cx >= 0 && cx < tx && !isnan_(s);
            const float g = binom5(ty - cy - 1) * binom5(tx - cx - 1);
            sum += use ? s * g : 0.f;
            count = use ? (int)((float)count + g) : count;
        }
    }
    return out_cast<T>(sum / (float)count);
}

template <class T> struct GlobalSrc {
    View<T> v;
    SICP_DEV T operator()(int y, int x) const { return v.at(y, x); }
};

template <class T>
__global__ __launch_bounds__(256) void pyrdown_kernel(View<T> src, View<T> dst) {
    const int x = blockIdx.x * kTX + threadIdx.x;
    const int y = blockIdx.y * kTY + threadIdx.y;
    if (x >= dst.cols || y >= dst.rows) return;
    dst.at(y, x) = pyrdown_px<T>(GlobalSrc<T>{src}, src.cols, src.rows, x, y);
}

// ImageToIntensityKernel (Pyrdown.cu:250-260): double weights of G, R, B.
SICP_DEV void intensity_px(const View<uint8_t>& rgb, const View<uint8_t>& gray, int x, int y) {
    const uint8_t* p = &rgb.at(y, 3 * x);
    int value = (int)(0.2125 * (double)p[1] + 0.7154 * (double)p[0] + 0.0721 * (double)p[2]);
    gray.at(y, x) = (uint8_t)value;
}

__global__ __launch_bounds__(256) void intensity_kernel(View<uint8_t> rgb, View<uint8_t> gray) {
    const int x = blockIdx.x * kTX + threadIdx.x;
    const int y = blockIdx.y * kTY + threadIdx.y;
    if (x >= gray.cols || y >= gray.rows) return;
    intensity_px(rgb, gray, x, y);
}

// ComputeDerivativeImageKernel (Pyrdown.cu:280-304): the tap index counts
// down from 8 over the clipped 3x3 window.
// Sobel pair of the reference, {1,0,-1, 2,0,-2, 1,0,-1} and its transpose,
// written as closed forms of the tap index below.

SICP_DEV int2 derivative_px(const View<uint8_t>& img, const View<int16_t>& dx,
                            const View<int16_t>& dy, int x, int y) {
    uint8_t tap[9];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const uint8_t* row = &img.at(min(max(y - 1 + j, 0), img.rows - 1), 0);
#pragma unroll
        for (int i = 0; i < 3; ++i) tap[j * 3 + i] = row[min(max(x - 1 + i, 0), img.cols - 1)];
    }
    float dxv = 0.f, dyv = 0.f;
    int k = 8;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            // Selects instead of branches, as in filter_px; k counts the
            // taps inside the image, as the reference's index does.
            const int cy = y - 1 + j, cx = x - 1 + i;
            const bool in = cy >= 0 && cy < img.rows && cx >= 0 && cx < img.cols;
            const float p = (float)tap[j * 3 + i];
            const int r = k / 3, c = k - 3 * r;
            const float wx = p * ((float)(1 - c) * (r == 1 ? 2.f : 1.f));
            const float wy = p * ((float)(1 - r) * (c == 1 ? 2.f : 1.f));
            dxv += in ? wx : 0.f;
            dyv += in ? wy : 0.f;
            k -= in ? 1 : 0;
        }
    }
    const int16_t gx = (int16_t)(int)dxv, gy = (int16_t)(int)dyv;
    dx.at(y, x) = gx;
    dy.at(y, x) = gy;
    return make_int2(gx, gy);
}

// The tests of RGBReduction::FindCorresp (Tracking/Reduction.cu:420-433) that
// do not depend on the pose: a 5-pixel border, a positive intensity over the
// window rows y-2..y+1, columns x-2..x+1 (the loops' bounds), a non-zero
// gradient (sqrtf(gx^2 + gy^2) > 0; the int sum cannot overflow for int16
// inputs), a valid vertex. Returns the pixel's intensity when they pass,
// else -1.
SICP_DEV int16_t rgb_key_px(const View<uint8_t>& img, int gx, int gy, float vx, int x, int y) {
    if (!(x >= 5 && x < img.cols - 5 && y >= 5 && y < img.rows - 5)) return -1;
    bool ok = true;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const uint8_t* row = &img.at(y - 2 + a, x - 2);
#pragma unroll
        for (int b = 0; b < 4; ++b) ok = ok && row[b] > 0;
    }
    if (!ok || gx * gx + gy * gy == 0 || isnan_(vx)) return -1;
    return (int16_t)img.at(y, x);
}

__global__ __launch_bounds__(256) void rgb_key_kernel(View<uint8_t> img, View<int16_t> dx,
                                                      View<int16_t> dy, View<float4> vmap,
                                                      View<int16_t> key) {
    const int x = blockIdx.x * kTX + threadIdx.x;
    const int y = blockIdx.y * kTY + threadIdx.y;
    if (x >= img.cols || y >= img.rows) return;
    key.at(y, x) = rgb_key_px(img, dx.at(y, x), dy.at(y, x), vmap.at(y, x).x, x, y);
}

__global__ __launch_bounds__(256) void derivative_kernel(View<uint8_t> img, View<int16_t> dx,
                                                         View<int16_t> dy) {
    const int x = blockIdx.x * kTX + threadIdx.x;
    const int y = blockIdx.y * kTY + threadIdx.y;
    if (x >= img.cols || y >= img.rows) return;
    derivative_px(img, dx, dy, x, y);
}

// ResizeMapKernel (Pyrdown.cu:318-346), guarded on the destination size.
// The 2x2 block of `v`/`n` (row-major 00, 01, 10, 11) to one coarser pixel.
SICP_DEV void resize_px(const float4 v[4], const float4 n[4], float4& vo, float4& no) {
    const float nan = qnanf();
    if (isnan_(v[0].x) || isnan_(v[1].x) || isnan_(v[2].x) || isnan_(v[3].x)) {
        vo = make_float4(nan, nan, nan, nan);
    } else {
        vo.x = (((v[0].x + v[1].x) + v[2].x) + v[3].x) / 4.0f;
        vo.y = (((v[0].y + v[1].y) + v[2].y) + v[3].y) / 4.0f;
        vo.z = (((v[0].z + v[1].z) + v[2].z) + v[3].z) / 4.0f;
        vo.w = (((v[0].w + v[1].w) + v[2].w) + v[3].w) / 4.0f;
    }
    if (isnan_(n[0].x) || isnan_(n[1].x) || isnan_(n[2].x) || isnan_(n[3].x)) {
        no = make_float4(nan, nan, nan, nan);
    } else {
        float mx = (((n[0].x + n[1].x) + n[2].x) + n[3].x) / 4.0f;
        float my = (((n[0].y + n[1].y) + n[2].y) + n[3].y) / 4.0f;
        float mz = (((n[0].z + n[1].z) + n[2].z) + n[3].z) / 4.0f;
        float mw = (((n[0].w + n[1].w) + n[2].w) + n[3].w) / 4.0f;
        float nn = sqrtf(mx * mx + my * my + mz * mz + mw * mw);
        no = make_float4(mx / nn, my / nn, mz / nn, mw / nn);
    }
}

SICP_DEV void resize_from(const View<float4>& vsrc, const View<float4>& nsrc, int x, int y,
                          float4& vo, float4& no) {
    float4 v[4], n[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[k] = vsrc.at(2 * y + (k >> 1), 2 * x + (k & 1));
        n[k] = nsrc.at(2 * y + (k >> 1), 2 * x + (k & 1));
    }
    resize_px(v, n, vo, no);
}

__global__ __launch_bounds__(256) void resize_map_kernel(View<float4> vsrc, View<float4> nsrc,
                                                         View<float4> vdst, View<float4> ndst) {
    const int x = blockIdx.x * kTX + threadIdx.x;
    const int y = blockIdx.y * kTY + threadIdx.y;
    if (x >= vdst.cols || y >= vdst.rows) return;
    float4 vo, no;
    resize_from(vsrc, nsrc, x, y, vo, no);
    vdst.at(y, x) = vo;
    ndst.at(y, x) = no;
}

// ------------------------------------------------ whole-frame pyramid ----
// Frame::Create / FillImages / ResizeImages (Core/Frame.cc:33-98) in three
// launches instead of seventeen. Each launch runs the same per-pixel
// functions as the single-image kernels above, so the results are
// identical; what changes is only which launch computes which pixel.
//
//  1. level 0: bilateral filter + intensity of the new frame, and the model
//     pyramid of the last frame (levels 1 and 2, the latter straight from
//     level 0 through the same 2x2 means);
//  2. PyrDown to levels 1 and 2 of depth and intensity: a workgroup owns an
//     8x4 tile of level 2, computes the 20x12 level-1 window it needs into
//     LDS (writing the level-1 pixels it owns), then its level-2 pixels;
//  3. vertex map, normal map (the two neighbour vertices recomputed from
//     depth) and derivatives for every level.

// Workgroup b of a launch that covers several 64x4-tiled images in turn.
SICP_DEV bool take_tile(int& b, int cols, int rows, int& x, int& y) {
    const int nx = (cols + kTX - 1) / kTX, n = nx * ((rows + kTY - 1) / kTY);
    if (b >= n) {
        b -= n;
        return false;
    }
    x = (b % nx) * kTX + threadIdx.x;
    y = (b / nx) * kTY + threadIdx.y;
    return true;
}

__global__ __launch_bounds__(256) void frame_level0_kernel(View<uint16_t> d16, View<uint8_t> rgb,
                                                           View<float> range, View<float> depth0,
                                                           View<uint8_t> gray0, float scale_inv,
                                                           float cutoff, ModelPyramid mp,
                                                           FrameReset fr) {
    int b = blockIdx.x, x, y;
    if (take_tile(b, d16.cols, d16.rows, x, y)) {
        if (x < d16.cols && y < d16.rows) {
            filter_px(d16, range, depth0, scale_inv, cutoff, x, y);
            intensity_px(rgb, gray0, x, y);
        }
        return;
    }
    // The model pyramid in tiles of 64 x 8 level-1 pixels (kept in LDS) and
    // the 32 x 4 level-2 pixels made of them, so level 0 is read once.
    __shared__ float4 lv[8][kModelTX], ln[8][kModelTX];
    const int nx1 = (mp.v[1].cols + kModelTX - 1) / kModelTX, ny1 = (mp.v[1].rows + 7) / 8;
    if (b < nx1 * ny1) {
        const int X1 = (b % nx1) * kModelTX, Y1 = (b / nx1) * 8;
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int ty = (int)threadIdx.y + 4 * r;
            const int x1 = X1 + (int)threadIdx.x, y1 = Y1 + ty;
            if (x1 < mp.v[1].cols && y1 < mp.v[1].rows) {
                float4 vo, no;
                resize_from(mp.v[0], mp.n[0], x1, y1, vo, no);
                mp.v[1].at(y1, x1) = vo;
                mp.n[1].at(y1, x1) = no;
                lv[ty][threadIdx.x] = vo;
                ln[ty][threadIdx.x] = no;
            }
        }
        __syncthreads();
        const int t = (int)(threadIdx.y * kTX + threadIdx.x);
        const int x2l = t & 31, y2l = t >> 5;
        const int x2 = X1 / 2 + x2l, y2 = Y1 / 2 + y2l;
        // level-2 pixel (x2, y2) reads level-1 pixels 2 x2 + {0,1}, 2 y2 + {0,1},
        // inside level 1 and inside this tile
        if (t < 128 && x2 < mp.v[2].cols && y2 < mp.v[2].rows) {
            float4 v[4], n[4], vo, no;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                v[k] = lv[2 * y2l + (k >> 1)][2 * x2l + (k & 1)];
                n[k] = ln[2 * y2l + (k >> 1)][2 * x2l + (k & 1)];
            }
            resize_px(v, n, vo, no);
            mp.v[2].at(y2, x2) = vo;
            mp.n[2].at(y2, x2) = no;
        }
        return;
    }
    b -= nx1 * ny1;
    // the frame's resets, in the workgroups after the tiles
    const int nz = fr.zmin.cols * fr.zmin.rows;
    const int i = b * 256 + (int)(threadIdx.y * kTX + threadIdx.x);
    if (i < nz) {
        const int zy = i / fr.zmin.cols, zx = i - zy * fr.zmin.cols;
        fr.zmin.at(zy, zx) = 100.f;
        fr.zmax.at(zy, zx) = 0.f;
    }
    if (b == 0 && threadIdx.x == 0 && threadIdx.y == 0) {
        *fr.render_tiles = 0u;
        *fr.rq_count = 0u;
        *fr.updated_voxels = 0u;
    }
}

// A row-major LDS window of `W` columns starting at (x0, y0).
template <class T, int W> struct LdsWin {
    const T* p;
    int x0, y0;
    SICP_DEV T operator()(int y, int x) const { return p[(y - y0) * W + (x - x0)]; }
};

// A workgroup owns a TX x TY tile of level 2. It computes the level-1 window
// its 5x5 taps reach (2 TX + 4 by 2 TY + 4 pixels, writing the level-1 pixels
// it owns) from the level-0 window those reach in turn (twice that plus 4),
// staged in LDS with every load in flight at once; then its level-2 pixels.
template <int TX, int TY>
__global__ __launch_bounds__(256) void pyrdown_levels_kernel(View<float> d0, View<uint8_t> g0,
                                                             View<float> d1, View<uint8_t> g1,
                                                             View<float> d2, View<uint8_t> g2) {
    constexpr int W1X = 2 * TX + 4, W1Y = 2 * TY + 4;  // level-1 window
    constexpr int W0X = 2 * W1X + 4, W0Y = 2 * W1Y + 4;  // level-0 window
    static_assert(TX * TY <= 256, "one level-2 pixel per thread");
    __shared__ float w0d[W0X * W0Y];
    __shared__ uint8_t w0g[W0X * W0Y];
    __shared__ float rd[W1X * W1Y];
    __shared__ uint8_t rg[W1X * W1Y];
    const int X0 = blockIdx.x * TX, Y0 = blockIdx.y * TY;  // level-2 tile origin
    const int rx0 = 2 * X0 - 2, ry0 = 2 * Y0 - 2;          // level-1 window origin
    const int wx0 = 2 * rx0 - 2, wy0 = 2 * ry0 - 2;        // level-0 window origin
    // Every load of the window is issued before the first LDS store (a loop
    // of load-then-store waits out one memory round trip per element);
    // entries outside the image are loaded from (0, 0) and not stored.
    constexpr int kWinIters = (W0X * W0Y + 255) / 256;
    float dv[kWinIters];
    uint8_t gv[kWinIters];
    bool in[kWinIters];
#pragma unroll
    for (int j = 0; j < kWinIters; ++j) {
        const int i = (int)threadIdx.x + 256 * j;
        const int wx = wx0 + i % W0X, wy = wy0 + i / W0X;
        in[j] = i < W0X * W0Y && wx >= 0 && wy >= 0 && wx < d0.cols && wy < d0.rows;
        const int cx = in[j] ? wx : 0, cy = in[j] ? wy : 0;
        dv[j] = d0.at(cy, cx);
        gv[j] = g0.at(cy, cx);
    }
#pragma unroll
    for (int j = 0; j < kWinIters; ++j) {
        const int i = (int)threadIdx.x + 256 * j;
        if (in[j]) {
            w0d[i] = dv[j];
            w0g[i] = gv[j];
        }
    }
    __syncthreads();
    // Level-1 pixels owned here: [2*X0, 2*X0 + 2 TX), the last tile column
    // (row) also the odd remainder up to the image edge.
    const int ox1 = blockIdx.x == gridDim.x - 1 ? d1.cols : 2 * X0 + 2 * TX;
    const int oy1 = blockIdx.y == gridDim.y - 1 ? d1.rows : 2 * Y0 + 2 * TY;
    for (int i = threadIdx.x; i < W1X * W1Y; i += 256) {
        const int rx = rx0 + i % W1X, ry = ry0 + i / W1X;
        if (rx < 0 || ry < 0 || rx >= d1.cols || ry >= d1.rows) continue;
        const float v =
            pyrdown_px<float>(LdsWin<float, W0X>{w0d, wx0, wy0}, d0.cols, d0.rows, rx, ry);
        const uint8_t u =
            pyrdown_px<uint8_t>(LdsWin<uint8_t, W0X>{w0g, wx0, wy0}, g0.cols, g0.rows, rx, ry);
        rd[i] = v;
        rg[i] = u;
        if (rx >= 2 * X0 && rx < ox1 && ry >= 2 * Y0 && ry < oy1) {
            d1.at(ry, rx) = v;
            g1.at(ry, rx) = u;
        }
    }
    __syncthreads();
    if (threadIdx.x < TX * TY) {
        const int x = X0 + (int)threadIdx.x % TX, y = Y0 + (int)threadIdx.x / TX;
        if (x < d2.cols && y < d2.rows) {
            d2.at(y, x) =
                pyrdown_px<float>(LdsWin<float, W1X>{rd, rx0, ry0}, d1.cols, d1.rows, x, y);
            g2.at(y, x) =
                pyrdown_px<uint8_t>(LdsWin<uint8_t, W1X>{rg, rx0, ry0}, g1.cols, g1.rows, x, y);
        }
    }
}
// The level-2 tile of pyrdown_levels_kernel: 8x4 gives 600 workgroups at
// 640x480, so the CUs hold two or three each and one's window loads overlap
// another's arithmetic (frame pyramid 36.6 us against 39.1 with 8x8 tiles,
// 38.0 with 10x8, 42.0 with 16x8, 38.3 with 4x4, 38.9 with 8x2; 4x8 as 8x4).
constexpr int kPyrTX = 8, kPyrTY = 4;

__global__ __launch_bounds__(256) void frame_maps_kernel(FramePyramid P, float cutoff) {
    int b = blockIdx.x, x = 0, y = 0, l = 0;
    for (; l < 3; ++l)
        if (take_tile(b, P.depth[l].cols, P.depth[l].rows, x, y)) break;
    if (l == 3) return;
    const View<float>& d = P.depth[l];
    if (x >= d.cols || y >= d.rows) return;
    const float fx = P.invfx[l], fy = P.invfy[l], cx = P.cx[l], cy = P.cy[l];
    const float4 c = vmap_px(d, x, y, fx, fy, cx, cy, cutoff);
    P.vmap[l].at(y, x) = c;
    if (x == d.cols - 1 || y == d.rows - 1) {
        const float nan = qnanf();
        P.nmap[l].at(y, x) = make_float4(nan, nan, nan, nan);
    } else {
        P.nmap[l].at(y, x) = nmap_of(c, vmap_px(d, x + 1, y, fx, fy, cx, cy, cutoff),
                                     vmap_px(d, x, y + 1, fx, fy, cx, cy, cutoff));
    }
    const int2 g = derivative_px(P.image[l], P.dx[l], P.dy[l], x, y);
    if (P.key[l].ptr) P.key[l].at(y, x) = rgb_key_px(P.image[l], g.x, g.y, c.x, x, y);
}

inline int tiles(int cols, int rows) { return (int)(div_up(cols, kTX) * div_up(rows, kTY)); }

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

int launch_rgb_keys(View<uint8_t> img, View<int16_t> dx, View<int16_t> dy, View<float4> vmap,
                    View<int16_t> key, hipStream_t s) {
    rgb_key_kernel<<<grid2d(img.cols, img.rows), dim3(kTX, kTY), 0, s>>>(img, dx, dy, vmap, key);
    SICP_LAUNCH_CHECK("rgb_key_kernel");
    return SICP_OK;
}

int launch_resize_map(View<float4> vsrc, View<float4> nsrc, View<float4> vdst, View<float4> ndst,
                      hipStream_t s) {
    resize_map_kernel<<<grid2d(vdst.cols, vdst.rows), dim3(kTX, kTY), 0, s>>>(vsrc, nsrc, vdst,
                                                                              ndst);
    SICP_LAUNCH_CHECK("resize_map_kernel");
    return SICP_OK;
}

int launch_frame_pyramid(View<uint16_t> depth_u16, View<uint8_t> rgb, View<float> range,
                         const FramePyramid& P, float scale_inv, float cutoff,
                         const ModelPyramid& last_model, const FrameReset& reset, hipStream_t s) {
    const ModelPyramid& mp = last_model;
    static_assert(kTX * kTY == 256, "reset workgroups of 256 threads");
    const int nr = (int)div_up(reset.zmin.cols * reset.zmin.rows, 256);
    const int nm = (int)(div_up(mp.v[1].cols, kModelTX) * div_up(mp.v[1].rows, 8));
    const int n0 = tiles(depth_u16.cols, depth_u16.rows) + nm + (nr > 0 ? nr : 1);
    frame_level0_kernel<<<n0, dim3(kTX, kTY), 0, s>>>(depth_u16, rgb, range, P.depth[0],
                                                      P.image[0], scale_inv, cutoff, mp, reset);
    SICP_LAUNCH_CHECK("frame_level0_kernel");
    const dim3 g2((unsigned)div_up(P.depth[2].cols, kPyrTX),
                  (unsigned)div_up(P.depth[2].rows, kPyrTY));
    pyrdown_levels_kernel<kPyrTX, kPyrTY><<<g2, 256, 0, s>>>(P.depth[0], P.image[0], P.depth[1],
                                                             P.image[1], P.depth[2], P.image[2]);
    SICP_LAUNCH_CHECK("pyrdown_levels_kernel");
    int n2 = 0;
    for (int l = 0; l < 3; ++l) n2 += tiles(P.depth[l].cols, P.depth[l].rows);
    frame_maps_kernel<<<n2, dim3(kTX, kTY), 0, s>>>(P, cutoff);
    SICP_LAUNCH_CHECK("frame_maps_kernel");
    return SICP_OK;
}

}  // namespace sicp
