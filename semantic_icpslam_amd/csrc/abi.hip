// abi.hip — the extern "C" entry points of include/sicp.h: argument
// validation, map ownership and the reference-shaped operations.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "se3_math.hpp"

#include "sicp_internal.hpp"

namespace sicp {

static thread_local std::string g_last_error = "";

void set_error(const std::string& msg) { g_last_error = msg; }

int hip_fail(hipError_t e, const char* what) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return SICP_ERR_HIP;
}

int invalid(const char* msg) {
    set_error(msg);
    return SICP_ERR_INVALID_ARGUMENT;
}

bool image_ok(const sicp_image& im, size_t elem, int cols, int rows) {
    if (!im.ptr || im.cols != cols || im.rows != rows || cols <= 0 || rows <= 0) return false;
    if (im.pitch < (size_t)cols * elem) return false;
    return true;
}

static bool pow2(uint32_t v) { return v && !(v & (v - 1)); }

template <class T> static int dmalloc(sicp_map* m, T** p, size_t count) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T) + 16);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc");
    m->owned.push_back(q);
    *p = (T*)q;
    return SICP_OK;
}

// IEEE half to float (exact), for downloads of the half form.
static float half_to_float(uint32_t h) {
    const uint32_t sign = (h & 0x8000u) << 16, e = (h >> 10) & 0x1fu, mnt = h & 0x3ffu;
    uint32_t bits;
    if (e == 0x1fu) {
        bits = sign | 0x7f800000u | (mnt << 13);
    } else if (e != 0) {
        bits = sign | ((e + 112u) << 23) | (mnt << 13);
    } else if (mnt == 0) {
        bits = sign;
    } else {  // subnormal: normalise
        int sh = 0;
        uint32_t m = mnt;
        while (!(m & 0x400u)) {
            m <<= 1;
            ++sh;
        }
        bits = sign | ((uint32_t)(113 - sh) << 23) | ((m & 0x3ffu) << 13);
    }
    float f;
    std::memcpy(&f, &bits, sizeof(f));
    return f;
}

int map_alloc(const sicp_map_config& cfg, sicp_map** out) {
    if (cfg.voxel_size <= 0.f || cfg.truncation <= 0.f) return invalid("voxel_size / truncation must be > 0");
    if (cfg.num_buckets <= 0 || cfg.num_blocks <= 0) return invalid("num_buckets / num_blocks must be > 0");
    if (cfg.num_blocks > (1 << 22)) return invalid("num_blocks above 4M");
    if (!pow2((uint32_t)cfg.request_capacity)) return invalid("request_capacity must be a power of two");
    if (cfg.max_cols <= 0 || cfg.max_rows <= 0) return invalid("max_cols / max_rows must be > 0");
    if ((size_t)cfg.max_cols * cfg.max_rows > (size_t)4096 * 1024) return invalid("image area above 4M pixels");
    if (cfg.max_weight < 1 || cfg.max_weight > 255) return invalid("max_weight outside 1..255");
    if (cfg.max_label_count < 1 || cfg.max_label_count > 65535) return invalid("max_label_count outside 1..65535");
    if (cfg.unify_labels != 0 && cfg.unify_labels != 1) return invalid("unify_labels must be 0 or 1");
    if (cfg.tsdf_half != 0 && cfg.tsdf_half != 1) return invalid("tsdf_half must be 0 or 1");
    sicp_map* m = new sicp_map();
    m->cfg = cfg;
    int rc = SICP_OK;
    const size_t nb = (size_t)cfg.num_buckets, nblk = (size_t)cfg.num_blocks;
    const size_t q = (size_t)cfg.request_capacity;
    const size_t px = (size_t)cfg.max_cols * cfg.max_rows;
    MapDev& d = m->dev;
    if ((rc = dmalloc(m, &d.table, 2 * nb)) || (rc = dmalloc(m, &d.excess, nblk)) ||
        (rc = dmalloc(m, &d.blocks, nblk)) ||
        (cfg.tsdf_half ? (rc = dmalloc(m, &d.geom_h, nblk * kBlockVoxels))
                       : (rc = dmalloc(m, &d.geom, nblk * kBlockVoxels))) ||
        (rc = dmalloc(m, &d.sem, nblk * kBlockVoxels)) || (rc = dmalloc(m, &m->visible, nblk)) ||
        (rc = dmalloc(m, &m->ctr, 1))) {
        map_free(m);
        return rc;
    }
    AllocScratch& a = m->alloc;
    if ((rc = dmalloc(m, &a.req_keys, q)) || (rc = dmalloc(m, &a.req_first, q)) ||
        (rc = dmalloc(m, &a.pix_mask, px)) || (rc = dmalloc(m, &a.pix_count, px)) ||
        (rc = dmalloc(m, &a.scan_partials, 4096 + 64)) ||
        (rc = dmalloc(m, &a.new_keys, q)) || (rc = dmalloc(m, &a.new_req_slot, q)) ||
        (rc = dmalloc(m, &a.bucket_leader, nb)) || (rc = dmalloc(m, &a.bucket_head, nb)) ||
        (rc = dmalloc(m, &a.pending_next, q)) || (rc = dmalloc(m, &a.vis_flags, nblk)) ||
        (rc = dmalloc(m, &a.fuse_counts, (size_t)kFuseBlocks * 4)) ||
        (rc = dmalloc(m, &d.rq, px)) || (rc = dmalloc(m, &d.rq_count, 1))) {
        map_free(m);
        return rc;
    }
    {
        int* one = nullptr;
        if ((rc = dmalloc(m, &one, 1))) {
            map_free(m);
            return rc;
        }
        const int v = 1;
        hipError_t e = hipMemcpy(one, &v, sizeof(v), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            map_free(m);
            return hip_fail(e, "hipMemcpy(one)");
        }
        d.one = one;
    }
    if (cfg.unify_labels) {
        // scratch of the neighbour label pass, here so that no later call allocates
        if ((rc = dmalloc(m, &m->unify_snap, nblk * kBlockVoxels)) ||
            (rc = dmalloc(m, &m->unify_mark, nblk))) {
            map_free(m);
            return rc;
        }
        hipError_t e = hipMemset(m->unify_mark, 0, nblk * sizeof(uint32_t));
        if (e != hipSuccess) {
            map_free(m);
            return hip_fail(e, "hipMemset(unify marks)");
        }
    }
    a.req_mask = (uint32_t)(q - 1);
    a.cap_pixels = (uint32_t)px;
    d.rq_cap = (uint32_t)px;
    d.nb = (uint32_t)nb;
    d.nb_mask = pow2((uint32_t)nb) ? (uint32_t)nb - 1 : 0u;
    d.num_blocks = cfg.num_blocks;
    d.voxel = cfg.voxel_size;
    d.voxel_inv = (float)(1.0 / (double)cfg.voxel_size);                 // DeviceMap.h:186
    d.trunc = cfg.truncation;
    d.step_scale = (float)(0.5 * (double)cfg.truncation * (double)d.voxel_inv);  // DeviceMap.h:189
    d.block_width = cfg.voxel_size * (float)kBlockSide;                  // DeviceMap.h:187
    d.dmin = cfg.depth_min;
    d.dmax = cfg.depth_max;
    d.max_weight = cfg.max_weight;
    d.max_count = cfg.max_label_count;
    d.half = cfg.tsdf_half;
    rc = launch_map_reset(m->dev, m->ctr, m->alloc, nullptr);
    if (rc == SICP_OK) {
        hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) rc = hip_fail(e, "map reset");
    }
    if (rc) {
        map_free(m);
        return rc;
    }
    *out = m;
    return SICP_OK;
}

void map_free(sicp_map* m) {
    if (!m) return;
    for (void* p : m->owned) (void)hipFree(p);
    delete m;
}

static PoseArg pose_arg(const sicp_pose_f* host, const sicp_pose_f* dev) {
    return make_pose_arg(host, dev);
}

}  // namespace sicp

using namespace sicp;

extern "C" {

const char* sicp_last_error(void) { return g_last_error.c_str(); }

const char* sicp_version(void) { return "semantic_icpslam_amd 0.1 (gfx950)"; }

int sicp_device_count(int* count) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        *count = 0;
        return hip_fail(e, "hipGetDeviceCount");
    }
    *count = n;
    return SICP_OK;
}

void sicp_map_config_default(sicp_map_config* cfg, float voxel_size, float truncation,
                             int32_t num_buckets, int32_t num_blocks) {
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->voxel_size = voxel_size;
    cfg->truncation = truncation;
    cfg->depth_min = 0.1f;   // DeviceMap::DepthMin (DeviceMap.h:168)
    cfg->depth_max = 3.0f;   // DeviceMap::DepthMax (DeviceMap.h:169)
    cfg->num_buckets = num_buckets;
    cfg->num_blocks = num_blocks;
    cfg->max_weight = 255;   // min(255, w + 1) (FuseMap.cu:505)
    cfg->max_label_count = 65535;
    cfg->unify_labels = 0;
    cfg->request_capacity = 1 << 20;
    cfg->max_cols = 640;
    cfg->max_rows = 480;
}

// ---------------------------------------------------------------- pyramid --
int sicp_filter_depth(sicp_image depth_u16, sicp_image raw_f32, sicp_image filtered_f32,
                      float depth_scale, float depth_cutoff, void* stream) {
    const int c = depth_u16.cols, r = depth_u16.rows;
    if (!image_ok(depth_u16, 2, c, r) || !image_ok(raw_f32, 4, c, r) ||
        !image_ok(filtered_f32, 4, c, r))
        return invalid("filter_depth: image shape / pitch mismatch");
    if (!(depth_scale > 0.f)) return invalid("filter_depth: depth_scale must be > 0");
    // FilterDepth passes 1.0 / depthScale rounded to float (Pyrdown.cu:117).
    return launch_filter_depth(view<uint16_t>(depth_u16), view<float>(raw_f32),
                               view<float>(filtered_f32), (float)(1.0 / (double)depth_scale),
                               depth_cutoff, (hipStream_t)stream);
}

int sicp_compute_vmap(sicp_image depth_f32, sicp_image vmap_f4, float fx, float fy, float cx,
                      float cy, float depth_cutoff, void* stream) {
    const int c = depth_f32.cols, r = depth_f32.rows;
    if (!image_ok(depth_f32, 4, c, r) || !image_ok(vmap_f4, 16, c, r))
        return invalid("compute_vmap: image shape / pitch mismatch");
    return launch_vmap(view<float>(depth_f32), view<float4>(vmap_f4), (float)(1.0 / (double)fx),
                       (float)(1.0 / (double)fy), cx, cy, depth_cutoff, (hipStream_t)stream);
}

int sicp_compute_nmap(sicp_image vmap_f4, sicp_image nmap_f4, void* stream) {
    const int c = vmap_f4.cols, r = vmap_f4.rows;
    if (!image_ok(vmap_f4, 16, c, r) || !image_ok(nmap_f4, 16, c, r))
        return invalid("compute_nmap: image shape / pitch mismatch");
    return launch_nmap(view<float4>(vmap_f4), view<float4>(nmap_f4), (hipStream_t)stream);
}

static bool half_of(const sicp_image& src, const sicp_image& dst) {
    return dst.cols == src.cols / 2 && dst.rows == src.rows / 2 && dst.cols > 0 && dst.rows > 0;
}

int sicp_pyrdown_gauss_f32(sicp_image src, sicp_image dst, void* stream) {
    if (!image_ok(src, 4, src.cols, src.rows) || !image_ok(dst, 4, dst.cols, dst.rows) ||
        !half_of(src, dst))
        return invalid("pyrdown_gauss_f32: dst must be src/2 (integer division)");
    return launch_pyrdown_f32(view<float>(src), view<float>(dst), (hipStream_t)stream);
}

int sicp_pyrdown_gauss_u8(sicp_image src, sicp_image dst, void* stream) {
    if (!image_ok(src, 1, src.cols, src.rows) || !image_ok(dst, 1, dst.cols, dst.rows) ||
        !half_of(src, dst))
        return invalid("pyrdown_gauss_u8: dst must be src/2 (integer division)");
    return launch_pyrdown_u8(view<uint8_t>(src), view<uint8_t>(dst), (hipStream_t)stream);
}

int sicp_image_to_intensity(sicp_image rgb_u8x3, sicp_image gray_u8, void* stream) {
    const int c = gray_u8.cols, r = gray_u8.rows;
    if (!image_ok(rgb_u8x3, 3, c, r) || !image_ok(gray_u8, 1, c, r))
        return invalid("image_to_intensity: image shape / pitch mismatch");
    return launch_intensity(view<uint8_t>(rgb_u8x3), view<uint8_t>(gray_u8), (hipStream_t)stream);
}

int sicp_compute_derivative_image(sicp_image gray_u8, sicp_image dx_i16, sicp_image dy_i16,
                                  void* stream) {
    const int c = gray_u8.cols, r = gray_u8.rows;
    if (!image_ok(gray_u8, 1, c, r) || !image_ok(dx_i16, 2, c, r) || !image_ok(dy_i16, 2, c, r))
        return invalid("compute_derivative_image: image shape / pitch mismatch");
    return launch_derivative(view<uint8_t>(gray_u8), view<int16_t>(dx_i16), view<int16_t>(dy_i16),
                             (hipStream_t)stream);
}

int sicp_resize_map(sicp_image vsrc, sicp_image nsrc, sicp_image vdst, sicp_image ndst,
                    void* stream) {
    const int c = vsrc.cols, r = vsrc.rows;
    if (!image_ok(vsrc, 16, c, r) || !image_ok(nsrc, 16, c, r) ||
        !image_ok(vdst, 16, c / 2, r / 2) || !image_ok(ndst, 16, c / 2, r / 2))
        return invalid("resize_map: dst must be src/2 (integer division)");
    return launch_resize_map(view<float4>(vsrc), view<float4>(nsrc), view<float4>(vdst),
                             view<float4>(ndst), (hipStream_t)stream);
}

// -------------------------------------------------------------------- ICP --
size_t sicp_icp_partials_size(void) { return sizeof(double) * kIcpMaxBlocks * kIcpSums + 256; }

}  // extern "C"

namespace sicp {
int icp_step(sicp_image vcurr, sicp_image vlast, sicp_image ncurr, sicp_image nlast,
                    const sicp_pose_f* curr, const sicp_pose_f* last, sicp_intrinsics K,
                    void* partials_dev, float* sums_dev, float* host_sums, double* A_host,
                    double* b_host, float* residual_host, bool k32, void* stream) {
    const int c = vcurr.cols, r = vcurr.rows;
    if (!image_ok(vcurr, 16, c, r) || !image_ok(vlast, 16, c, r) || !image_ok(ncurr, 16, c, r) ||
        !image_ok(nlast, 16, c, r))
        return invalid("icp_step: the four maps must share one shape");
    if (!curr || !last || !partials_dev || !sums_dev) return invalid("icp_step: null argument");
    if (((uintptr_t)partials_dev & 15) != 0)
        return invalid("icp_step: partials not 16-byte aligned");
    IcpScratch sc{(double*)partials_dev};
    hipStream_t s = (hipStream_t)stream;
    int rc = launch_icp_sums(view<float4>(vcurr), view<float4>(vlast), view<float4>(ncurr),
                             view<float4>(nlast), make_pose_arg(curr, nullptr),
                             make_pose_arg(last, nullptr), K, sc, sums_dev, k32, s);
    if (rc) return rc;
    if (host_sums || A_host || b_host || residual_host) {
        float h[kIcpSums];
        SICP_HIP(hipMemcpyAsync(h, sums_dev, sizeof(h), hipMemcpyDeviceToHost, s));
        SICP_HIP(hipStreamSynchronize(s));
        if (host_sums) std::memcpy(host_sums, h, sizeof(h));
        double A[36], b[6];
        create_matrix(h, A, b);
        if (A_host) std::memcpy(A_host, A, sizeof(A));
        if (b_host) std::memcpy(b_host, b, sizeof(b));
        if (residual_host) {
            residual_host[0] = h[27];
            residual_host[1] = h[28];
        }
    }
    return SICP_OK;
}
}  // namespace sicp

extern "C" {

int sicp_icp_step(sicp_image vcurr, sicp_image vlast, sicp_image ncurr, sicp_image nlast,
                  const sicp_pose_f* curr, const sicp_pose_f* last, sicp_intrinsics K,
                  void* partials_dev, float* sums_dev, float* host_sums, double* A_host,
                  double* b_host, float* residual_host, void* stream) {
    return icp_step(vcurr, vlast, ncurr, nlast, curr, last, K, partials_dev, sums_dev, host_sums,
                    A_host, b_host, residual_host, false, stream);
}

int sicp_icp_step_fp32(sicp_image vcurr, sicp_image vlast, sicp_image ncurr, sicp_image nlast,
                       const sicp_pose_f* curr, const sicp_pose_f* last, sicp_intrinsics K,
                       void* partials_dev, float* sums_dev, float* host_sums, double* A_host,
                       double* b_host, float* residual_host, void* stream) {
    return icp_step(vcurr, vlast, ncurr, nlast, curr, last, K, partials_dev, sums_dev, host_sums,
                    A_host, b_host, residual_host, true, stream);
}

int sicp_so3_step(sicp_image next_image, sicp_image last_image, sicp_image didx,
                  sicp_image didy, const sicp_pose_f* curr, const sicp_pose_f* last,
                  sicp_intrinsics K, void* partials_dev, float* sums_dev, float* host_sums,
                  double* A_host, double* b_host, float* residual_host, void* stream) {
    const int c = next_image.cols, r = next_image.rows;
    if (!image_ok(next_image, 1, c, r) || !image_ok(last_image, 1, c, r) ||
        !image_ok(didx, 2, c, r) || !image_ok(didy, 2, c, r))
        return invalid("so3_step: the four images must share one shape");
    if (!curr || !last || !partials_dev || !sums_dev) return invalid("so3_step: null argument");
    if (((uintptr_t)partials_dev & 15) != 0)
        return invalid("so3_step: partials not 16-byte aligned");
    IcpScratch sc{(double*)partials_dev};
    hipStream_t s = (hipStream_t)stream;
    int rc = launch_so3_sums(view<uint8_t>(next_image), view<uint8_t>(last_image),
                             view<int16_t>(didx), view<int16_t>(didy), make_pose_arg(curr, nullptr),
                             make_pose_arg(last, nullptr), K, sc, sums_dev, s);
    if (rc) return rc;
    if (host_sums || A_host || b_host || residual_host) {
        float h[11];
        SICP_HIP(hipMemcpyAsync(h, sums_dev, sizeof(h), hipMemcpyDeviceToHost, s));
        SICP_HIP(hipStreamSynchronize(s));
        if (host_sums) std::memcpy(host_sums, h, sizeof(h));
        // CreateMatrix<3,4> (Reduction.cu:6-16)
        double A[9], b[3];
        int shift = 0;
        for (int i = 0; i < 3; ++i)
            for (int j = i; j < 4; ++j) {
                const double value = (double)h[shift++];
                if (j == 3)
                    b[i] = value;
                else
                    A[j * 3 + i] = A[i * 3 + j] = value;
            }
        if (A_host) std::memcpy(A_host, A, sizeof(A));
        if (b_host) std::memcpy(b_host, b, sizeof(b));
        if (residual_host) {
            residual_host[0] = h[9];
            residual_host[1] = h[10];
        }
    }
    return SICP_OK;
}

// The photometric scratch: rows, records, counts, totals, keys.
static size_t rgb_scratch_bytes(size_t n) {
    const size_t pr = (size_t)kIcpSums * kIcpMaxBlocks * sizeof(double);
    const size_t ip = 2 * kIcpMaxBlocks * sizeof(uint32_t);
    return pr + n * sizeof(int2) + ip + 16 + n * sizeof(int16_t);
}

size_t sicp_rgb_scratch_size(int32_t cols, int32_t rows) {
    if (cols <= 0 || rows <= 0) return 0;
    return rgb_scratch_bytes((size_t)cols * rows);
}

int sicp_rgb_step(sicp_image next_image, sicp_image last_image, sicp_image next_vmap,
                  sicp_image last_vmap, sicp_image didx, sicp_image didy,
                  const sicp_pose_f* curr, const sicp_pose_f* last, sicp_intrinsics K,
                  void* scratch_dev, size_t scratch_bytes, float* sums_dev, float* host_sums,
                  int32_t* corresp_host, float* sigma_host, double* A_host, double* b_host,
                  float* residual_host, void* stream) {
    const int c = next_image.cols, r = next_image.rows;
    if (!image_ok(next_image, 1, c, r) || !image_ok(last_image, 1, c, r) ||
        !image_ok(next_vmap, 16, c, r) || !image_ok(last_vmap, 16, c, r) ||
        !image_ok(didx, 2, c, r) || !image_ok(didy, 2, c, r))
        return invalid("rgb_step: the six images must share one shape");
    if (!curr || !last || !scratch_dev || !sums_dev) return invalid("rgb_step: null argument");
    const size_t n = (size_t)c * r;
    if (scratch_bytes < rgb_scratch_bytes(n))
        return invalid("rgb_step: scratch smaller than sicp_rgb_scratch_size(cols, rows)");
    if (((uintptr_t)scratch_dev & 15) != 0) return invalid("rgb_step: scratch not 16-byte aligned");
    char* b = (char*)scratch_dev;
    const size_t pr = (size_t)kIcpSums * kIcpMaxBlocks * sizeof(double);
    const size_t ip = 2 * kIcpMaxBlocks * sizeof(uint32_t);
    RgbScratch rs;
    rs.partials = (double*)b;
    rs.rec = (int2*)(b + pr);
    rs.ipart = (uint32_t*)(b + pr + n * sizeof(int2));
    rs.totals = (int32_t*)(b + pr + n * sizeof(int2) + ip);
    int16_t* key = (int16_t*)(b + pr + n * sizeof(int2) + ip + 16);
    hipStream_t s = (hipStream_t)stream;
    int rc = launch_rgb_sums(view<uint8_t>(next_image), view<uint8_t>(last_image),
                             view<float4>(next_vmap), view<float4>(last_vmap),
                             view<int16_t>(didx), view<int16_t>(didy),
                             make_pose_arg(curr, nullptr), make_pose_arg(last, nullptr), K, key,
                             rs, sums_dev, s);
    if (rc) return rc;
    if (host_sums || corresp_host || sigma_host || A_host || b_host || residual_host) {
        float h[kIcpSums];
        int32_t t[3];
        SICP_HIP(hipMemcpyAsync(h, sums_dev, sizeof(h), hipMemcpyDeviceToHost, s));
        SICP_HIP(hipMemcpyAsync(t, rs.totals, sizeof(t), hipMemcpyDeviceToHost, s));
        SICP_HIP(hipStreamSynchronize(s));
        if (host_sums) std::memcpy(host_sums, h, sizeof(h));
        if (corresp_host) {
            corresp_host[0] = t[0];
            corresp_host[1] = t[1];
        }
        if (sigma_host) std::memcpy(sigma_host, &t[2], sizeof(float));
        double A[36], bv[6];
        create_matrix(h, A, bv);
        if (A_host) std::memcpy(A_host, A, sizeof(A));
        if (b_host) std::memcpy(b_host, bv, sizeof(bv));
        if (residual_host) {
            residual_host[0] = h[27];
            residual_host[1] = h[28];
        }
    }
    return SICP_OK;
}

int sicp_se3_update_host(const double* A, const double* b, double* delta,
                         const double* last_pose, double* next_pose, double* x_out) {
    if (!A || !b || !delta || !last_pose || !next_pose) return invalid("se3_update_host: null argument");
    se3_update_host(A, b, delta, last_pose, next_pose, x_out);
    return SICP_OK;
}

// -------------------------------------------------------------------- map --
int sicp_map_create(const sicp_map_config* cfg, sicp_map** out) {
    if (!cfg || !out) return invalid("map_create: null argument");
    return map_alloc(*cfg, out);
}

int sicp_map_destroy(sicp_map* map) {
    map_free(map);
    return SICP_OK;
}

int sicp_map_get_config(const sicp_map* map, sicp_map_config* cfg) {
    if (!map || !cfg) return invalid("map_get_config: null argument");
    *cfg = map->cfg;
    return SICP_OK;
}

int sicp_map_reset(sicp_map* map, void* stream) {
    if (!map) return invalid("map_reset: null map");
    return launch_map_reset(map->dev, map->ctr, map->alloc, (hipStream_t)stream);
}

static int read_u32(const uint32_t* dev, uint32_t* host, hipStream_t s) {
    SICP_HIP(hipMemcpyAsync(host, dev, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    SICP_HIP(hipStreamSynchronize(s));
    return SICP_OK;
}

size_t sicp_raycast_backward_scratch_size(void) { return raycast_adjoint_scratch_bytes(); }

int sicp_raycast_backward(sicp_map* map, sicp_image zmin, sicp_image zmax,
                          const sicp_pose_f* pose, float invfx, float invfy, float cx, float cy,
                          int32_t cols, int32_t rows, const int32_t* pixels_dev,
                          int32_t n_pixels, const float* grad_vertex_dev,
                          const float* grad_normal_dev, float* grad_sdf_dev,
                          float* grad_pose_px_dev, double* grad_pose_dev, void* scratch_dev,
                          size_t scratch_bytes, void* stream) {
    if (!map || !pose || !grad_vertex_dev || !grad_pose_dev || !scratch_dev ||
        (n_pixels > 0 && !pixels_dev))
        return invalid("raycast_backward: null argument");
    if (n_pixels < 0) return invalid("raycast_backward: negative pixel count");
    if (cols <= 0 || rows <= 0) return invalid("raycast_backward: empty image");
    const int zc = (cols + 7) / 8, zr = (rows + 7) / 8;
    if (!image_ok(zmin, 4, zc, zr) || !image_ok(zmax, 4, zc, zr))
        return invalid("raycast_backward: zRange images must be ceil(cols/8) x ceil(rows/8)");
    if (scratch_bytes < raycast_adjoint_scratch_bytes())
        return invalid("raycast_backward: scratch smaller than sicp_raycast_backward_scratch_size()");
    if (((uintptr_t)scratch_dev & 7) != 0) return invalid("raycast_backward: scratch not 8-byte aligned");
    return launch_raycast_adjoint(map->dev, view<float>(zmin), view<float>(zmax),
                                  make_pose_arg(pose, nullptr), invfx, invfy, cx, cy, cols, rows,
                                  pixels_dev, n_pixels, grad_vertex_dev, grad_normal_dev,
                                  grad_sdf_dev, grad_pose_px_dev, (double*)scratch_dev,
                                  grad_pose_dev, (hipStream_t)stream);
}

int sicp_unify_neighbor_labels(sicp_map* map, void* stream) {
    if (!map) return invalid("unify_neighbor_labels: null map");
    return launch_unify(map, nullptr, (hipStream_t)stream);
}

int sicp_fuse_map_color(sicp_map* map, sicp_image depth_f32, sicp_image rgb_u8x3,
                        sicp_image nmap_f4, sicp_image mask_u8, int num_detection,
                        const sicp_pose_f* pose_host, const sicp_pose_f* pose_dev,
                        sicp_intrinsics K, uint32_t* visible_host, void* stream) {
    if (!map) return invalid("fuse_map_color: null map");
    const int c = depth_f32.cols, r = depth_f32.rows;
    if (!image_ok(depth_f32, 4, c, r) || !image_ok(rgb_u8x3, 3, c, r))
        return invalid("fuse_map_color: depth / rgb shape mismatch");
    if (num_detection > 0 && (!image_ok(nmap_f4, 16, c, r) || !image_ok(mask_u8, 1, c, r)))
        return invalid("fuse_map_color: nmap / mask required when num_detection > 0");
    if (!pose_host && !pose_dev) return invalid("fuse_map_color: no pose");
    if ((size_t)c * r > (size_t)map->cfg.max_cols * map->cfg.max_rows)
        return invalid("fuse_map_color: image larger than the map's max_cols x max_rows");
    hipStream_t s = (hipStream_t)stream;
    PoseArg pa = pose_arg(pose_host, pose_dev);
    // FuseMapColor checks visibility against the map's own depth range
    // (FuseMap.cu:1080-1081); the allocation's last launch flags the blocks.
    int rc = launch_allocate(map->dev, map->ctr, map->alloc, view<float>(depth_f32), pa, K,
                             VisFrustum{c, r, map->cfg.depth_max, map->cfg.depth_min}, nullptr, s);
    if (rc) return rc;
    rc = launch_visible_compact(map->dev, map->ctr, map->alloc, map->visible, nullptr, s);
    if (rc) return rc;
    sicp_image nm = nmap_f4, mk = mask_u8;
    if (num_detection <= 0) {
        nm = depth_f32;  // unused by the kernel when num_detection <= 0
        mk = depth_f32;
    }
    rc = launch_fuse(map->dev, map->ctr, map->visible, view<float>(depth_f32),
                     view<uint8_t>(rgb_u8x3), view<float4>(nm), view<uint8_t>(mk), num_detection,
                     pa, K, map->alloc.fuse_counts, nullptr, s);
    if (rc) return rc;
    rc = launch_fuse_count(map->ctr, map->alloc.fuse_counts, nullptr, s);
    if (rc) return rc;
    if (map->cfg.unify_labels && num_detection > 0 && mask_u8.ptr) {
        rc = launch_unify(map, nullptr, s);
        if (rc) return rc;
    }
    if (visible_host) {
        rc = read_u32(&map->ctr->visible, visible_host, s);
        if (rc) return rc;
        if (*visible_host == 0) {
            set_error("no visible blocks found");
            return SICP_ERR_NO_VISIBLE;
        }
    }
    return SICP_OK;
}

int sicp_check_block_visibility(sicp_map* map, const sicp_pose_f* pose_host,
                                const sicp_pose_f* pose_dev, int cols, int rows,
                                sicp_intrinsics K, float depth_max, float depth_min,
                                uint32_t* visible_host, void* stream) {
    if (!map) return invalid("check_block_visibility: null map");
    if (!pose_host && !pose_dev) return invalid("check_block_visibility: no pose");
    hipStream_t s = (hipStream_t)stream;
    int rc = launch_visibility(map->dev, map->ctr, map->alloc, map->visible,
                               pose_arg(pose_host, pose_dev), K, cols, rows, depth_max, depth_min,
                               nullptr, s);
    if (rc) return rc;
    if (visible_host) {
        rc = read_u32(&map->ctr->visible, visible_host, s);
        if (rc) return rc;
        if (*visible_host == 0) {
            set_error("no visible blocks found");
            return SICP_ERR_NO_VISIBLE;
        }
    }
    return SICP_OK;
}

int sicp_create_rendering_blocks(sicp_map* map, sicp_image zmin, sicp_image zmax,
                                 float depth_max, float depth_min,
                                 const sicp_pose_f* pose_host, const sicp_pose_f* pose_dev,
                                 sicp_intrinsics K, int* nonempty_host, void* stream) {
    if (!map) return invalid("create_rendering_blocks: null map");
    if (!image_ok(zmin, 4, zmin.cols, zmin.rows) || !image_ok(zmax, 4, zmin.cols, zmin.rows))
        return invalid("create_rendering_blocks: zmin / zmax shape mismatch");
    if (!pose_host && !pose_dev) return invalid("create_rendering_blocks: no pose");
    hipStream_t s = (hipStream_t)stream;
    int rc = launch_render_prepass(map->dev, map->ctr, map->visible, view<float>(zmin),
                                   view<float>(zmax), pose_arg(pose_host, pose_dev), K, depth_max,
                                   depth_min, nullptr, s);
    if (rc) return rc;
    if (nonempty_host) {
        uint32_t n = 0;
        rc = read_u32(&map->ctr->render_tiles, &n, s);
        if (rc) return rc;
        *nonempty_host = n > 0 ? 1 : 0;
        if (!n) {
            set_error("nothing to render");
            return SICP_ERR_NOTHING_TO_RENDER;
        }
    }
    return SICP_OK;
}

static int raycast_common(sicp_map* map, sicp_image vmap_f4, sicp_image nmap_f4,
                          sicp_image zmin, sicp_image zmax, const sicp_pose_f* pose_host,
                          const sicp_pose_f* pose_dev, float invfx, float invfy, float cx,
                          float cy, void* stream) {
    if (!map) return invalid("raycast: null map");
    const int c = vmap_f4.cols, r = vmap_f4.rows;
    if (!image_ok(vmap_f4, 16, c, r) || !image_ok(nmap_f4, 16, c, r))
        return invalid("raycast: vmap / nmap shape mismatch");
    if (!image_ok(zmin, 4, (c + 7) / 8, (r + 7) / 8) || !image_ok(zmax, 4, (c + 7) / 8, (r + 7) / 8))
        return invalid("raycast: zmin / zmax must be ceil(W/8) x ceil(H/8)");
    if (!pose_host && !pose_dev) return invalid("raycast: no pose");
    return launch_raycast(map->dev, view<float4>(vmap_f4), view<float4>(nmap_f4),
                          view<float>(zmin), view<float>(zmax), pose_arg(pose_host, pose_dev),
                          invfx, invfy, cx, cy, nullptr, (hipStream_t)stream);
}

int sicp_raycast(sicp_map* map, sicp_image vmap_f4, sicp_image nmap_f4, sicp_image zmin,
                 sicp_image zmax, const sicp_pose_f* pose_host, const sicp_pose_f* pose_dev,
                 sicp_intrinsics K, void* stream) {
    // Mapping::RayTrace passes 1.0 / fx rounded to float (Mapping.cc:158-159).
    return raycast_common(map, vmap_f4, nmap_f4, zmin, zmax, pose_host, pose_dev,
                          (float)(1.0 / (double)K.fx), (float)(1.0 / (double)K.fy), K.cx, K.cy,
                          stream);
}

int sicp_raycast_inv(sicp_map* map, sicp_image vmap_f4, sicp_image nmap_f4, sicp_image zmin,
                     sicp_image zmax, const sicp_pose_f* pose_host, const sicp_pose_f* pose_dev,
                     float invfx, float invfy, float cx, float cy, void* stream) {
    return raycast_common(map, vmap_f4, nmap_f4, zmin, zmax, pose_host, pose_dev, invfx, invfy,
                          cx, cy, stream);
}

int sicp_map_visible_list(sicp_map* map, sicp_block** list_dev, uint32_t* capacity) {
    if (!map || !list_dev || !capacity) return invalid("map_visible_list: null argument");
    *list_dev = map->visible;
    *capacity = (uint32_t)map->cfg.num_blocks;
    return SICP_OK;
}

int sicp_map_get_stats(sicp_map* map, sicp_map_stats* out, void* stream) {
    if (!map || !out) return invalid("map_get_stats: null argument");
    MapCounters c;
    hipStream_t s = (hipStream_t)stream;
    SICP_HIP(hipMemcpyAsync(&c, map->ctr, sizeof(c), hipMemcpyDeviceToHost, s));
    SICP_HIP(hipStreamSynchronize(s));
    out->allocated_blocks = c.heap_used;
    out->visible_blocks = c.visible;
    out->new_blocks = c.new_blocks;
    out->dropped_heap = c.dropped_heap;
    out->dropped_requests = c.dropped_requests;
    out->dropped_range = c.dropped_range;
    out->updated_voxels = c.updated_voxels;
    out->render_tiles = c.render_tiles;
    uint32_t handover = 0;
    SICP_HIP(hipMemcpy(&handover, map->dev.rq_count, sizeof(handover), hipMemcpyDeviceToHost));
    out->ray_handover = handover;
    out->updated_total = c.updated_total;
    out->visible_total = c.visible_total;
    out->fusions = c.fusions;
    out->heap_scanned_total = c.heap_total;
    return SICP_OK;
}

int sicp_map_download(sicp_map* map, sicp_hash_entry* table, sicp_hash_entry* excess,
                      sicp_block* blocks, sicp_voxel_geom* geom, sicp_voxel_sem* sem,
                      uint32_t n_blocks_limit) {
    if (!map) return invalid("map_download: null map");
    SICP_HIP(hipDeviceSynchronize());
    const size_t nb = (size_t)map->cfg.num_buckets;
    size_t nblk = (size_t)map->cfg.num_blocks;
    if (n_blocks_limit && n_blocks_limit < nblk) nblk = n_blocks_limit;
    if (table)
        SICP_HIP(hipMemcpy(table, map->dev.table, 2 * nb * sizeof(sicp_hash_entry),
                           hipMemcpyDeviceToHost));
    if (excess)
        SICP_HIP(hipMemcpy(excess, map->dev.excess, nblk * sizeof(sicp_hash_entry),
                           hipMemcpyDeviceToHost));
    if (blocks)
        SICP_HIP(hipMemcpy(blocks, map->dev.blocks, nblk * sizeof(sicp_block),
                           hipMemcpyDeviceToHost));
    if (geom && !map->dev.half)
        SICP_HIP(hipMemcpy(geom, map->dev.geom, nblk * kBlockVoxels * sizeof(sicp_voxel_geom),
                           hipMemcpyDeviceToHost));
    if (sem)
        SICP_HIP(hipMemcpy(sem, map->dev.sem, nblk * kBlockVoxels * sizeof(sicp_voxel_sem),
                           hipMemcpyDeviceToHost));
    if (geom && map->dev.half) {
        // the half form widened to the FP32 record; the colour from the
        // semantic record's spare bytes (cleared in the returned records)
        const size_t n = nblk * kBlockVoxels;
        std::vector<uint32_t> w(n);
        std::vector<sicp_voxel_sem> s(n);
        SICP_HIP(hipMemcpy(w.data(), map->dev.geom_h, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
        SICP_HIP(hipMemcpy(s.data(), map->dev.sem, n * sizeof(sicp_voxel_sem),
                           hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) {
            geom[i].sdf = half_to_float(w[i] & 0xffffu);
            geom[i].weight = (uint8_t)((w[i] >> 16) & 0xffu);
            geom[i].r = s[i].pad[0];
            geom[i].g = s[i].pad[1];
            geom[i].b = s[i].pad[2];
        }
        if (sem)
            for (size_t i = 0; i < n; ++i) sem[i].pad[0] = sem[i].pad[1] = sem[i].pad[2] = 0;
    }
    return SICP_OK;
}

int sicp_map_download_visible(sicp_map* map, sicp_block* visible, uint32_t cap,
                              uint32_t* count) {
    if (!map) return invalid("map_download_visible: null map");
    SICP_HIP(hipDeviceSynchronize());
    MapCounters c;
    SICP_HIP(hipMemcpy(&c, map->ctr, sizeof(c), hipMemcpyDeviceToHost));
    if (count) *count = c.visible;
    uint32_t n = c.visible < cap ? c.visible : cap;
    if (visible && n)
        SICP_HIP(hipMemcpy(visible, map->visible, n * sizeof(sicp_block), hipMemcpyDeviceToHost));
    return SICP_OK;
}

// ---- snapshot -------------------------------------------------------------
// System::WriteMapToDisk / ReadMapFromDisk (Core/System.cc:327-403) dump
// every map buffer raw. The snapshot here is versioned and holds only what
// the map uses: the table, the first heap_used blocks' overflow entries,
// records and voxels, the visible list and the counters.
}  // extern "C"

namespace {
struct SnapshotHeader {
    char magic[8];
    uint32_t version, header_bytes;
    sicp_map_config cfg;
    MapCounters ctr;
    uint64_t table_entries, blocks, visible;
};
constexpr char kSnapMagic[8] = {'S', 'I', 'C', 'P', 'M', 'A', 'P', '\0'};
constexpr uint32_t kSnapVersion = 2;  // 2: MapCounters with heap_total

struct File {
    FILE* f;
    ~File() {
        if (f) fclose(f);
    }
};

template <class T> int put(FILE* f, const T* dev, size_t n, const char* what) {
    if (!n) return SICP_OK;
    std::vector<T> h(n);
    SICP_HIP(hipMemcpy(h.data(), dev, n * sizeof(T), hipMemcpyDeviceToHost));
    if (fwrite(h.data(), sizeof(T), n, f) != n) return invalid(what);
    return SICP_OK;
}
template <class T> int read_all(FILE* f, std::vector<T>& h, size_t n) {
    h.resize(n);
    if (n && fread(h.data(), sizeof(T), n, f) != n) return invalid("map_load: truncated file");
    return SICP_OK;
}
template <class T> int upload(T* dev, const std::vector<T>& h) {
    if (!h.empty())
        SICP_HIP(hipMemcpy(dev, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return SICP_OK;
}

// The block key and bucket of the device hash (map_common.hpp pack_key /
// bucket_of), on the host.
uint64_t host_pack_key(int x, int y, int z) {
    const uint64_t msk = (1ull << 21) - 1;
    return (((uint64_t)(uint32_t)x & msk) << 42) | (((uint64_t)(uint32_t)y & msk) << 21) |
           ((uint64_t)(uint32_t)z & msk);
}
uint32_t host_bucket(const sicp_block& b, uint32_t nb) {
    const uint32_t h = ((uint32_t)b.x * 73856093u) ^ ((uint32_t)b.y * 19349669u) ^
                       ((uint32_t)b.z * 83492791u);
    return h % nb;
}

// Everything a loaded map's kernels index by: every entry's slot and chain
// link inside the used heap, each overflow chain acyclic and owned by one
// bucket, each entry's key that of the block record at its slot, the block
// records at their own slots, the visible list a subset of the records, and
// the counters equal to the section sizes.
int validate_snapshot(const SnapshotHeader& h, const std::vector<sicp_hash_entry>& table,
                      const std::vector<sicp_hash_entry>& excess,
                      const std::vector<sicp_block>& blocks,
                      const std::vector<sicp_block>& visible) {
    const int64_t nb = (int64_t)h.blocks;
    if (h.ctr.heap_used != h.blocks || h.ctr.visible != h.visible)
        return invalid("map_load: counters disagree with the snapshot's sections");
    auto in_heap = [&](int32_t v) { return v >= -1 && (int64_t)v < nb; };
    const uint32_t nbk = (uint32_t)(table.size() / 2);
    // an entry's key is its block record's, and the block hashes to `bucket`
    auto key_ok = [&](const sicp_hash_entry& e, size_t bucket) {
        if (e.slot < 0) return true;
        const sicp_block& b = blocks[(size_t)e.slot];
        return e.key == host_pack_key(b.x, b.y, b.z) && host_bucket(b, nbk) == bucket;
    };
    for (int64_t i = 0; i < nb; ++i)
        if (blocks[(size_t)i].slot != (int32_t)i)
            return invalid("map_load: a block record is not at its own heap slot");
    for (const sicp_hash_entry& e : table)
        if (!in_heap(e.slot) || !in_heap(e.next))
            return invalid("map_load: a table entry points outside the heap");
    for (const sicp_hash_entry& e : excess)
        if (!in_heap(e.slot) || !in_heap(e.next))
            return invalid("map_load: an overflow entry points outside the heap");
    for (size_t i = 0; i < table.size(); ++i)
        if (!key_ok(table[i], i / 2)) return invalid("map_load: a table entry's key is wrong");
    // every chain from a bucket's second entry: each overflow entry visited
    // at most once over all chains (no cycle, no shared tail)
    std::vector<uint8_t> seen((size_t)nb, 0);
    for (size_t b = 1; b < table.size(); b += 2) {
        for (int32_t n = table[b].next; n >= 0; n = excess[(size_t)n].next) {
            if (seen[(size_t)n]) return invalid("map_load: an overflow chain loops or is shared");
            seen[(size_t)n] = 1;
            if (!key_ok(excess[(size_t)n], b / 2))
                return invalid("map_load: an overflow entry's key is wrong");
        }
    }
    for (const sicp_block& v : visible) {
        if (v.slot < 0 || (int64_t)v.slot >= nb)
            return invalid("map_load: a visible block is outside the heap");
        const sicp_block& b = blocks[(size_t)v.slot];
        if (b.x != v.x || b.y != v.y || b.z != v.z)
            return invalid("map_load: a visible block differs from its record");
    }
    return SICP_OK;
}
}  // namespace

extern "C" {

int sicp_map_save(sicp_map* map, const char* path) {
    if (!map || !path) return invalid("map_save: null argument");
    SICP_HIP(hipDeviceSynchronize());
    SnapshotHeader h{};
    std::memcpy(h.magic, kSnapMagic, sizeof(h.magic));
    h.version = kSnapVersion;
    h.header_bytes = sizeof(SnapshotHeader);
    h.cfg = map->cfg;
    SICP_HIP(hipMemcpy(&h.ctr, map->ctr, sizeof(h.ctr), hipMemcpyDeviceToHost));
    h.table_entries = 2 * (uint64_t)map->cfg.num_buckets;
    h.blocks = h.ctr.heap_used;
    h.visible = h.ctr.visible;
    File file{fopen(path, "wb")};
    if (!file.f) return invalid("map_save: cannot open the file for writing");
    if (fwrite(&h, sizeof(h), 1, file.f) != 1) return invalid("map_save: write failed");
    int rc;
    const size_t nb = h.blocks;
    if ((rc = put(file.f, map->dev.table, h.table_entries, "map_save: write failed")) ||
        (rc = put(file.f, map->dev.excess, nb, "map_save: write failed")) ||
        (rc = put(file.f, map->dev.blocks, nb, "map_save: write failed")) ||
        (map->dev.half
             ? (rc = put(file.f, map->dev.geom_h, nb * kBlockVoxels, "map_save: write failed"))
             : (rc = put(file.f, map->dev.geom, nb * kBlockVoxels, "map_save: write failed"))) ||
        (rc = put(file.f, map->dev.sem, nb * kBlockVoxels, "map_save: write failed")) ||
        (rc = put(file.f, map->visible, (size_t)h.visible, "map_save: write failed")))
        return rc;
    return SICP_OK;
}

}  // extern "C"

namespace {
int map_load_file(sicp_map* map, FILE* f) {
    SnapshotHeader h;
    if (fread(&h, sizeof(h), 1, f) != 1) return invalid("map_load: truncated header");
    if (std::memcmp(h.magic, kSnapMagic, sizeof(h.magic)) != 0)
        return invalid("map_load: not a map snapshot");
    if (h.version != kSnapVersion || h.header_bytes != sizeof(SnapshotHeader))
        return invalid("map_load: unsupported snapshot version");
    const sicp_map_config& a = map->cfg;
    const sicp_map_config& b = h.cfg;
    if (a.num_buckets != b.num_buckets || a.voxel_size != b.voxel_size ||
        a.truncation != b.truncation || a.depth_min != b.depth_min ||
        a.depth_max != b.depth_max || a.max_weight != b.max_weight ||
        a.max_label_count != b.max_label_count || a.tsdf_half != b.tsdf_half)
        return invalid("map_load: the snapshot was written by a map of another configuration");
    if (h.blocks > (uint64_t)a.num_blocks || h.visible > h.blocks ||
        h.table_entries != 2 * (uint64_t)a.num_buckets)
        return invalid("map_load: the snapshot does not fit this map");
    // the whole file into host memory and checked before any upload
    const size_t nb = h.blocks;
    std::vector<sicp_hash_entry> table, excess;
    std::vector<sicp_block> blocks, visible;
    std::vector<uint32_t> geom_h;
    std::vector<sicp_voxel_geom> geom;
    std::vector<sicp_voxel_sem> sem;
    int rc;
    if ((rc = read_all(f, table, h.table_entries)) || (rc = read_all(f, excess, nb)) ||
        (rc = read_all(f, blocks, nb)) ||
        (map->dev.half ? (rc = read_all(f, geom_h, nb * kBlockVoxels))
                       : (rc = read_all(f, geom, nb * kBlockVoxels))) ||
        (rc = read_all(f, sem, nb * kBlockVoxels)) || (rc = read_all(f, visible, h.visible)))
        return rc;
    if (fgetc(f) != EOF) return invalid("map_load: trailing bytes after the snapshot");
    if ((rc = validate_snapshot(h, table, excess, blocks, visible))) return rc;
    if ((rc = upload(map->dev.table, table)) || (rc = upload(map->dev.excess, excess)) ||
        (rc = upload(map->dev.blocks, blocks)) ||
        (map->dev.half ? (rc = upload(map->dev.geom_h, geom_h))
                       : (rc = upload(map->dev.geom, geom))) ||
        (rc = upload(map->dev.sem, sem)) || (rc = upload(map->visible, visible)))
        return rc;
    SICP_HIP(hipMemcpy(map->ctr, &h.ctr, sizeof(h.ctr), hipMemcpyHostToDevice));
    return SICP_OK;
}
}  // namespace

extern "C" {

// The map is reset first and left reset when any check or copy fails.
int sicp_map_load(sicp_map* map, const char* path) {
    if (!map || !path) return invalid("map_load: null argument");
    File file{fopen(path, "rb")};
    if (!file.f) return invalid("map_load: cannot open the file");
    int rc = launch_map_reset(map->dev, map->ctr, map->alloc, nullptr);
    if (rc) return rc;
    SICP_HIP(hipDeviceSynchronize());
    rc = map_load_file(map, file.f);
    if (rc) {
        const std::string msg = sicp_last_error();
        if (launch_map_reset(map->dev, map->ctr, map->alloc, nullptr) == SICP_OK)
            (void)hipDeviceSynchronize();
        set_error(msg);  // keep the load's own message
        return rc;
    }
    return SICP_OK;
}

}  // extern "C"
