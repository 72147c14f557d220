// raycast.hip — rendering-block depth-range prepass and the TSDF raycast
// for gfx950.
//
// Prepass (CreateRenderingBlocks, Mapping/RenderScene.cu:210-289): one thread
// per visible block projects its eight corners, clips the 1/8-resolution box
// and applies the reference's tile coverage directly with integer
// atomicMin / atomicMax on the float bits (all depths are positive, where the
// IEEE bit order is the numeric order) — no rendering-block list, no
// host round trip.
// Raycast (Rendering::operator(), RenderScene.cu:380-488): one thread per
// pixel, 8x8-pixel workgroups (one wavefront) so neighbouring rays march
// through the same voxel blocks; each ray keeps the reference's one-entry
// block cache.
#include "map_common.hpp"

namespace sicp {

namespace {

struct PrepassArgs {
    float fx, fy, cx, cy;
    float dmax, dmin;
    int zcols, zrows;
};

__global__ __launch_bounds__(256) void zrange_init_kernel(View<float> zmin, View<float> zmax,
                                                          MapCounters* ctr) {
    const int n = zmin.cols * zmin.rows;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int y = i / zmin.cols, x = i - y * zmin.cols;
        zmin.at(y, x) = 100.f;
        zmax.at(y, x) = 0.f;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) ctr->render_tiles = 0u;
}

// Projection::projectBlock (RenderScene.cu:54-112).
SICP_DEV bool project_block(const sicp_block& b, const PoseF& pose, const PrepassArgs& g,
                            float voxel, int& ulx, int& uly, int& lrx, int& lry, float& z0,
                            float& z1) {
    ulx = g.zcols;
    uly = g.zrows;
    lrx = -1;
    lry = -1;
    z0 = g.dmax;
    z1 = g.dmin;
    for (int corner = 0; corner < 8; ++corner) {
        int tx = b.x + ((corner & 1) ? 1 : 0);
        int ty = b.y + ((corner & 2) ? 1 : 0);
        int tz = b.z + ((corner & 4) ? 1 : 0);
        v3 p{(float)(tx * kBlockSide) * voxel, (float)(ty * kBlockSide) * voxel,
             (float)(tz * kBlockSide) * voxel};
        p = mv(pose.Rinv, sub(p, pose.t));
        if ((double)p.z < 2e-1) continue;
        float px = (g.fx * p.x / p.z + g.cx) / 8.0f;
        float py = (g.fy * p.y / p.z + g.cy) / 8.0f;
        float fx_ = floorf(px), cx_ = ceilf(px), fy_ = floorf(py), cy_ = ceilf(py);
        if ((float)ulx > fx_) ulx = (int16_t)(int)fx_;
        if ((float)lrx < cx_) lrx = (int16_t)(int)cx_;
        if ((float)uly > fy_) uly = (int16_t)(int)fy_;
        if ((float)lry < cy_) lry = (int16_t)(int)cy_;
        if (z0 > p.z) z0 = p.z;
        if (z1 < p.z) z1 = p.z;
    }
    if (ulx < 0) ulx = 0;
    if (uly < 0) uly = 0;
    if (lrx >= g.zcols) lrx = g.zcols - 1;
    if (lry >= g.zrows) lry = g.zrows - 1;
    if (ulx > lrx) return false;
    if (uly > lry) return false;
    if (z0 < g.dmin) z0 = g.dmin;
    if (z1 < g.dmin) return false;
    return true;
}

// createRenderingBlockList (:114-132) + fillBlocks (:167-190): the tile loop
// runs x < ny and does not offset the tiles' lower-right corner, so a box
// [ul, lr] covers columns ul.x .. min(ul.x+15, lr.x) (+ ul.x+16 when ny >= 2
// and lr.x >= ul.x+16) times rows ul.y .. min(ul.y+15, lr.y) (+ ul.y+16 on
// the same condition).
__global__ __launch_bounds__(256) void prepass_kernel(MapDev m, MapCounters* ctr,
                                                      const sicp_block* visible, View<float> zmin,
                                                      View<float> zmax, PoseArg pa,
                                                      PrepassArgs g, const int* gate) {
    if (gate && *gate == 0) return;
    const uint32_t nvis = ctr->visible;
    const PoseF pose = load_pose(pa);
    uint32_t valid_blocks = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nvis;
         i += gridDim.x * blockDim.x) {
        const sicp_block b = visible[i];
        int ulx, uly, lrx, lry;
        float z0, z1;
        if (!project_block(b, pose, g, m.voxel, ulx, uly, lrx, lry, z0, z1)) continue;
        float dy = (float)lry - (float)uly + 1.0f;
        int ny = (int)ceilf(dy / 16.0f);
        if (ny <= 0) continue;
        ++valid_blocks;
        const int xe = min(ulx + 15, lrx), ye = min(uly + 15, lry);
        const bool xx = ny >= 2 && lrx >= ulx + 16;
        const bool yy = ny >= 2 && lry >= uly + 16;
        const int lo = __float_as_int(z0), hi = __float_as_int(z1);
        for (int yi = uly; yi <= ye + (yy ? 1 : 0); ++yi) {
            int Y = (yi > ye) ? uly + 16 : yi;
            if (Y >= g.zrows) continue;
            for (int xi = ulx; xi <= xe + (xx ? 1 : 0); ++xi) {
                int X = (xi > xe) ? ulx + 16 : xi;
                if (X >= g.zcols) continue;
                atomicMin((int*)&zmin.at(Y, X), lo);
                atomicMax((int*)&zmax.at(Y, X), hi);
            }
        }
    }
    // One atomic per wavefront on the shared counter.
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) valid_blocks += __shfl_xor(valid_blocks, off, 64);
    if ((threadIdx.x & 63) == 0 && valid_blocks) atomicAdd(&ctr->render_tiles, valid_blocks);
}

// ------------------------------------------------------------------ raycast --
struct BlockCache {
    int x, y, z, slot;  // slot -1: empty
};

// DeviceMap::FindVoxel(float3, cache, valid) (DeviceMap.cu:116-133); the cache
// starts empty (the reference's default entry would match block (0,0,0)).
SICP_DEV long find_voxel(const MapDev& m, v3 pt, BlockCache& c, bool& valid) {
    i3 p = floor_i3(pt);
    i3 bp = voxel_to_block(p);
    int idx = voxel_local_idx(p);
    if (c.slot >= 0 && bp.x == c.x && bp.y == c.y && bp.z == c.z) {
        valid = true;
        return (long)c.slot * kBlockVoxels + idx;
    }
    int slot = map_lookup(m, bp.x, bp.y, bp.z);
    if (slot < 0) {
        valid = false;
        return -1;
    }
    valid = true;
    c = BlockCache{bp.x, bp.y, bp.z, slot};
    return (long)slot * kBlockVoxels + idx;
}

SICP_DEV float sdf_at(const MapDev& m, long k) { return k < 0 ? qnanf() : m.geom[k].sdf; }

// Rendering::readSdfInterped (RenderScene.cu:310-343): `valid` is that of the
// last of the eight lookups.
SICP_DEV float read_sdf_interped(const MapDev& m, v3 pt, BlockCache& c, bool& valid) {
    v3 xyz{pt.x - floorf(pt.x), pt.y - floorf(pt.y), pt.z - floorf(pt.z)};
    float s0, s1, r0, r1, r2, r3;
    s0 = sdf_at(m, find_voxel(m, pt, c, valid));
    s1 = sdf_at(m, find_voxel(m, add(pt, mk(1, 0, 0)), c, valid));
    r0 = (1.0f - xyz.x) * s0 + xyz.x * s1;
    s0 = sdf_at(m, find_voxel(m, add(pt, mk(0, 1, 0)), c, valid));
    s1 = sdf_at(m, find_voxel(m, add(pt, mk(1, 1, 0)), c, valid));
    r1 = (1.0f - xyz.x) * s0 + xyz.x * s1;
    r2 = (1.0f - xyz.y) * r0 + xyz.y * r1;
    s0 = sdf_at(m, find_voxel(m, add(pt, mk(0, 0, 1)), c, valid));
    s1 = sdf_at(m, find_voxel(m, add(pt, mk(1, 0, 1)), c, valid));
    r0 = (1.0f - xyz.x) * s0 + xyz.x * s1;
    s0 = sdf_at(m, find_voxel(m, add(pt, mk(0, 1, 1)), c, valid));
    s1 = sdf_at(m, find_voxel(m, add(pt, mk(1, 1, 1)), c, valid));
    r1 = (1.0f - xyz.x) * s0 + xyz.x * s1;
    r3 = (1.0f - xyz.y) * r0 + xyz.y * r1;
    return (1.0f - xyz.z) * r2 + xyz.z * r3;
}

// Rendering::readNormal (RenderScene.cu:345-378).
SICP_DEV bool read_normal(const MapDev& m, v3 pt, BlockCache& c, const PoseF& pose, v3& n) {
    bool valid;
    float s[6];
    const v3 off[6] = {mk(1, 0, 0), mk(-1, 0, 0), mk(0, 1, 0),
                       mk(0, -1, 0), mk(0, 0, 1), mk(0, 0, -1)};
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        s[i] = read_sdf_interped(m, add(pt, off[i]), c, valid);
        if (isnan_(s[i]) || s[i] == 1.0f || !valid) return false;
    }
    n = normalised(mv(pose.Rinv, mk(s[0] - s[1], s[2] - s[3], s[4] - s[5])));
    return true;
}

struct RayArgs {
    float invfx, invfy, cx, cy;
    int cols, rows;
};

__global__ __launch_bounds__(64) void raycast_kernel(MapDev m, View<float4> vmap,
                                                     View<float4> nmap, View<float> zmin,
                                                     View<float> zmax, PoseArg pa, RayArgs g,
                                                     const int* gate, const uint32_t* nonempty) {
    if (gate && *gate == 0) return;
    if (nonempty && *nonempty == 0) return;  // CreateRenderingBlocks returned false
    const int x = blockIdx.x * 8 + (threadIdx.x & 7);
    const int y = blockIdx.y * 8 + (threadIdx.x >> 3);
    if (x >= g.cols || y >= g.rows) return;
    const float nan = qnanf();
    float4 vout = make_float4(nan, nan, nan, nan);
    float4 nout = vout;
    const PoseF pose = load_pose(pa);
    const float z0 = zmin.at(y >> 3, x >> 3);
    const float z1 = zmax.at(y >> 3, x >> 3);
    if (!((double)z1 < 1e-3 || (double)z0 < 1e-3 || isnan_(z0) || isnan_(z1))) {
        float sdf = 1.0f;
        v3 p;
        p.z = z0;
        p.x = p.z * ((float)x - g.cx) * g.invfx;
        p.y = p.z * ((float)y - g.cy) * g.invfy;
        float dist_s = norm(p) * m.voxel_inv;
        v3 block_s = scale(add(mv(pose.R, p), pose.t), m.voxel_inv);
        p.z = z1;
        p.x = p.z * ((float)x - g.cx) * g.invfx;
        p.y = p.z * ((float)y - g.cy) * g.invfy;
        float dist_e = norm(p) * m.voxel_inv;
        v3 block_e = scale(add(mv(pose.R, p), pose.t), m.voxel_inv);
        v3 dir = normalised(sub(block_e, block_s));
        v3 result = block_s;
        bool valid_sdf = false, back_face = false;
        float step, prev_sdf;
        BlockCache cache{0, 0, 0, -1};
        while (dist_s < dist_e) {
            prev_sdf = sdf;
            long k = find_voxel(m, result, cache, valid_sdf);
            if (k < 0) {
                sdf = nan;
            } else {
                uint2 w = ((const uint2*)m.geom)[k];
                sdf = __uint_as_float(w.x);
                if ((w.y & 0xffu) == 0u) valid_sdf = false;
            }
            if (!valid_sdf) {
                step = (float)kBlockSide;
            } else {
                if (sdf <= 0.1f && sdf >= -0.5f) sdf = read_sdf_interped(m, result, cache, valid_sdf);
                if (prev_sdf < 0.f && sdf >= 0.f) {
                    back_face = true;
                    break;
                }
                if (prev_sdf > 0.f && sdf <= 0.f) break;
                step = 1.0f;
            }
            result = add(result, scale(step, dir));
            dist_s += step;
        }
        if (!back_face && sdf <= 0.0f) {
            step = sdf * m.step_scale;
            result = add(result, scale(step, dir));
            sdf = read_sdf_interped(m, result, cache, valid_sdf);
            step = sdf * m.step_scale;
            result = add(result, scale(step, dir));
            v3 normal;
            if (read_normal(m, result, cache, pose, normal)) {
                v3 r = mv(pose.Rinv, sub(scale(result, m.voxel), pose.t));
                vout = make_float4(r.x, r.y, r.z, 1.0f);
                nout = make_float4(normal.x, normal.y, normal.z, 1.0f);
            }
        }
    }
    vmap.at(y, x) = vout;
    nmap.at(y, x) = nout;
}

}  // namespace

int launch_render_prepass(MapDev m, MapCounters* ctr, const sicp_block* visible, View<float> zmin,
                          View<float> zmax, PoseArg pose, sicp_intrinsics K, float dmax,
                          float dmin, const int* gate, hipStream_t s) {
    int n = zmin.cols * zmin.rows;
    zrange_init_kernel<<<div_up(n, 256) > 0 ? div_up(n, 256) : 1, 256, 0, s>>>(zmin, zmax, ctr);
    SICP_LAUNCH_CHECK("zrange_init_kernel");
    PrepassArgs g{K.fx, K.fy, K.cx, K.cy, dmax, dmin, zmin.cols, zmin.rows};
    prepass_kernel<<<512, 256, 0, s>>>(m, ctr, visible, zmin, zmax, pose, g, gate);
    SICP_LAUNCH_CHECK("prepass_kernel");
    return SICP_OK;
}

int launch_raycast(MapDev m, View<float4> vmap, View<float4> nmap, View<float> zmin,
                   View<float> zmax, PoseArg pose, float invfx, float invfy, float cx, float cy,
                   const int* gate, hipStream_t s) {
    RayArgs g{invfx, invfy, cx, cy, vmap.cols, vmap.rows};
    dim3 grid(div_up(vmap.cols, 8), div_up(vmap.rows, 8));
    raycast_kernel<<<grid, 64, 0, s>>>(m, vmap, nmap, zmin, zmax, pose, g, gate, nullptr);
    SICP_LAUNCH_CHECK("raycast_kernel");
    return SICP_OK;
}

int launch_raycast_gated(MapDev m, View<float4> vmap, View<float4> nmap, View<float> zmin,
                         View<float> zmax, PoseArg pose, float invfx, float invfy, float cx,
                         float cy, const int* gate, const uint32_t* nonempty, hipStream_t s) {
    RayArgs g{invfx, invfy, cx, cy, vmap.cols, vmap.rows};
    dim3 grid(div_up(vmap.cols, 8), div_up(vmap.rows, 8));
    raycast_kernel<<<grid, 64, 0, s>>>(m, vmap, nmap, zmin, zmax, pose, g, gate, nonempty);
    SICP_LAUNCH_CHECK("raycast_kernel");
    return SICP_OK;
}

}  // namespace sicp
