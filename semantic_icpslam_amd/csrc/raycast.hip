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
// the same condition). `tile(Y, X, lo, hi)` receives every covered tile with
// the block's depth range as int bit patterns (positive floats order as
// their bits do). Returns whether the block produced tiles.
template <class TileFn>
SICP_DEV bool block_tiles(const sicp_block& b, const PoseF& pose, const PrepassArgs& g,
                          float voxel, TileFn tile) {
    int ulx, uly, lrx, lry;
    float z0, z1;
    if (!project_block(b, pose, g, voxel, ulx, uly, lrx, lry, z0, z1)) return false;
    float dy = (float)lry - (float)uly + 1.0f;
    int ny = (int)ceilf(dy / 16.0f);
    if (ny <= 0) return false;
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
            tile(Y, X, lo, hi);
        }
    }
    return true;
}

// What a prepass thread reads before it can decide anything, loaded at once:
// the gate, the counters, the pose, its first visible block (at a clamped
// index, used only below the visible count) and its share of the fusion's
// per-wavefront counts. Each is a round trip to memory the previous launches
// wrote; issued one after another, as early returns would order them, they
// set the length of a launch this small.
struct PrepassIn {
    bool gated;
    uint32_t nvis, heap_used, fuse_sum;
    uint64_t visible_total, heap_total, fusions;
    PoseF pose;
    sicp_block b0;
};
// kThreads, kGrid: the launch's workgroup size and count (constants, so the
// kernel does not load them from the dispatch packet, one more round trip).
template <int kThreads, int kGrid>
SICP_DEV PrepassIn prepass_loads(const MapDev& m, const MapCounters* ctr,
                                 const sicp_block* visible, const PoseArg& pa, const int* gate,
                                 const uint32_t* fuse_counts) {
    PrepassIn in;
    const uint32_t i0 = blockIdx.x * kThreads + threadIdx.x;
    in.nvis = early_load(&ctr->visible);
    in.heap_used = early_load(&ctr->heap_used);
    in.visible_total = early_load(&ctr->visible_total);
    in.heap_total = early_load(&ctr->heap_total);
    in.fusions = early_load(&ctr->fusions);
    in.b0 = visible[i0 < (uint32_t)m.num_blocks ? i0 : 0u];
    static_assert(kThreads * kGrid >= kFuseBlocks * 4, "one fusion count per thread at most");
    in.fuse_sum = fuse_counts && i0 < (uint32_t)(kFuseBlocks * 4) ? fuse_counts[i0] : 0u;
    in.pose = load_pose(pa);
    in.gated = early_load(gate) == 0;  // never null: MapDev::one for an absent gate
    return in;
}

// The fusion's sums (fuse_end_kernel's, updated_voxels reset before the
// fusion) from its kFuseBlocks * 4 per-wavefront counts: integer sums, one
// atomic per wavefront; workgroup 0 also advances the running totals.
SICP_DEV void fusion_sums(MapCounters* ctr, const PrepassIn& in) {
    uint32_t c = in.fuse_sum;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
    if ((threadIdx.x & 63) == 0 && c) {
        atomicAdd(&ctr->updated_voxels, c);
        atomicAdd(&ctr->updated_total, (unsigned long long)c);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctr->visible_total = in.visible_total + in.nvis;
        ctr->heap_total = in.heap_total + in.heap_used;
        ctr->fusions = in.fusions + 1ull;
    }
}

SICP_DEV void count_render_blocks(MapCounters* ctr, uint32_t valid_blocks) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) valid_blocks += __shfl_xor(valid_blocks, off, 64);
    if ((threadIdx.x & 63) == 0 && valid_blocks) atomicAdd(&ctr->render_tiles, valid_blocks);
}

// The visible blocks of this thread (the first one preloaded), each with
// its tiles passed to `tile`; returns the number that produced tiles.
template <int kThreads, int kGrid, class TileFn>
SICP_DEV uint32_t prepass_blocks(const MapDev& m, const sicp_block* visible, const PrepassIn& in,
                                 const PrepassArgs& g, TileFn tile) {
    uint32_t valid_blocks = 0;
    const uint32_t stride = kGrid * kThreads;
    for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < in.nvis; i += stride)
        valid_blocks += block_tiles(i < stride ? in.b0 : visible[i], in.pose, g, m.voxel, tile);
    return valid_blocks;
}

// The tiles straight into the zRange images: two device-scope atomics per
// covered tile (the fallback for zRange images above kPrepassLdsTiles).
// One-wavefront workgroups: the visible blocks (a thread each) spread over
// more CUs than 256-thread groups would give them.
constexpr int kPrepassGrid = 2048;
__global__ __launch_bounds__(64) void prepass_kernel(MapDev m, MapCounters* ctr,
                                                     const sicp_block* visible, View<float> zmin,
                                                     View<float> zmax, PoseArg pa,
                                                     PrepassArgs g, const int* gate,
                                                     const uint32_t* fuse_counts) {
    const PrepassIn in = prepass_loads<64, kPrepassGrid>(m, ctr, visible, pa, gate, fuse_counts);
    if (in.gated) return;
    if (fuse_counts) fusion_sums(ctr, in);
    auto to_images = [&](int Y, int X, int lo, int hi) {
        atomicMin((int*)&zmin.at(Y, X), lo);
        atomicMax((int*)&zmax.at(Y, X), hi);
    };
    const uint32_t valid_blocks = prepass_blocks<64, kPrepassGrid>(m, visible, in, g, to_images);
    count_render_blocks(ctr, valid_blocks);
}

// The same with the workgroup's tiles gathered in LDS first: its 256
// consecutive visible blocks (heap-slot order, so mostly neighbours) cover
// overlapping tiles, and LDS atomics merge them; the workgroup then applies
// one atomic per touched tile and bound to the zRange images (device-scope
// atomics from eight XCDs are carried out beyond the L2s, one memory
// transaction each). min / max are order-independent, so the images are the
// same as the direct kernel's.
constexpr int kPrepassLdsTiles = 8000;  // 2 x 32000 bytes of LDS
constexpr int kPrepassLdsGrid = 512;
__global__ __launch_bounds__(256) void prepass_lds_kernel(MapDev m, MapCounters* ctr,
                                                          const sicp_block* visible,
                                                          View<float> zmin, View<float> zmax,
                                                          PoseArg pa, PrepassArgs g,
                                                          const int* gate,
                                                          const uint32_t* fuse_counts) {
    __shared__ int tmin[kPrepassLdsTiles], tmax[kPrepassLdsTiles];
    const PrepassIn in =
        prepass_loads<256, kPrepassLdsGrid>(m, ctr, visible, pa, gate, fuse_counts);
    if (in.gated) return;
    if (fuse_counts) fusion_sums(ctr, in);
    if (blockIdx.x * 256 >= in.nvis) return;  // no blocks for this workgroup
    const int ntiles = g.zcols * g.zrows;
    for (int t = threadIdx.x; t < ntiles; t += 256) {
        tmin[t] = 0x7fffffff;
        tmax[t] = 0;
    }
    __syncthreads();
    auto to_lds = [&](int Y, int X, int lo, int hi) {
        atomicMin(&tmin[Y * g.zcols + X], lo);
        atomicMax(&tmax[Y * g.zcols + X], hi);
    };
    const uint32_t valid_blocks =
        prepass_blocks<256, kPrepassLdsGrid>(m, visible, in, g, to_lds);
    count_render_blocks(ctr, valid_blocks);
    __syncthreads();
    for (int t = threadIdx.x; t < ntiles; t += 256) {
        const int lo = tmin[t], hi = tmax[t];
        if (lo == 0x7fffffff) continue;  // untouched (every touch sets both)
        const int Y = t / g.zcols, X = t - Y * g.zcols;
        atomicMin((int*)&zmin.at(Y, X), lo);
        atomicMax((int*)&zmax.at(Y, X), hi);
    }
}

// ------------------------------------------------------------------ raycast --
// Every march iteration is a dependent chain (address, lookup, voxel load,
// decision), run by all of the launch's wavefronts at once (about five share
// each SIMD), so voxel addressing is written for few instructions: floor
// once per distinct float coordinate, block = v >> 3 and
// local = v & 7 (the same values as voxelPosToBlockPos / voxelPosToLocalIdx,
// DeviceMap.cu:169-201, for every int), and one cache compare per block.
struct BlockCache {
    int x, y, z, slot;  // slot -1: empty
    // The first pass's per-wavefront table of recent lookups (null elsewhere).
    unsigned long long* lds = nullptr;
};

// A wavefront's 64 rays cross the same few blocks, so the first pass keeps a
// direct-mapped table of its recent lookups in LDS, hits and misses alike:
// 256 entries of one 64-bit word {tag = 1 + the block coordinates in 14 bits
// each, slot + 1 (0: no block)}, written by atomic exchange so that lanes
// storing the same entry at once never leave a torn word. Blocks outside
// +-8192 per axis, and heaps above 2^21 - 2 blocks, go to the table directly.
// The map is read-only during the raycast, so a cached answer is the table's.
constexpr int kRayLdsEntries = 256;
SICP_DEV int lds_lookup(const MapDev& m, unsigned long long* t, int bx, int by, int bz) {
    const bool fits = (unsigned)(bx + 8192) < 16384u && (unsigned)(by + 8192) < 16384u &&
                      (unsigned)(bz + 8192) < 16384u;
    if (!fits) return map_lookup(m, bx, by, bz);
    const unsigned long long tag = 1ull + (((unsigned long long)(bx & 0x3fff) << 28) |
                                           ((unsigned long long)(by & 0x3fff) << 14) |
                                           (unsigned long long)(bz & 0x3fff));
    const uint32_t i = (((uint32_t)bx * 0x9E3779B1u) ^ ((uint32_t)by * 0x85EBCA77u) ^
                        ((uint32_t)bz * 0xC2B2AE3Du)) >> 24;
    const unsigned long long e = t[i];
    if ((e >> 21) == tag) return (int)(e & 0x1fffffull) - 1;
    const int s = map_lookup(m, bx, by, bz);
    atomicExch(&t[i], (tag << 21) | (unsigned long long)(s + 1));
    return s;
}
SICP_DEV int cached_lookup(const MapDev& m, const BlockCache& c, int bx, int by, int bz) {
    return c.lds ? lds_lookup(m, c.lds, bx, by, bz) : map_lookup(m, bx, by, bz);
}

// make_int3 of one coordinate (VectorMath.h:36-42): the truncation, one less
// where it exceeds the value — the floor, for every float (NaN and values
// past the int range convert as the truncation does). v_floor_f32 and one
// conversion, the floor shared with the fraction of an interpolated read.
SICP_DEV int floor_i(float a) { return (int)floorf(a); }

SICP_DEV int local_of(int vx, int vy, int vz) { return ((vz & 7) << 6) | ((vy & 7) << 3) | (vx & 7); }

SICP_DEV bool cache_has(const BlockCache& c, int bx, int by, int bz) {
    return c.slot >= 0 && bx == c.x && by == c.y && bz == c.z;
}

// DeviceMap::FindVoxel(int3, cache, valid) (DeviceMap.cu:116-133) on integer
// voxel coordinates; the cache starts empty (the reference's default entry
// would match block (0,0,0)).
SICP_DEV long find_voxel_i(const MapDev& m, int vx, int vy, int vz, BlockCache& c, bool& valid) {
    const int bx = vx >> 3, by = vy >> 3, bz = vz >> 3;
    if (cache_has(c, bx, by, bz)) {
        valid = true;
        return (long)c.slot * kBlockVoxels + local_of(vx, vy, vz);
    }
    const int slot = cached_lookup(m, c, bx, by, bz);
    if (slot < 0) {
        valid = false;
        return -1;
    }
    valid = true;
    c = BlockCache{bx, by, bz, slot, c.lds};
    return (long)slot * kBlockVoxels + local_of(vx, vy, vz);
}

SICP_DEV float sdf_at(const MapDev& m, long k) { return k < 0 ? qnanf() : vox_sdf(m, (size_t)k); }

// One axis of an interpolated read centred at coordinate q: the voxel
// coordinates of q and q + 1 (the reference adds 1.0f to the float point)
// and the fraction q - floor(q).
struct AxisPair {
    int v0, v1;
    float frac;
};
SICP_DEV AxisPair axis_pair(float q) { return AxisPair{floor_i(q), floor_i(q + 1.0f), q - floorf(q)}; }

// Slots of the (up to) eight blocks the corners of an interpolated read touch:
// per axis the blocks of v0 and v1 (equal, or neighbours). One lookup site,
// looped, instead of one per corner.
struct CornerBlocks {
    int s[8];  // [dx + 2 dy + 4 dz]; dx = 1 only when the x blocks differ, etc.
};
SICP_DEV void corner_blocks(const MapDev& m, const AxisPair& X, const AxisPair& Y,
                            const AxisPair& Z, const BlockCache& c, CornerBlocks& cb) {
    const int bx0 = X.v0 >> 3, by0 = Y.v0 >> 3, bz0 = Z.v0 >> 3;
    const int nx = (X.v1 >> 3) != bx0, ny = (Y.v1 >> 3) != by0, nz = (Z.v1 >> 3) != bz0;
    int s0 = -1, s1 = -1, s2 = -1, s3 = -1, s4 = -1, s5 = -1, s6 = -1, s7 = -1;
#pragma unroll 1
    for (int b = 0; b < 8; ++b) {
        const int dx = b & 1, dy = (b >> 1) & 1, dz = b >> 2;
        if (dx > nx || dy > ny || dz > nz) continue;
        const int bx = dx ? (X.v1 >> 3) : bx0, by = dy ? (Y.v1 >> 3) : by0,
                  bz = dz ? (Z.v1 >> 3) : bz0;
        const int sl = cache_has(c, bx, by, bz) ? c.slot : cached_lookup(m, c, bx, by, bz);
        s0 = b == 0 ? sl : s0;
        s1 = b == 1 ? sl : s1;
        s2 = b == 2 ? sl : s2;
        s3 = b == 3 ? sl : s3;
        s4 = b == 4 ? sl : s4;
        s5 = b == 5 ? sl : s5;
        s6 = b == 6 ? sl : s6;
        s7 = b == 7 ? sl : s7;
    }
    cb.s[0] = s0;
    cb.s[1] = s1;
    cb.s[2] = s2;
    cb.s[3] = s3;
    cb.s[4] = s4;
    cb.s[5] = s5;
    cb.s[6] = s6;
    cb.s[7] = s7;
}

// The slot of corner (cx, cy, cz) in {0,1}^3 (selected without indexing, so
// the slots stay in registers).
SICP_DEV int corner_slot(const CornerBlocks& cb, const AxisPair& X, const AxisPair& Y,
                         const AxisPair& Z, int cx, int cy, int cz) {
    const int dx = cx && ((X.v1 >> 3) != (X.v0 >> 3));
    const int dy = cy && ((Y.v1 >> 3) != (Y.v0 >> 3));
    const int dz = cz && ((Z.v1 >> 3) != (Z.v0 >> 3));
    const int a0 = dx ? cb.s[1] : cb.s[0], a1 = dx ? cb.s[3] : cb.s[2];
    const int a2 = dx ? cb.s[5] : cb.s[4], a3 = dx ? cb.s[7] : cb.s[6];
    const int b0 = dy ? a1 : a0, b1 = dy ? a3 : a2;
    return dz ? b1 : b0;
}

// Rendering::readSdfInterped (RenderScene.cu:310-343) in three parts: the
// voxel indices of the eight corners (-1 for a missing block) with `valid`
// of the last corner's lookup, the eight loads, the blend. Corner order,
// NaN for a missing block and the blend order are the reference's. When all
// eight corners lie in the cached block (the common case) no lookup is made.
SICP_DEV void interp_corners(const MapDev& m, const AxisPair& X, const AxisPair& Y,
                             const AxisPair& Z, const BlockCache& c, int (&idx)[8], bool& valid) {
    if (cache_has(c, X.v0 >> 3, Y.v0 >> 3, Z.v0 >> 3) && (X.v1 >> 3) == c.x &&
        (Y.v1 >> 3) == c.y && (Z.v1 >> 3) == c.z) {
        const int base = c.slot * kBlockVoxels;
#pragma unroll
        for (int i = 0; i < 8; ++i)
            idx[i] = base + local_of((i & 1) ? X.v1 : X.v0, (i & 2) ? Y.v1 : Y.v0,
                                     (i & 4) ? Z.v1 : Z.v0);
        valid = true;
    } else {
        CornerBlocks cb;
        corner_blocks(m, X, Y, Z, c, cb);
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int sl = corner_slot(cb, X, Y, Z, i & 1, (i >> 1) & 1, i >> 2);
            idx[i] = sl < 0 ? -1
                            : sl * kBlockVoxels + local_of((i & 1) ? X.v1 : X.v0,
                                                           (i & 2) ? Y.v1 : Y.v0,
                                                           (i & 4) ? Z.v1 : Z.v0);
        }
        valid = corner_slot(cb, X, Y, Z, 1, 1, 1) >= 0;
    }
}
// Every load issued (a missing corner reads voxel 0 and is replaced by NaN).
SICP_DEV void interp_loads(const MapDev& m, const int (&idx)[8], float (&sd)[8]) {
#pragma unroll
    for (int i = 0; i < 8; ++i) sd[i] = vox_sdf(m, (size_t)(idx[i] < 0 ? 0 : idx[i]));
#pragma unroll
    for (int i = 0; i < 8; ++i) sd[i] = idx[i] < 0 ? qnanf() : sd[i];
}
SICP_DEV float interp_blend(const float (&sd)[8], const AxisPair& X, const AxisPair& Y,
                            const AxisPair& Z) {
    const float fx = X.frac, fy = Y.frac, fz = Z.frac;
    float r0, r1, r2, r3;
    r0 = (1.0f - fx) * sd[0] + fx * sd[1];
    r1 = (1.0f - fx) * sd[2] + fx * sd[3];
    r2 = (1.0f - fy) * r0 + fy * r1;
    r0 = (1.0f - fx) * sd[4] + fx * sd[5];
    r1 = (1.0f - fx) * sd[6] + fx * sd[7];
    r3 = (1.0f - fy) * r0 + fy * r1;
    return (1.0f - fz) * r2 + fz * r3;
}
SICP_DEV float interp_axes(const MapDev& m, AxisPair X, AxisPair Y, AxisPair Z, BlockCache& c,
                           bool& valid) {
    int idx[8];
    float sd[8];
    interp_corners(m, X, Y, Z, c, idx, valid);
    interp_loads(m, idx, sd);
    return interp_blend(sd, X, Y, Z);
}

SICP_DEV float read_sdf_interped(const MapDev& m, v3 pt, BlockCache& c, bool& valid) {
    return interp_axes(m, axis_pair(pt.x), axis_pair(pt.y), axis_pair(pt.z), c, valid);
}

// Three interpolated reads with their 24 loads issued together.
SICP_DEV void interp3(const MapDev& m, const BlockCache& c, const AxisPair (&P)[3][3],
                      float (&s)[3], bool (&v)[3]) {
    int i0[8], i1[8], i2[8];
    interp_corners(m, P[0][0], P[0][1], P[0][2], c, i0, v[0]);
    interp_corners(m, P[1][0], P[1][1], P[1][2], c, i1, v[1]);
    interp_corners(m, P[2][0], P[2][1], P[2][2], c, i2, v[2]);
    float a0[8], a1[8], a2[8];
    interp_loads(m, i0, a0);
    interp_loads(m, i1, a1);
    interp_loads(m, i2, a2);
    s[0] = interp_blend(a0, P[0][0], P[0][1], P[0][2]);
    s[1] = interp_blend(a1, P[1][0], P[1][1], P[1][2]);
    s[2] = interp_blend(a2, P[2][0], P[2][1], P[2][2]);
}

// Rendering::readNormal (RenderScene.cu:345-378): six interpolated reads at
// pt +- 1 voxel along each axis (pt.a + 1.0f, pt.a + -1.0f), in order, the
// first failing one deciding. The three centre values per axis are
// decomposed once and shared by the six reads. The reads have no side
// effect, so they run three at a time (their loads in flight together) and
// the tests follow in the reference's order: the same result.
SICP_DEV bool read_normal(const MapDev& m, v3 pt, BlockCache& c, const PoseF& pose, v3& n) {
    const AxisPair xm = axis_pair(pt.x + -1.0f), x0 = axis_pair(pt.x), xp = axis_pair(pt.x + 1.0f);
    const AxisPair ym = axis_pair(pt.y + -1.0f), y0 = axis_pair(pt.y), yp = axis_pair(pt.y + 1.0f);
    const AxisPair zm = axis_pair(pt.z + -1.0f), z0 = axis_pair(pt.z), zp = axis_pair(pt.z + 1.0f);
    float s[6];
    bool ok[6];
    {
        const AxisPair P[3][3] = {{xp, y0, z0}, {xm, y0, z0}, {x0, yp, z0}};
        float t[3];
        bool v[3];
        interp3(m, c, P, t, v);
        for (int i = 0; i < 3; ++i) {
            s[i] = t[i];
            ok[i] = v[i];
        }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
        if (isnan_(s[i]) || s[i] == 1.0f || !ok[i]) return false;
    {
        const AxisPair P[3][3] = {{x0, ym, z0}, {x0, y0, zp}, {x0, y0, zm}};
        float t[3];
        bool v[3];
        interp3(m, c, P, t, v);
        for (int i = 0; i < 3; ++i) {
            s[3 + i] = t[i];
            ok[3 + i] = v[i];
        }
    }
#pragma unroll
    for (int i = 3; i < 6; ++i)
        if (isnan_(s[i]) || s[i] == 1.0f || !ok[i]) return false;
    n = normalised(mv(pose.Rinv, mk(s[0] - s[1], s[2] - s[3], s[4] - s[5])));
    return true;
}

struct RayArgs {
    float invfx, invfy, cx, cy;
    int cols, rows;
    int first_pass_samples;  // the first pass's hand-over point (raycast_kernel only)
    int lds_block_table;     // raycast_kernel: block lookups through its LDS table
};

// The ray of pixel (x, y) (RenderScene.cu:393-421): start point and unit
// direction in voxel units, start and end distances along it.
struct Ray {
    v3 start, dir;
    float dist_s, dist_e;
};
SICP_DEV bool ray_setup(const MapDev& m, float z0, float z1, const PoseF& pose, const RayArgs& g,
                        int x, int y, Ray& r) {
    if ((double)z1 < 1e-3 || (double)z0 < 1e-3 || isnan_(z0) || isnan_(z1)) return false;
    v3 p;
    p.z = z0;
    p.x = p.z * ((float)x - g.cx) * g.invfx;
    p.y = p.z * ((float)y - g.cy) * g.invfy;
    r.dist_s = norm(p) * m.voxel_inv;
    const v3 block_s = scale(add(mv(pose.R, p), pose.t), m.voxel_inv);
    p.z = z1;
    p.x = p.z * ((float)x - g.cx) * g.invfx;
    p.y = p.z * ((float)y - g.cy) * g.invfy;
    r.dist_e = norm(p) * m.voxel_inv;
    const v3 block_e = scale(add(mv(pose.R, p), pose.t), m.voxel_inv);
    r.dir = normalised(sub(block_e, block_s));
    r.start = block_s;
    return true;
}
SICP_DEV bool ray_setup(const MapDev& m, const View<float>& zmin, const View<float>& zmax,
                        const PoseF& pose, const RayArgs& g, int x, int y, Ray& r) {
    return ray_setup(m, zmin.at(y >> 3, x >> 3), zmax.at(y >> 3, x >> 3), pose, g, x, y, r);
}

// RayArgs::first_pass_samples is the number of samples one ray marches in
// the first pass; rays still marching then are handed to the second pass.
// Most rays end well before; the few that graze a surface for hundreds of
// one-voxel steps would otherwise set the length of the whole launch.

// First pass: one lane per ray, 8x8-pixel wavefronts (the zRange tile).
__global__ __launch_bounds__(64, 5) void raycast_kernel(MapDev m, View<float4> vmap,
                                                     View<float4> nmap, View<float> zmin,
                                                     View<float> zmax, PoseArg pa, RayArgs g,
                                                     const int* gate, const uint32_t* nonempty) {
    const int x = blockIdx.x * 8 + (threadIdx.x & 7);
    const int y = blockIdx.y * 8 + (threadIdx.x >> 3);
    // the tile's depth range, the pose and the gates, loaded together
    const int ty = min(y, g.rows - 1) >> 3, tx = min(x, g.cols - 1) >> 3;
    const float z0 = early_load(&zmin.at(ty, tx)), z1 = early_load(&zmax.at(ty, tx));
    // gate and nonempty (CreateRenderingBlocks returned true) are never null
    // here: the launch passes MapDev::one for an absent one
    const int open = early_load(gate);
    const uint32_t any = early_load(nonempty);
    const PoseF pose = load_pose(pa);
    if ((open == 0) | (any == 0u)) return;
    // One wavefront per workgroup: the table is cleared before any lane uses
    // it, and its accesses are ordered by the wavefront's own instruction
    // order (no barrier needed).
    __shared__ unsigned long long lookup_table[kRayLdsEntries];
#pragma unroll
    for (int i = 0; i < kRayLdsEntries / 64; ++i) lookup_table[threadIdx.x + 64 * i] = 0ull;
    if (x >= g.cols || y >= g.rows) return;
    const float nan = qnanf();
    float4 vout = make_float4(nan, nan, nan, nan);
    float4 nout = vout;
    Ray ray;
    bool deferred = false;
    v3 result{0.f, 0.f, 0.f};
    float dist_s = 0.f, sdf = 1.0f;
    if (ray_setup(m, z0, z1, pose, g, x, y, ray)) {
        result = ray.start;
        dist_s = ray.dist_s;
        bool valid_sdf = false, back_face = false;
        float step, prev_sdf;
        BlockCache cache{0, 0, 0, -1,
                         g.lds_block_table && m.num_blocks < (1 << 21) - 2 ? lookup_table
                                                                            : nullptr};
        int iters = 0;
        while (dist_s < ray.dist_e) {
            if (iters == g.first_pass_samples) {
                deferred = true;
                break;
            }
            ++iters;
            prev_sdf = sdf;
            long k = find_voxel_i(m, floor_i(result.x), floor_i(result.y), floor_i(result.z), cache,
                                  valid_sdf);
            if (k < 0) {
                sdf = nan;
            } else {
                uint32_t wt;
                vox_load(m, (size_t)k, sdf, wt);
                if (wt == 0u) valid_sdf = false;
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
            result = add(result, scale(step, ray.dir));
            dist_s += step;
        }
        if (!deferred && !back_face && sdf <= 0.0f) {
            step = sdf * m.step_scale;
            result = add(result, scale(step, ray.dir));
            sdf = read_sdf_interped(m, result, cache, valid_sdf);
            step = sdf * m.step_scale;
            result = add(result, scale(step, ray.dir));
            v3 normal;
            if (read_normal(m, result, cache, pose, normal)) {
                v3 r = mv(pose.Rinv, sub(scale(result, m.voxel), pose.t));
                vout = make_float4(r.x, r.y, r.z, 1.0f);
                nout = make_float4(normal.x, normal.y, normal.z, 1.0f);
            }
        }
    }
    // Hand-over: one queue reservation per wavefront.
    const unsigned long long mask = __ballot(deferred);
    if (mask) {
        const int lane = threadIdx.x & 63;
        const int leader = __ffsll((long long)mask) - 1;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(m.rq_count, (uint32_t)__popcll(mask));
        base = __shfl(base, leader, 64);
        if (deferred) {
            const uint32_t q = base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
            if (q < m.rq_cap)
                m.rq[q] = RayState{y * g.cols + x, result.x, result.y, result.z, dist_s, sdf,
                                   ray.dir.x, ray.dir.y, ray.dir.z, ray.dist_e, {0, 0}};
        }
        if (deferred) return;  // the second pass writes this pixel
    }
    vmap.at(y, x) = vout;
    nmap.at(y, x) = nout;
}

// ---- second pass: kRayGroup lanes per ray --------------------------------
// Lane k of a group evaluates the sample the scalar loop reaches after k more
// one-voxel steps (result + dir added k times, as the loop adds it). A ballot
// finds the first sample where the scalar loop would do anything but a plain
// one-voxel step — out of range, missing or empty voxel, the interpolation
// band, a sign change — and the group applies the loop body there exactly;
// the samples before it are one-voxel steps taken in one go. Interpolated
// reads and the normal split their corners over the group's lanes.
constexpr int kRayGroup = 8;
static_assert(48 % kRayGroup == 0 && kRayGroup >= 8, "corner split of the normal read");

template <class T> SICP_DEV T gshfl(T v, int src) { return __shfl(v, src, kRayGroup); }
SICP_DEV v3 gshfl3(v3 v, int src) { return v3{gshfl(v.x, src), gshfl(v.y, src), gshfl(v.z, src)}; }

// Rendering::readSdfInterped with corner c on lane c of the group; every lane
// returns the blended value and the validity of the last lookup.
SICP_DEV float coop_interp(const MapDev& m, v3 pt, int k, BlockCache& c, bool& valid) {
    const AxisPair X = axis_pair(pt.x), Y = axis_pair(pt.y), Z = axis_pair(pt.z);
    float mine = qnanf();
    bool ok = false;
    if (k < 8) {
        long idx = find_voxel_i(m, (k & 1) ? X.v1 : X.v0, (k & 2) ? Y.v1 : Y.v0,
                                (k & 4) ? Z.v1 : Z.v0, c, ok);
        mine = sdf_at(m, idx);
    }
    float sd[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) sd[i] = gshfl(mine, i);
    valid = gshfl((int)ok, 7) != 0;
    const float fx = X.frac, fy = Y.frac, fz = Z.frac;
    float r0, r1, r2, r3;
    r0 = (1.0f - fx) * sd[0] + fx * sd[1];
    r1 = (1.0f - fx) * sd[2] + fx * sd[3];
    r2 = (1.0f - fy) * r0 + fy * r1;
    r0 = (1.0f - fx) * sd[4] + fx * sd[5];
    r1 = (1.0f - fx) * sd[6] + fx * sd[7];
    r3 = (1.0f - fy) * r0 + fy * r1;
    return (1.0f - fz) * r2 + fz * r3;
}

// Rendering::readNormal with its 48 corners spread over the group's lanes.
SICP_DEV bool coop_normal(const MapDev& m, v3 pt, int k, BlockCache& c, const PoseF& pose,
                          v3& n) {
    const AxisPair xm = axis_pair(pt.x + -1.0f), x0 = axis_pair(pt.x), xp = axis_pair(pt.x + 1.0f);
    const AxisPair ym = axis_pair(pt.y + -1.0f), y0 = axis_pair(pt.y), yp = axis_pair(pt.y + 1.0f);
    const AxisPair zm = axis_pair(pt.z + -1.0f), z0 = axis_pair(pt.z), zp = axis_pair(pt.z + 1.0f);
    constexpr int kPer = 48 / kRayGroup;  // corners per lane
    float val[kPer];
    int ok[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const int q = k + kRayGroup * j;  // corner q of the 48: read q / 8, corner q % 8
        const int i = q >> 3, cc = q & 7;
        const AxisPair X = i == 0 ? xp : (i == 1 ? xm : x0);
        const AxisPair Y = i == 2 ? yp : (i == 3 ? ym : y0);
        const AxisPair Z = i == 4 ? zp : (i == 5 ? zm : z0);
        bool v;
        long idx = find_voxel_i(m, (cc & 1) ? X.v1 : X.v0, (cc & 2) ? Y.v1 : Y.v0,
                                (cc & 4) ? Z.v1 : Z.v0, c, v);
        val[j] = sdf_at(m, idx);
        ok[j] = v ? 1 : 0;
    }
    float s[6];
    bool valid[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        float sd[8];
#pragma unroll
        for (int cc = 0; cc < 8; ++cc) {
            const int q = i * 8 + cc;
            sd[cc] = gshfl(val[q / kRayGroup], q % kRayGroup);
        }
        valid[i] = gshfl(ok[(i * 8 + 7) / kRayGroup], (i * 8 + 7) % kRayGroup) != 0;
        const AxisPair X = i == 0 ? xp : (i == 1 ? xm : x0);
        const AxisPair Y = i == 2 ? yp : (i == 3 ? ym : y0);
        const AxisPair Z = i == 4 ? zp : (i == 5 ? zm : z0);
        const float fx = X.frac, fy = Y.frac, fz = Z.frac;
        float r0, r1, r2, r3;
        r0 = (1.0f - fx) * sd[0] + fx * sd[1];
        r1 = (1.0f - fx) * sd[2] + fx * sd[3];
        r2 = (1.0f - fy) * r0 + fy * r1;
        r0 = (1.0f - fx) * sd[4] + fx * sd[5];
        r1 = (1.0f - fx) * sd[6] + fx * sd[7];
        r3 = (1.0f - fy) * r0 + fy * r1;
        s[i] = (1.0f - fz) * r2 + fz * r3;
    }
#pragma unroll
    for (int i = 0; i < 6; ++i)
        if (isnan_(s[i]) || s[i] == 1.0f || !valid[i]) return false;
    n = normalised(mv(pose.Rinv, mk(s[0] - s[1], s[2] - s[3], s[4] - s[5])));
    return true;
}

constexpr int kRayLongThreads = 256;
__global__ __launch_bounds__(kRayLongThreads) void raycast_long_kernel(
    MapDev m, View<float4> vmap, View<float4> nmap, View<float> zmin, View<float> zmax, PoseArg pa,
    RayArgs g, const int* gate, const uint32_t* nonempty) {
    constexpr uint32_t kGroupsPerWg = kRayLongThreads / kRayGroup;
    const uint32_t i0 = blockIdx.x * kGroupsPerWg + (threadIdx.x / kRayGroup);
    // the queue's count, this group's first entry (at a clamped index, used
    // only below the count), the pose and the gates, loaded together
    const uint32_t queued = early_load(m.rq_count);
    const RayState st0 = m.rq[i0 < m.rq_cap ? i0 : 0u];
    const int open = early_load(gate);  // gate and nonempty: never null here
    const uint32_t any = early_load(nonempty);
    const PoseF pose = load_pose(pa);
    if ((open == 0) | (any == 0u)) return;
    const uint32_t n = min(queued, m.rq_cap);
    const int k = threadIdx.x & (kRayGroup - 1);
    const int gbase = (threadIdx.x & 63) & ~(kRayGroup - 1);  // first lane of the group
    const uint32_t groups = gridDim.x * kGroupsPerWg;
    const float nan = qnanf();
    for (uint32_t i = i0; i < n; i += groups) {
        const RayState st = i == i0 ? st0 : m.rq[i];
        const int x = st.px % g.cols, y = st.px / g.cols;
        Ray ray;  // the first pass's ray_setup
        ray.dir = v3{st.dx, st.dy, st.dz};
        ray.dist_e = st.dist_e;
        v3 result{st.rx, st.ry, st.rz};
        float dist_s = st.dist_s, sdf = st.sdf;
        bool back_face = false;
        BlockCache cache{0, 0, 0, -1};
        while (dist_s < ray.dist_e) {
            v3 P = result;
            float D = dist_s;
#pragma unroll
            for (int j = 0; j < kRayGroup - 1; ++j)
                if (j < k) {
                    P = add(P, scale(1.0f, ray.dir));
                    D += 1.0f;
                }
            const bool inr = D < ray.dist_e;
            float sv = nan;
            int exists = 0, wok = 0;
            if (inr) {
                bool v;
                const long idx = find_voxel_i(m, floor_i(P.x), floor_i(P.y), floor_i(P.z), cache, v);
                if (idx >= 0) {
                    uint32_t wt;
                    vox_load(m, (size_t)idx, sv, wt);
                    exists = 1;
                    wok = wt != 0u;
                }
            }
            float sp = __shfl_up(sv, 1, kRayGroup);
            if (k == 0) sp = sdf;
            const bool band = wok && sv <= 0.1f && sv >= -0.5f;
            const bool cross = wok && !band && ((sp < 0.f && sv >= 0.f) || (sp > 0.f && sv <= 0.f));
            const bool event = !inr || !wok || band || cross;
            const uint32_t gm = (uint32_t)(__ballot(event) >> gbase) & ((1u << kRayGroup) - 1u);
            if (gm == 0u) {  // kRayGroup one-voxel steps
                result = add(gshfl3(P, kRayGroup - 1), scale(1.0f, ray.dir));
                dist_s = gshfl(D, kRayGroup - 1) + 1.0f;
                sdf = gshfl(sv, kRayGroup - 1);
                continue;
            }
            const int ks = __ffs(gm) - 1;
            result = gshfl3(P, ks);
            dist_s = gshfl(D, ks);
            const float prev = gshfl(sp, ks);
            if (!gshfl((int)inr, ks)) {  // the loop condition fails here
                sdf = prev;
                break;
            }
            const float sk = gshfl(sv, ks);
            if (!gshfl(wok, ks)) {  // missing block (NaN) or empty voxel: block-width step
                sdf = gshfl(exists, ks) ? sk : nan;
                result = add(result, scale((float)kBlockSide, ray.dir));
                dist_s += (float)kBlockSide;
                continue;
            }
            float s1 = sk;
            if (gshfl((int)band, ks)) {
                bool v;
                s1 = coop_interp(m, result, k, cache, v);
            }
            if (prev < 0.f && s1 >= 0.f) {
                back_face = true;
                break;
            }
            if (prev > 0.f && s1 <= 0.f) {
                sdf = s1;
                break;
            }
            sdf = s1;
            result = add(result, scale(1.0f, ray.dir));
            dist_s += 1.0f;
        }
        float4 vout = make_float4(nan, nan, nan, nan);
        float4 nout = vout;
        if (!back_face && sdf <= 0.0f) {
            float step = sdf * m.step_scale;
            result = add(result, scale(step, ray.dir));
            bool v;
            sdf = coop_interp(m, result, k, cache, v);
            step = sdf * m.step_scale;
            result = add(result, scale(step, ray.dir));
            v3 normal;
            if (coop_normal(m, result, k, cache, pose, normal)) {
                v3 r = mv(pose.Rinv, sub(scale(result, m.voxel), pose.t));
                vout = make_float4(r.x, r.y, r.z, 1.0f);
                nout = make_float4(normal.x, normal.y, normal.z, 1.0f);
            }
        }
        if (k == 0) {
            vmap.at(y, x) = vout;
            nmap.at(y, x) = nout;
        }
    }
}

__global__ void ray_queue_reset_kernel(uint32_t* count) { *count = 0u; }

// ------------------------------------------------------------------ adjoint --
// The raycast's vertex and normal as functions of the camera pose and of the
// TSDF values, with the zRange images and every discrete decision of the
// march held fixed. Per sampled pixel the march is recomputed (nothing is
// stored by the forward pass), then the chain
//   r0 = block_s + M dir                      (M: the march's summed steps)
//   r1 = r0 + k s0 dir, s0 = F(r0) or raw voxel (k: stepScale)
//   r2 = r1 + k s1 dir, s1 = F(r1)
//   v  = R^T (voxel r2 - t)
//   n  = normalised(R^T g), g_a = F(r2 + e_a) - F(r2 - e_a)
// is run backwards (F: the trilinear read). The pose perturbation is
// T' = T exp(xi), xi = (translation, rotation) in the camera frame:
// R' = R (I + [w]x), t' = t + R v.

// A trilinear read (readSdfInterped, RenderScene.cu:310-343) with its eight
// corner voxels, their weights and its gradient in voxel units. The value is
// formed exactly as the forward read forms it.
struct TriRead {
    long k[8];
    float w[8];
    v3 grad;
    float val;
    bool ok;  // every corner exists, the value is not NaN
};
SICP_DEV void tri_read(const MapDev& m, v3 p, TriRead& t) {
    const AxisPair X = axis_pair(p.x), Y = axis_pair(p.y), Z = axis_pair(p.z);
    float sd[8];
    t.ok = true;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int vx = (i & 1) ? X.v1 : X.v0, vy = (i & 2) ? Y.v1 : Y.v0,
                  vz = (i & 4) ? Z.v1 : Z.v0;
        const int slot = map_lookup(m, vx >> 3, vy >> 3, vz >> 3);
        if (slot < 0) {
            t.k[i] = -1;
            sd[i] = qnanf();
            t.ok = false;
        } else {
            t.k[i] = (long)slot * kBlockVoxels + local_of(vx, vy, vz);
            sd[i] = vox_sdf(m, (size_t)t.k[i]);
        }
    }
    const float fx = X.frac, fy = Y.frac, fz = Z.frac;
    float r0, r1, r2, r3;
    r0 = (1.0f - fx) * sd[0] + fx * sd[1];
    r1 = (1.0f - fx) * sd[2] + fx * sd[3];
    r2 = (1.0f - fy) * r0 + fy * r1;
    const float a0 = r0, a1 = r1;
    r0 = (1.0f - fx) * sd[4] + fx * sd[5];
    r1 = (1.0f - fx) * sd[6] + fx * sd[7];
    r3 = (1.0f - fy) * r0 + fy * r1;
    t.val = (1.0f - fz) * r2 + fz * r3;
    t.ok = t.ok && !isnan_(t.val);
#pragma unroll
    for (int i = 0; i < 8; ++i)
        t.w[i] = ((i & 1) ? fx : 1.0f - fx) * ((i & 2) ? fy : 1.0f - fy) * ((i & 4) ? fz : 1.0f - fz);
    t.grad.x = (1.0f - fz) * ((1.0f - fy) * (sd[1] - sd[0]) + fy * (sd[3] - sd[2])) +
               fz * ((1.0f - fy) * (sd[5] - sd[4]) + fy * (sd[7] - sd[6]));
    t.grad.y = (1.0f - fz) * (a1 - a0) + fz * (r1 - r0);
    t.grad.z = r3 - r2;
}

// dL/d(values) += g * weights.
SICP_DEV void tri_scatter(const TriRead& t, float g, float* grad_sdf) {
    if (!grad_sdf || g == 0.0f) return;
#pragma unroll
    for (int i = 0; i < 8; ++i)
        if (t.k[i] >= 0 && t.w[i] != 0.0f) atomicAdd(grad_sdf + t.k[i], g * t.w[i]);
}

struct AdjointArgs {
    RayArgs g;
    const int32_t* pixels;
    int n;
    const float* grad_v;   // n x 3, d loss / d vertex (camera frame)
    const float* grad_n;   // n x 3, d loss / d normal; may be null
    float* grad_sdf;       // num_blocks x 512, accumulated; may be null
    float* grad_px;        // n x 6 per-pixel pose gradients; may be null
    double* partials;      // 6 x kAdjMaxBlocks
};
constexpr int kAdjThreads = 256;
constexpr int kAdjMaxBlocks = 256;

// One pixel: false (and no gradient) where the forward pass writes NaN.
SICP_DEV bool adjoint_pixel(const MapDev& m, const View<float>& zmin, const View<float>& zmax,
                            const PoseF& pose, const RayArgs& g, int x, int y, v3 gv, v3 gn,
                            float* grad_sdf, float gxi[6]) {
    Ray ray;
    if (!ray_setup(m, zmin, zmax, pose, g, x, y, ray)) return false;
    // the march of raycast_kernel, uncapped, with its summed step
    v3 result = ray.start;
    float dist_s = ray.dist_s, sdf = 1.0f, msum = 0.0f;
    bool valid_sdf = false, interp0 = false;
    BlockCache cache{0, 0, 0, -1};
    long k0 = -1;
    while (dist_s < ray.dist_e) {
        const float prev = sdf;
        const long k = find_voxel_i(m, floor_i(result.x), floor_i(result.y), floor_i(result.z),
                                    cache, valid_sdf);
        if (k < 0) {
            sdf = qnanf();
        } else {
            uint32_t wt;
            vox_load(m, (size_t)k, sdf, wt);
            if (wt == 0u) valid_sdf = false;
        }
        k0 = k;
        interp0 = false;
        float step;
        if (!valid_sdf) {
            step = (float)kBlockSide;
        } else {
            if (sdf <= 0.1f && sdf >= -0.5f) {
                sdf = read_sdf_interped(m, result, cache, valid_sdf);
                interp0 = true;
            }
            if (prev < 0.f && sdf >= 0.f) return false;  // back face
            if (prev > 0.f && sdf <= 0.f) break;
            step = 1.0f;
        }
        result = add(result, scale(step, ray.dir));
        dist_s += step;
        msum += step;
    }
    if (!(sdf <= 0.0f)) return false;
    const float ks = m.step_scale;
    const v3 r0 = result;
    TriRead t0;
    if (interp0) tri_read(m, r0, t0);
    const float s0 = sdf;
    const v3 r1 = add(r0, scale(s0 * ks, ray.dir));
    TriRead t1;
    tri_read(m, r1, t1);
    const float s1 = t1.val;
    const v3 r2 = add(r1, scale(s1 * ks, ray.dir));
    // readNormal's six reads: pt.a + 1.0f / pt.a + -1.0f on one axis, pt on
    // the others, in order; the first failing one ends the pixel (the reads
    // are repeated on the way back rather than kept)
    auto normal_point = [&](int a) {
        const float d = (a & 1) ? -1.0f : 1.0f;
        const int ax = a >> 1;
        return v3{ax == 0 ? r2.x + d : r2.x, ax == 1 ? r2.y + d : r2.y, ax == 2 ? r2.z + d : r2.z};
    };
    float sv[6];
#pragma unroll 1
    for (int a = 0; a < 6; ++a) {
        TriRead tn;
        tri_read(m, normal_point(a), tn);
        sv[a] = tn.val;
        if (!tn.ok || sv[a] == 1.0f) return false;
    }
    const v3 gvec = mk(sv[0] - sv[1], sv[2] - sv[3], sv[4] - sv[5]);
    const v3 mvec = mv(pose.Rinv, gvec);
    const float mlen = norm(mvec);
    const v3 nrm = divs(mvec, mlen);
    const v3 vout = mv(pose.Rinv, sub(scale(r2, m.voxel), pose.t));
    // ---- backwards ----
    v3 gw{0.f, 0.f, 0.f}, gt{0.f, 0.f, 0.f}, gr2{0.f, 0.f, 0.f}, gdir{0.f, 0.f, 0.f};
    // normal: n = m / |m|, m = R'^T g
    const float ndg = dot(nrm, gn);
    const v3 gm = divs(sub(gn, scale(ndg, nrm)), mlen);
    gw = add(gw, cross(gm, mvec));
    const v3 gg = mv(pose.R, gm);
#pragma unroll 1
    for (int a = 0; a < 6; ++a) {
        const float ga = (a >> 1) == 0 ? gg.x : ((a >> 1) == 1 ? gg.y : gg.z);
        const float gsa = (a & 1) ? -ga : ga;
        TriRead tn;
        tri_read(m, normal_point(a), tn);
        tri_scatter(tn, gsa, grad_sdf);
        gr2 = add(gr2, scale(gsa, tn.grad));
    }
    // vertex: v = R'^T (voxel r2 - t')
    gr2 = add(gr2, scale(m.voxel, mv(pose.R, gv)));
    gt = sub(gt, gv);
    gw = add(gw, cross(gv, vout));
    // r2 = r1 + k s1 dir
    v3 gr1 = gr2;
    const float gs1 = ks * dot(ray.dir, gr2);
    gdir = add(gdir, scale(ks * s1, gr2));
    tri_scatter(t1, gs1, grad_sdf);
    gr1 = add(gr1, scale(gs1, t1.grad));
    // r1 = r0 + k s0 dir
    v3 gr0 = gr1;
    const float gs0 = ks * dot(ray.dir, gr1);
    gdir = add(gdir, scale(ks * s0, gr1));
    if (interp0) {
        tri_scatter(t0, gs0, grad_sdf);
        gr0 = add(gr0, scale(gs0, t0.grad));
    } else if (grad_sdf && k0 >= 0 && gs0 != 0.0f) {
        atomicAdd(grad_sdf + k0, gs0);
    }
    // r0 = block_s + M dir
    gdir = add(gdir, scale(msum, gr0));
    // dir = R' u, u the ray in the camera frame
    const v3 u = mv(pose.Rinv, ray.dir);
    gw = add(gw, cross(u, mv(pose.Rinv, gdir)));
    // block_s = (R' p_s + t') / voxel
    const float z0 = zmin.at(y >> 3, x >> 3);
    const v3 ps{z0 * ((float)x - g.cx) * g.invfx, z0 * ((float)y - g.cy) * g.invfy, z0};
    const v3 gb = scale(m.voxel_inv, mv(pose.Rinv, gr0));
    gt = add(gt, gb);
    gw = add(gw, cross(ps, gb));
    gxi[0] = gt.x;
    gxi[1] = gt.y;
    gxi[2] = gt.z;
    gxi[3] = gw.x;
    gxi[4] = gw.y;
    gxi[5] = gw.z;
    return true;
}

__global__ __launch_bounds__(kAdjThreads) void raycast_adjoint_kernel(MapDev m, View<float> zmin,
                                                                      View<float> zmax,
                                                                      PoseArg pa, AdjointArgs a) {
    const PoseF pose = load_pose(pa);
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += gridDim.x * blockDim.x) {
        const int px = a.pixels[i];
        if (px < 0 || px >= a.g.cols * a.g.rows) {  // outside the image: no gradient
            if (a.grad_px)
                for (int j = 0; j < 6; ++j) a.grad_px[6 * i + j] = 0.f;
            continue;
        }
        const int y = px / a.g.cols, x = px - y * a.g.cols;
        const v3 gv{a.grad_v[3 * i], a.grad_v[3 * i + 1], a.grad_v[3 * i + 2]};
        const v3 gn = a.grad_n ? v3{a.grad_n[3 * i], a.grad_n[3 * i + 1], a.grad_n[3 * i + 2]}
                               : v3{0.f, 0.f, 0.f};
        float gxi[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        if (!adjoint_pixel(m, zmin, zmax, pose, a.g, x, y, gv, gn, a.grad_sdf, gxi))
            for (int j = 0; j < 6; ++j) gxi[j] = 0.f;
#pragma unroll
        for (int j = 0; j < 6; ++j) acc[j] += (double)gxi[j];
        if (a.grad_px)
            for (int j = 0; j < 6; ++j) a.grad_px[6 * i + j] = gxi[j];
    }
    // fixed-order workgroup sum: rows through LDS, thread j sums column j
    __shared__ double tile[kAdjThreads][7];
#pragma unroll
    for (int j = 0; j < 6; ++j) tile[threadIdx.x][j] = acc[j];
    __syncthreads();
    if (threadIdx.x < 6) {
        double s = 0.0;
        for (int r = 0; r < kAdjThreads; ++r) s += tile[r][threadIdx.x];
        a.partials[threadIdx.x * kAdjMaxBlocks + blockIdx.x] = s;
    }
}

__global__ void adjoint_finish_kernel(const double* partials, int nb, double* grad_pose) {
    if (threadIdx.x < 6) {
        double s = 0.0;
        for (int b = 0; b < nb; ++b) s += partials[threadIdx.x * kAdjMaxBlocks + b];
        grad_pose[threadIdx.x] = s;
    }
}

}  // namespace

int launch_raycast_adjoint(MapDev m, View<float> zmin, View<float> zmax, PoseArg pose,
                           float invfx, float invfy, float cx, float cy, int cols, int rows,
                           const int32_t* pixels, int n, const float* grad_v,
                           const float* grad_n, float* grad_sdf, float* grad_px,
                           double* partials, double* grad_pose, hipStream_t s) {
    AdjointArgs a{RayArgs{invfx, invfy, cx, cy, cols, rows, 0, 0}, pixels, n, grad_v, grad_n,
                  grad_sdf, grad_px, partials};
    int nb = (int)div_up(n > 0 ? n : 1, kAdjThreads);
    if (nb > kAdjMaxBlocks) nb = kAdjMaxBlocks;
    raycast_adjoint_kernel<<<nb, kAdjThreads, 0, s>>>(m, zmin, zmax, pose, a);
    SICP_LAUNCH_CHECK("raycast_adjoint_kernel");
    adjoint_finish_kernel<<<1, 64, 0, s>>>(partials, nb, grad_pose);
    SICP_LAUNCH_CHECK("adjoint_finish_kernel");
    return SICP_OK;
}

size_t raycast_adjoint_scratch_bytes() { return sizeof(double) * (6 * kAdjMaxBlocks + 8); }

int launch_render_prepass(MapDev m, MapCounters* ctr, const sicp_block* visible, View<float> zmin,
                          View<float> zmax, PoseArg pose, sicp_intrinsics K, float dmax,
                          float dmin, const int* gate, hipStream_t s, bool zrange_cleared,
                          const uint32_t* fuse_counts) {
    int n = zmin.cols * zmin.rows;
    if (!zrange_cleared) {
        zrange_init_kernel<<<div_up(n, 256) > 0 ? div_up(n, 256) : 1, 256, 0, s>>>(zmin, zmax,
                                                                                  ctr);
        SICP_LAUNCH_CHECK("zrange_init_kernel");
    }
    PrepassArgs g{K.fx, K.fy, K.cx, K.cy, dmax, dmin, zmin.cols, zmin.rows};
    if (!gate) gate = m.one;
    if (n <= kPrepassLdsTiles) {
        prepass_lds_kernel<<<kPrepassLdsGrid, 256, 0, s>>>(m, ctr, visible, zmin, zmax, pose, g, gate,
                                               fuse_counts);
        SICP_LAUNCH_CHECK("prepass_lds_kernel");
    } else {
        prepass_kernel<<<kPrepassGrid, 64, 0, s>>>(m, ctr, visible, zmin, zmax, pose, g, gate,
                                           fuse_counts);
        SICP_LAUNCH_CHECK("prepass_kernel");
    }
    return SICP_OK;
}

// Both passes; the queue count is cleared first (unless the caller's
// launch_fuse_count already did).
int launch_raycast_gated(MapDev m, View<float4> vmap, View<float4> nmap, View<float> zmin,
                         View<float> zmax, PoseArg pose, float invfx, float invfy, float cx,
                         float cy, const int* gate, const uint32_t* nonempty, hipStream_t s,
                         bool queue_cleared, RayTuning tune) {
    if ((size_t)vmap.cols * vmap.rows > m.rq_cap) return invalid("raycast: image above the map's max_cols x max_rows");
    if (tune.first_pass_samples < 1 || tune.first_pass_samples > kRayTuneMaxSamples ||
        tune.second_pass_workgroups < 1 || tune.second_pass_workgroups > kRayTuneMaxWorkgroups ||
        (tune.lds_block_table != 0 && tune.lds_block_table != 1))
        return invalid("raycast: tuning outside 1..4096 samples, 1..65536 workgroups, table 0/1");
    RayArgs g{invfx, invfy, cx, cy, vmap.cols, vmap.rows, tune.first_pass_samples,
              tune.lds_block_table};
    if (!gate) gate = m.one;
    if (!nonempty) nonempty = reinterpret_cast<const uint32_t*>(m.one);
    if (!queue_cleared) {
        ray_queue_reset_kernel<<<1, 1, 0, s>>>(m.rq_count);
        SICP_LAUNCH_CHECK("ray_queue_reset_kernel");
    }
    dim3 grid(div_up(vmap.cols, 8), div_up(vmap.rows, 8));
    raycast_kernel<<<grid, 64, 0, s>>>(m, vmap, nmap, zmin, zmax, pose, g, gate, nonempty);
    SICP_LAUNCH_CHECK("raycast_kernel");
    raycast_long_kernel<<<tune.second_pass_workgroups, kRayLongThreads, 0, s>>>(
        m, vmap, nmap, zmin, zmax, pose, g, gate, nonempty);
    SICP_LAUNCH_CHECK("raycast_long_kernel");
    return SICP_OK;
}

int launch_raycast(MapDev m, View<float4> vmap, View<float4> nmap, View<float> zmin,
                   View<float> zmax, PoseArg pose, float invfx, float invfy, float cx, float cy,
                   const int* gate, hipStream_t s) {
    return launch_raycast_gated(m, vmap, nmap, zmin, zmax, pose, invfx, invfy, cx, cy, gate,
                                nullptr, s, false, RayTuning{});
}

}  // namespace sicp
