// mesh.hip — marching-cubes surface extraction of the TSDF (MeshScene,
// Mapping/MeshScene.cu:4-356) for gfx950.
//
// One workgroup per allocated block, one thread per voxel. The 27 blocks
// around a block are resolved once into LDS, so every corner and
// neighbour read is one load. Two passes: the first counts each block's
// triangles, a device scan turns the counts into offsets, and the second
// recomputes each voxel and writes its triangles at its block's offset plus
// its rank within the block. The output order is therefore canonical
// (heap slot, then voxel index z*64 + y*8 + x, then the table's order),
// where the reference appends with an atomic counter in whatever order the
// threads arrive.
#include <cstdio>
#include <cstring>
#include <vector>

#include "map_common.hpp"
#include "sicp_internal.hpp"

namespace sicp {

// ------------------------------------------------------------- tables -----
// Edge j of a voxel cube (its two corners), in the reference's numbering
// (MeshScene.cu:200-291); corner k sits at (x, y, z) = (0,0,0), (1,0,0),
// (1,1,0), (0,1,0), then the same four at z = 1.
static const int kEdge[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                 {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

struct MeshTables {
    int32_t edge[256];      // bit j: edge j is cut
    int8_t tri[256][16];    // edge triples, -1 terminated
    int32_t nvert[256];     // 3 x triangles
};

// The triangle table is the reference's (the classic published
// marching-cubes table, Mapping/Constant.h:39-296), so the triangles of a
// voxel, and the corner order within each, are the reference's; it is held
// as one hexadecimal string of edge indices per sign case (mc_table.inc,
// written by gen_mc_table.py). The edge table (the cut edges: corners of
// different sign) and the vertex counts follow from it.
#include "mc_table.inc"

static void build_tables(MeshTables& t) {
    for (int cs = 0; cs < 256; ++cs) {
        int emask = 0;
        for (int j = 0; j < 12; ++j)
            if (((cs >> kEdge[j][0]) & 1) != ((cs >> kEdge[j][1]) & 1)) emask |= 1 << j;
        t.edge[cs] = emask;
        const char* row = kMcTriangles[cs];
        const int n = (int)strlen(row);
        for (int j = 0; j < 16; ++j) {
            const char ch = j < n ? row[j] : 0;
            t.tri[cs][j] = j < n ? (int8_t)(ch <= '9' ? ch - '0' : ch - 'a' + 10) : (int8_t)-1;
        }
        t.nvert[cs] = n;
    }
}

static const MeshTables& tables() {
    static MeshTables t = [] {
        MeshTables x;
        build_tables(x);
        return x;
    }();
    return t;
}

namespace {

constexpr int kMeshThreads = 512;  // one per voxel of a block

struct MeshOut {
    float* vertices;   // 9 floats per triangle
    float* normals;    // 9 floats per triangle
    uint8_t* colors;   // 9 bytes per triangle
    uint32_t max_triangles;
};

// The 27 blocks around block `b` (offsets -1..1 per axis) into LDS.
SICP_DEV void neighbour_slots(const MapDev& m, const sicp_block& b, int* nslot) {
    const int t = threadIdx.x;
    if (t < 27) {
        const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
        nslot[t] = map_lookup(m, b.x + dx, b.y + dy, b.z + dz);
    }
}

// A voxel at block-relative voxel coordinates in [-8, 15] (FindVoxel,
// DeviceMap.cu:97-103): the default voxel (NaN) when its block is missing.
SICP_DEV long voxel_at(const int* nslot, int x, int y, int z) {
    const int s = nslot[((z >> 3) + 1) * 9 + ((y >> 3) + 1) * 3 + ((x >> 3) + 1)];
    if (s < 0) return -1;
    return (long)s * 512 + ((z & 7) << 6) + ((y & 7) << 3) + (x & 7);
}
SICP_DEV float sdf_at(const MapDev& m, const int* nslot, int x, int y, int z) {
    const long k = voxel_at(nslot, x, y, z);
    return k >= 0 ? vox_sdf(m, (size_t)k) : qnanf();
}

// MeshEngine::interp (MeshScene.cu:156-164).
SICP_DEV float mc_interp(float v1, float v2) {
    if ((double)fabsf(0 - v1) < 1e-6) return 0;
    if ((double)fabsf(0 - v2) < 1e-6) return 1;
    if ((double)fabsf(v1 - v2) < 1e-6) return 0;
    return (0 - v1) / (v2 - v1);
}

struct Cell {
    float s[8];
    v3 n[8];
    uint8_t c[8][3];
    int cube;  // sign case, 0 when the voxel makes no triangles
};

// The corner coordinates on the device, as closed forms of the corner number.
SICP_DEV int d_corner(int k, int axis) {
    return axis == 0 ? ((k == 1 || k == 2 || k == 5 || k == 6) ? 1 : 0)
                     : (axis == 1 ? ((k == 2 || k == 3 || k == 6 || k == 7) ? 1 : 0) : (k >> 2));
}
SICP_DEV int d_edge(int e, int end) {
    // edges 0-3 and 4-7 run round the z = 0 and z = 1 faces, 8-11 are vertical
    if (e < 8) {
        const int base = e & 4, i = e & 3;
        return base + (end ? ((i + 1) & 3) : i);
    }
    return end ? (e - 8) + 4 : e - 8;
}

SICP_DEV int corner_index(int ox, int oy, int oz) {
    return oz * 4 + (oy ? (ox ? 2 : 3) : (ox ? 1 : 0));
}

// readVertexAndColor + readNormal + the case (MeshScene.cu:39-154,166-198):
// the eight corner values (none NaN or exactly 1), their colours, and at
// every corner the difference of the two voxels either side along each axis
// (a neighbouring corner's value where the neighbour is a corner); none of
// those neighbours NaN.
SICP_DEV void build_cell(const MapDev& m, const int* nslot, const MeshTables& tb, int x, int y,
                         int z, Cell& cl) {
    cl.cube = 0;
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int cx = d_corner(k, 0), cy = d_corner(k, 1), cz = d_corner(k, 2);
        const long q = voxel_at(nslot, x + cx, y + cy, z + cz);
        cl.s[k] = q >= 0 ? vox_sdf(m, (size_t)q) : qnanf();
        uint8_t r = 0, g = 0, b = 0;
        if (q >= 0) vox_rgb(m, (size_t)q, r, g, b);
        cl.c[k][0] = r;
        cl.c[k][1] = g;
        cl.c[k][2] = b;
        ok = ok && !(cl.s[k] == 1.0f || isnan_(cl.s[k]));
    }
    if (!ok) return;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int cx = d_corner(k, 0), cy = d_corner(k, 1), cz = d_corner(k, 2);
        float g[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int o = a == 0 ? cx : (a == 1 ? cy : cz);
            const int px = x + cx + (a == 0 ? (o ? 1 : -1) : 0);
            const int py = y + cy + (a == 1 ? (o ? 1 : -1) : 0);
            const int pz = z + cz + (a == 2 ? (o ? 1 : -1) : 0);
            const float far = sdf_at(m, nslot, px, py, pz);
            ok = ok && !isnan_(far);
            // the other corner along this axis
            const int kk = corner_index(a == 0 ? 1 - cx : cx, a == 1 ? 1 - cy : cy,
                                        a == 2 ? 1 - cz : cz);
            g[a] = o ? far - cl.s[kk] : cl.s[kk] - far;
        }
        cl.n[k] = v3{g[0], g[1], g[2]};
    }
    if (!ok) return;
    int cube = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) cube |= (cl.s[k] < 0) ? (1 << k) : 0;
    cl.cube = tb.edge[cube] == 0 ? 0 : cube;
}
// One vertex of a triangle, on cut edge `e` of the cell at voxel `pos`
// (MeshScene.cu:200-291 and :312-320): the position pos + the edge's corner
// offsets with the parameter, times the voxel size; the corner normals
// interpolated, then normalised; the colour of the edge's first corner,
// or of its second when the parameter is exactly 1 (the reference
// multiplies the uchar3 difference by the float parameter converted to int).
SICP_DEV void edge_vertex(const Cell& cl, int e, int px, int py, int pz, float voxel,
                          float* vo, float* no, uint8_t* co) {
    const int a = d_edge(e, 0), b = d_edge(e, 1);
    const float val = mc_interp(cl.s[a], cl.s[b]);
    const int p[3] = {px, py, pz};
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const int ca = d_corner(a, ax), cb = d_corner(b, ax);
        const float comp = ca == cb ? (float)ca : (ca == 0 ? val : 1 - val);
        vo[ax] = ((float)p[ax] + comp) * voxel;
    }
    const v3 na = cl.n[a], nb = cl.n[b];
    const v3 n = v3{na.x + val * (nb.x - na.x), na.y + val * (nb.y - na.y),
                    na.z + val * (nb.z - na.z)};
    const v3 nn = normalised(n);
    no[0] = nn.x;
    no[1] = nn.y;
    no[2] = nn.z;
    const int iv = (int)val;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        const uint8_t d = (uint8_t)(cl.c[b][ch] - cl.c[a][ch]);
        const uint8_t m = (uint8_t)(d * iv);
        co[ch] = (uint8_t)(cl.c[a][ch] + m);
    }
}

// Exclusive prefix of `v` over the workgroup (512 threads); `total` gets the
// sum. `part` is LDS for the eight wavefront totals.
SICP_DEV uint32_t block_exclusive_scan(uint32_t v, uint32_t* part, uint32_t& total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t up = __shfl_up(inc, o, 64);
        if (lane >= o) inc += up;
    }
    if (lane == 63) part[w] = inc;
    __syncthreads();
    uint32_t before = 0;
    total = 0;
#pragma unroll
    for (int k = 0; k < kMeshThreads / 64; ++k) {
        before += k < w ? part[k] : 0u;
        total += part[k];
    }
    return before + inc - v;
}

// Pass 1: the triangle count of every allocated block.
__global__ __launch_bounds__(kMeshThreads) void mesh_count_kernel(MapDev m,
                                                                  const MapCounters* ctr,
                                                                  const MeshTables* tb,
                                                                  uint32_t* tri_count) {
    __shared__ int nslot[27];
    __shared__ uint32_t part[kMeshThreads / 64];
    const uint32_t nblk = ctr->heap_used;
    const int t = threadIdx.x, x = t & 7, y = (t >> 3) & 7, z = t >> 6;
    for (uint32_t b = blockIdx.x; b < nblk; b += gridDim.x) {
        const sicp_block blk = m.blocks[b];
        neighbour_slots(m, blk, nslot);
        __syncthreads();
        Cell cl;
        build_cell(m, nslot, *tb, x, y, z, cl);
        const uint32_t n = cl.cube ? (uint32_t)tb->nvert[cl.cube] / 3u : 0u;
        uint32_t total;
        (void)block_exclusive_scan(n, part, total);
        if (t == 0) tri_count[b] = total;
        __syncthreads();
    }
}

// Pass 2: every voxel's triangles at its block's offset plus its rank
// within the block; triangles past `out.max_triangles` are dropped.
__global__ __launch_bounds__(kMeshThreads) void mesh_emit_kernel(MapDev m,
                                                                 const MapCounters* ctr,
                                                                 const MeshTables* tb,
                                                                 const uint32_t* tri_base,
                                                                 float voxel, MeshOut out) {
    __shared__ int nslot[27];
    __shared__ uint32_t part[kMeshThreads / 64];
    const uint32_t nblk = ctr->heap_used;
    const int t = threadIdx.x, x = t & 7, y = (t >> 3) & 7, z = t >> 6;
    for (uint32_t b = blockIdx.x; b < nblk; b += gridDim.x) {
        const sicp_block blk = m.blocks[b];
        neighbour_slots(m, blk, nslot);
        __syncthreads();
        Cell cl;
        build_cell(m, nslot, *tb, x, y, z, cl);
        const uint32_t n = cl.cube ? (uint32_t)tb->nvert[cl.cube] / 3u : 0u;
        uint32_t total;
        const uint32_t rank = block_exclusive_scan(n, part, total);
        const uint32_t base = tri_base[b] + rank;
        const int px = blk.x * 8 + x, py = blk.y * 8 + y, pz = blk.z * 8 + z;
        for (uint32_t j = 0; j < n; ++j) {
            const uint32_t tid = base + j;
            if (tid >= out.max_triangles) break;
#pragma unroll
            for (int v = 0; v < 3; ++v) {
                const int e = tb->tri[cl.cube][3 * j + v];
                float vo[3], no[3];
                uint8_t co[3];
                edge_vertex(cl, e, px, py, pz, voxel, vo, no, co);
                const size_t o = ((size_t)tid * 3 + v) * 3;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    out.vertices[o + a] = vo[a];
                    out.normals[o + a] = no[a];
                    out.colors[o + a] = co[a];
                }
            }
        }
        __syncthreads();
    }
}

}  // namespace

int launch_mesh(sicp_map* map, float* vertices, float* normals, uint8_t* colors,
                uint32_t max_triangles, uint32_t* n_out, uint32_t* total_out, hipStream_t s) {
    uint32_t nblk = 0;
    SICP_HIP(hipMemcpyAsync(&nblk, &map->ctr->heap_used, sizeof(nblk), hipMemcpyDeviceToHost, s));
    SICP_HIP(hipStreamSynchronize(s));
    if (nblk == 0) {
        *n_out = 0;
        if (total_out) *total_out = 0;
        return SICP_OK;
    }
    const uint32_t cap = (uint32_t)map->cfg.num_blocks;
    const size_t tb_bytes = (sizeof(MeshTables) + 255) & ~(size_t)255;
    const size_t bytes = tb_bytes + sizeof(uint32_t) * ((size_t)2 * cap + 4096 + 64 + 4);
    char* scratch = nullptr;
    SICP_HIP(hipMalloc((void**)&scratch, bytes));
    MeshTables* tb = (MeshTables*)scratch;
    uint32_t* tri_count = (uint32_t*)(scratch + tb_bytes);
    uint32_t* tri_base = tri_count + cap;
    uint32_t* partials = tri_base + cap;
    uint32_t* total = partials + 4096 + 64;
    int rc = SICP_OK;
    uint32_t tot = 0;
    do {
        hipError_t e = hipMemcpyAsync(tb, &tables(), sizeof(MeshTables), hipMemcpyHostToDevice, s);
        if (e != hipSuccess) {
            rc = hip_fail(e, "hipMemcpyAsync(mesh tables)");
            break;
        }
        const int grid = (int)(nblk < 16384u ? nblk : 16384u);
        mesh_count_kernel<<<grid, kMeshThreads, 0, s>>>(map->dev, map->ctr, tb, tri_count);
        if ((e = hipGetLastError()) != hipSuccess) {
            rc = hip_fail(e, "mesh_count_kernel");
            break;
        }
        if ((rc = launch_scan_u32(tri_count, tri_base, &map->ctr->heap_used, cap, partials, total,
                                  s)))
            break;
        MeshOut out{vertices, normals, colors, max_triangles};
        if (max_triangles > 0) {
            mesh_emit_kernel<<<grid, kMeshThreads, 0, s>>>(map->dev, map->ctr, tb, tri_base,
                                                           map->cfg.voxel_size, out);
            if ((e = hipGetLastError()) != hipSuccess) {
                rc = hip_fail(e, "mesh_emit_kernel");
                break;
            }
        }
        if ((e = hipMemcpyAsync(&tot, total, sizeof(tot), hipMemcpyDeviceToHost, s)) !=
                hipSuccess ||
            (e = hipStreamSynchronize(s)) != hipSuccess) {
            rc = hip_fail(e, "mesh total");
            break;
        }
    } while (false);
    (void)hipStreamSynchronize(s);
    (void)hipFree(scratch);
    if (rc) return rc;
    *n_out = tot < max_triangles ? tot : max_triangles;
    if (total_out) *total_out = tot;
    return SICP_OK;
}

}  // namespace sicp

using namespace sicp;

extern "C" {

int sicp_mesh_tables(int32_t* edge_table, int8_t* triangle_table, int32_t* vertex_table) {
    const MeshTables& t = tables();
    if (edge_table) std::memcpy(edge_table, t.edge, sizeof(t.edge));
    if (triangle_table) std::memcpy(triangle_table, t.tri, sizeof(t.tri));
    if (vertex_table) std::memcpy(vertex_table, t.nvert, sizeof(t.nvert));
    return SICP_OK;
}

int sicp_mesh_scene(sicp_map* map, float* vertices_dev, float* normals_dev, uint8_t* colors_dev,
                    uint32_t max_triangles, uint32_t* n_triangles, uint32_t* n_total,
                    void* stream) {
    if (!map || !n_triangles) return invalid("mesh_scene: null argument");
    if (max_triangles > 0 && (!vertices_dev || !normals_dev || !colors_dev))
        return invalid("mesh_scene: null output buffer");
    return launch_mesh(map, vertices_dev, normals_dev, colors_dev, max_triangles, n_triangles,
                       n_total, (hipStream_t)stream);
}

int sicp_mesh_write_ply(const char* path, const float* vertices, const float* normals,
                        const uint8_t* colors, uint32_t n_triangles) {
    if (!path || (n_triangles > 0 && (!vertices || !normals || !colors)))
        return invalid("mesh_write_ply: null argument");
    FILE* f = std::fopen(path, "w");
    if (!f) return invalid("mesh_write_ply: cannot open the file");
    const unsigned long long nv = 3ull * n_triangles;
    std::fprintf(f,
                 "ply\nformat ascii 1.0\nelement vertex %llu\nproperty float x\nproperty float y\n"
                 "property float z\nproperty float nx\nproperty float ny\nproperty float nz\n"
                 "property uchar red\nproperty uchar green\nproperty uchar blue\n"
                 "element face %u\nproperty list uchar uint vertex_indices\nend_header\n",
                 nv, n_triangles);
    for (unsigned long long i = 0; i < nv; ++i)
        std::fprintf(f, "%g %g %g %g %g %g %d %d %d\n", vertices[3 * i], vertices[3 * i + 1],
                     vertices[3 * i + 2], normals[3 * i], normals[3 * i + 1], normals[3 * i + 2],
                     (int)colors[3 * i], (int)colors[3 * i + 1], (int)colors[3 * i + 2]);
    for (uint32_t i = 0; i < n_triangles; ++i)
        std::fprintf(f, "3 %u %u %u\n", 3 * i, 3 * i + 1, 3 * i + 2);
    const bool ok = std::ferror(f) == 0;
    if (std::fclose(f) != 0 || !ok) return invalid("mesh_write_ply: write failed");
    return SICP_OK;
}

}  // extern "C"
