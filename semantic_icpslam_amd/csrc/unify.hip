// unify.hip — neighbour label smoothing after label fusion
// (Fusion::UnifyNeighbor / UpdateNeighborLabel, Mapping/FuseMap.cu:757-948),
// made deterministic, for gfx950.
//
// The reference reads neighbouring voxels while other threads rewrite them,
// so its result depends on timing. Here the pass is double-buffered: a
// snapshot kernel copies the label state (weight != 0, label, count) of
// every voxel of the blocks about to be rewritten, and stamps those blocks;
// the update kernel then reads every neighbour of a stamped block from the
// snapshot and every other neighbour from the map itself (which this pass
// does not write). Each voxel is written only by its own thread, after its
// own reads, so the outcome is that of a serial run in which every voxel sees
// the state before the pass.
//
// One workgroup of 64 threads per block: thread (x, y) walks z = 0..7 in
// order, keeping its neighbour tallies across the walk, as the reference's
// per-thread arrays do (they are declared outside its z loop). The 27 blocks
// around a block are looked up once into LDS.
#include "map_common.hpp"
#include "sicp_internal.hpp"

namespace sicp {

namespace {

// Snapshot word: bits 0-7 label, bit 8 weight != 0, bits 16-31 count.
SICP_DEV uint32_t pack_label(const MapDev& m, size_t k) {
    const sicp_voxel_sem& s = m.sem[k];
    return (uint32_t)s.label | (vox_weight(m, k) != 0 ? 0x100u : 0u) | ((uint32_t)s.count << 16);
}

__global__ __launch_bounds__(256) void unify_snapshot_kernel(MapDev m, const MapCounters* ctr,
                                                             const sicp_block* visible,
                                                             uint32_t* snap, uint32_t* mark,
                                                             uint32_t epoch, const int* gate) {
    if (gate && *gate == 0) return;
    const uint32_t nv = ctr->visible;
    for (uint32_t b = blockIdx.x; b < nv; b += gridDim.x) {
        const int slot = visible[b].slot;
        if (threadIdx.x == 0) mark[slot] = epoch;
        const size_t base = (size_t)slot * 512;
        for (int i = threadIdx.x; i < 512; i += blockDim.x)
            snap[base + i] = pack_label(m, base + i);
    }
}

// The 26 neighbour offsets in the order UnifyNeighbor visits them
// (FuseMap.cu:838-892): +x with (y, z) in {0, +y, -y, +z, -z, ++, +-, -+, --}
// order, then -x likewise, then x = 0: +y, +y+z, +y-z, -y, -y+z, -y-z, +z, -z.
SICP_DEV void neighbour_offset(int k, int& dx, int& dy, int& dz) {
    if (k < 18) {
        dx = k < 9 ? 1 : -1;
        const int j = k % 9;
        // (dy, dz) for j = 0..8
        dy = (j == 1 || j == 5 || j == 6) ? 1 : ((j == 2 || j == 7 || j == 8) ? -1 : 0);
        dz = (j == 3 || j == 5 || j == 7) ? 1 : ((j == 4 || j == 6 || j == 8) ? -1 : 0);
    } else {
        dx = 0;
        const int j = k - 18;
        dy = j < 3 ? 1 : (j < 6 ? -1 : 0);
        const int q = j < 6 ? j % 3 : j - 6 + 1;
        dz = q == 1 ? 1 : (q == 2 ? -1 : 0);
    }
}

// Floor of a float coordinate to the voxel lattice (make_int3(float3),
// Utility/VectorMath.h:36-42).
SICP_DEV int floor_f(float a) {
    int b = f2i(a);
    return ((float)b > a) ? b - 1 : b;
}

__global__ __launch_bounds__(64) void unify_kernel(MapDev m, const MapCounters* ctr,
                                                   const sicp_block* visible,
                                                   const uint32_t* snap, const uint32_t* mark,
                                                   uint32_t epoch, uint32_t count_mask,
                                                   const int* gate) {
    if (gate && *gate == 0) return;
    __shared__ int nslot[27];
    __shared__ int nsnap[27];
    const uint32_t nv = ctr->visible;
    const int tx = threadIdx.x & 7, ty = threadIdx.x >> 3;
    const float vs = m.voxel;
    for (uint32_t b = blockIdx.x; b < nv; b += gridDim.x) {
        const sicp_block blk = visible[b];
        if (threadIdx.x < 27) {
            const int t = threadIdx.x;
            const int s = map_lookup(m, blk.x + t % 3 - 1, blk.y + (t / 3) % 3 - 1,
                                     blk.z + t / 9 - 1);
            nslot[t] = s;
            nsnap[t] = s >= 0 && mark[s] == epoch;
        }
        __syncthreads();
        int n_label[4] = {-1, -1, -1, -1};
        int n_count[4] = {0, 0, 0, 0};
        int n_lc[4] = {0, 0, 0, 0};
        const size_t own = (size_t)blk.slot * 512;
        for (int z = 0; z < 8; ++z) {
            const int li = z * 64 + ty * 8 + tx;
            const uint32_t cw = snap[own + li];
            const int cur_label = (cw & 0x100u) ? (int)(cw & 0xffu) : -1;
            if (cur_label == -1) continue;
            const int vx = blk.x * 8 + tx, vy = blk.y * 8 + ty, vz = blk.z * 8 + z;
            // voxelPosToWorldPos (DeviceMap.cu:165-167), then +-VoxelSize
            const float px = (float)vx * vs, py = (float)vy * vs, pz = (float)vz * vs;
            const float step = vs * (float)1;
            for (int k = 0; k < 26; ++k) {
                int dx, dy, dz;
                neighbour_offset(k, dx, dy, dz);
                const float qx = dx > 0 ? px + step : (dx < 0 ? px - step : px);
                const float qy = dy > 0 ? py + step : (dy < 0 ? py - step : py);
                const float qz = dz > 0 ? pz + step : (dz < 0 ? pz - step : pz);
                // FindVoxel(float3, Voxel&) (DeviceMap.cu:83-95, 154-163)
                const int nx = floor_f(qx / vs), ny = floor_f(qy / vs), nz = floor_f(qz / vs);
                const int bx = (nx >> 3) - blk.x, by = (ny >> 3) - blk.y, bz = (nz >> 3) - blk.z;
                int slot, from_snap;
                if (bx >= -1 && bx <= 1 && by >= -1 && by <= 1 && bz >= -1 && bz <= 1) {
                    const int t = (bz + 1) * 9 + (by + 1) * 3 + (bx + 1);
                    slot = nslot[t];
                    from_snap = nsnap[t];
                } else {
                    slot = map_lookup(m, nx >> 3, ny >> 3, nz >> 3);
                    from_snap = slot >= 0 && mark[slot] == epoch;
                }
                if (slot < 0) continue;
                const size_t vi = (size_t)slot * 512 + ((nz & 7) << 6) + ((ny & 7) << 3) + (nx & 7);
                uint32_t w;
                if (from_snap)
                    w = snap[vi];
                else
                    w = pack_label(m, vi);
                // UpdateNeighborLabel (FuseMap.cu:757-797)
                if (!(w & 0x100u)) continue;
                const int tl = (int)(w & 0xffu), tc = (int)(w >> 16);
                for (int l = 0; l < 4; ++l) {
                    if (n_label[l] == tl) {
                        n_count[l]++;
                        n_lc[l] = n_lc[l] > tc ? n_lc[l] : tc;
                        break;
                    } else if (n_label[l] == -1) {
                        n_label[l] = tl;
                        n_count[l] = 1;
                        n_lc[l] = tc;
                        break;
                    }
                }
            }
            // the most frequent neighbour label with a non-zero count
            // (FuseMap.cu:903-917); the first of equals wins
            int nl = -1, nc = 0, nlc = 0;
            for (int j = 0; j < 4; ++j)
                if (n_label[j] != -1 && n_lc[j] != 0 && n_count[j] > nc) {
                    nl = n_label[j];
                    nc = n_count[j];
                    nlc = n_lc[j];
                }
            // FuseMap.cu:926-944; narrowing to the count's width
            sicp_voxel_sem& cur = m.sem[own + li];
            if (cur_label != nl) {
                const int cur_count = (int)(cw >> 16);
                int tmp;
                if (cur_label == (int)cur.label2)
                    tmp = nlc > (int)cur.count2 ? nlc : ((int)cur.count2 + 5);
                else
                    tmp = nlc;
                cur.label2 = (uint8_t)cur_label;
                cur.count2 = (uint16_t)cur_count;
                cur.label = (uint8_t)nl;
                cur.count = (uint16_t)((uint32_t)tmp & count_mask);
            } else {
                cur.count = (uint16_t)((uint32_t)nlc & count_mask);
            }
        }
        __syncthreads();
    }
}

}  // namespace

int launch_unify(sicp_map* map, const int* gate, hipStream_t s) {
    if (!map->unify_snap) {
        const size_t n = (size_t)map->cfg.num_blocks;
        hipError_t e = hipMalloc((void**)&map->unify_snap, n * 512 * sizeof(uint32_t));
        if (e == hipSuccess) e = hipMalloc((void**)&map->unify_mark, n * sizeof(uint32_t));
        if (e == hipSuccess) e = hipMemset(map->unify_mark, 0, n * sizeof(uint32_t));
        if (e != hipSuccess) return hip_fail(e, "hipMalloc(unify scratch)");
        map->owned.push_back(map->unify_snap);
        map->owned.push_back(map->unify_mark);
    }
    // Epochs distinguish the blocks stamped by this call from earlier ones;
    // the mark array starts at 0 and the first epoch is 1.
    const uint32_t epoch = ++map->unify_epoch;
    const uint32_t mask = map->cfg.max_label_count <= 255 ? 0xffu : 0xffffu;
    unify_snapshot_kernel<<<4096, 256, 0, s>>>(map->dev, map->ctr, map->visible, map->unify_snap,
                                                map->unify_mark, epoch, gate);
    SICP_LAUNCH_CHECK("unify_snapshot_kernel");
    unify_kernel<<<8192, 64, 0, s>>>(map->dev, map->ctr, map->visible, map->unify_snap,
                                      map->unify_mark, epoch, mask, gate);
    SICP_LAUNCH_CHECK("unify_kernel");
    return SICP_OK;
}

}  // namespace sicp
