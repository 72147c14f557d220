// map.hip — voxel-block hash allocation, visibility compaction and
// TSDF / colour / label fusion for gfx950.
//
// Allocation (Fusion::CreateBlocks, Mapping/FuseMap.cu:92-132) is made
// deterministic. The outcome is that of a serial pixel-major run of the
// reference kernel: a new block's heap slot is its rank in the order in which
// that run would first request it. The pipeline:
//   mark    per pixel: march the +-T/2 band; every block missing from the table
//           goes into a per-frame request set (open addressing, CAS) together
//           with the smallest request id (pixel << 16 | step) that asked for
//           it (atomicMin), and the step is flagged in the pixel's mask;
//   count   per pixel: how many of its flagged steps are their block's first
//           request;
//   scan    exclusive scan of those counts -> the rank of every new block;
//   emit    per pixel: write new blocks at their ranks;
//   link    per new block: elect the smallest rank of each bucket (atomicMin)
//           and push the rank onto the bucket's pending list;
//   insert  per elected bucket: insert its pending blocks in rank order into
//           the two bucket entries, then the overflow chain (the overflow
//           entry of a block is the one at its own heap slot, so no counter
//           is shared between buckets); reset the per-frame state.
// Visibility (CheckFullVisibility, FuseMap.cu:134-168) is a scan over the
// allocated heap slots compacted in slot order.
// Fusion runs one 64-lane wavefront per visible 8^3 block: lane l owns the
// voxel column (l % 8, l / 8) and walks z, so each z step reads 64
// consecutive voxel records of each plane.
#include "map_common.hpp"

namespace sicp {

namespace {

constexpr uint64_t kEmptyKey = ~0ull;
constexpr unsigned long long kNoId = ~0ull;
constexpr int kScanThreads = 256;
constexpr int kScanPerThread = 4;
constexpr int kScanChunk = kScanThreads * kScanPerThread;  // 1024 elements per workgroup
constexpr int kScanMaxChunks = 4096;

SICP_DEV uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// ----------------------------------------------------------------- reset --
// ResetMap (FuseMap.cu:1251-1304): entries empty; the heap starts full. Voxels
// of a block are cleared when the block is created.
__global__ void reset_table_kernel(MapDev m, AllocScratch a, uint32_t n_req) {
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const uint4 empty = make_uint4(0u, 0u, 0xffffffffu, 0xffffffffu);
    for (size_t i = tid; i < 2 * (size_t)m.nb; i += stride) ((uint4*)m.table)[i] = empty;
    for (size_t i = tid; i < (size_t)m.num_blocks; i += stride) {
        ((uint4*)m.excess)[i] = empty;
        m.blocks[i] = sicp_block{0, 0, 0, -1};
    }
    for (size_t i = tid; i < (size_t)m.nb; i += stride) {
        a.bucket_leader[i] = 0x7fffffff;
        a.bucket_head[i] = -1;
    }
    for (size_t i = tid; i < n_req; i += stride) {
        a.req_keys[i] = kEmptyKey;
        a.req_first[i] = kNoId;
    }
}

__global__ void reset_counters_kernel(MapCounters* ctr) {
    MapCounters z{};
    *ctr = z;
}

// ------------------------------------------------------------ allocation --
struct MarchSetup {
    v3 pt, dir;
    int nsteps;
};

// Fusion::CreateBlocks up to the step loop (FuseMap.cu:98-123), with
// Fusion::unproject (:35-44).
SICP_DEV bool march_setup(const MapDev& m, const View<float>& depth, int x, int y,
                          const PoseF& pose, float invfx, float invfy, float cx, float cy,
                          MarchSetup& ms) {
    float z = depth.at(y, x);
    if (isnan_(z) || z < m.dmin || z > m.dmax) return false;
    float thresh = m.trunc / 2.0f;
    float z_near = fmaxf(m.dmin, z - thresh);
    float z_far = fminf(m.dmax, z + thresh);
    if (z_near >= z_far) return false;
    v3 cn{z_near * ((float)x - cx) * invfx, z_near * ((float)y - cy) * invfy, z_near};
    v3 cf{z_far * ((float)x - cx) * invfx, z_far * ((float)y - cy) * invfy, z_far};
    v3 pn = scale(add(mv(pose.R, cn), pose.t), m.voxel_inv);
    v3 pf = scale(add(mv(pose.R, cf), pose.t), m.voxel_inv);
    v3 dir = sub(pf, pn);
    float length = norm(dir);
    int nsteps = (int)ceil(2.0 * (double)length);
    ms.dir = divs(dir, (float)(nsteps - 1));
    ms.pt = pn;
    ms.nsteps = nsteps;
    return true;
}

SICP_DEV int req_find(const AllocScratch& a, uint64_t key) {
    uint32_t h = (uint32_t)mix64(key) & a.req_mask;
    for (uint32_t p = 0; p <= a.req_mask; ++p) {
        uint64_t k = a.req_keys[h];
        if (k == key) return (int)h;
        if (k == kEmptyKey) return -1;
        h = (h + 1) & a.req_mask;
    }
    return -1;
}

SICP_DEV int req_insert(const AllocScratch& a, uint64_t key) {
    uint32_t h = (uint32_t)mix64(key) & a.req_mask;
    for (uint32_t p = 0; p <= a.req_mask; ++p) {
        uint64_t k = a.req_keys[h];
        if (k == key) return (int)h;
        if (k == kEmptyKey) {
            unsigned long long old = atomicCAS((unsigned long long*)&a.req_keys[h],
                                               (unsigned long long)kEmptyKey,
                                               (unsigned long long)key);
            if (old == kEmptyKey || old == key) return (int)h;
        }
        h = (h + 1) & a.req_mask;
    }
    return -1;
}

struct AllocArgs {
    float invfx, invfy, cx, cy;
    int cols, rows;
};

__global__ __launch_bounds__(256) void alloc_mark_kernel(MapDev m, MapCounters* ctr,
                                                         AllocScratch a, View<float> depth,
                                                         PoseArg pa, AllocArgs g,
                                                         const int* gate) {
    if (gate && *gate == 0) return;
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= g.cols || y >= g.rows) return;
    const int i = y * g.cols + x;
    const PoseF pose = load_pose(pa);
    MarchSetup ms;
    unsigned long long mask = 0ull;
    if (march_setup(m, depth, x, y, pose, g.invfx, g.invfy, g.cx, g.cy, ms)) {
        i3 prev{0, 0, 0};
        uint32_t lost_range = 0, lost_req = 0;
        for (int s = 0; s < ms.nsteps; ++s) {
            i3 bp = voxel_to_block(floor_i3(ms.pt));
            ms.pt = add(ms.pt, ms.dir);
            if (s > 0 && bp.x == prev.x && bp.y == prev.y && bp.z == prev.z) continue;
            prev = bp;
            if (!key_in_range(bp.x, bp.y, bp.z)) {
                ++lost_range;
                continue;
            }
            if (map_lookup(m, bp.x, bp.y, bp.z) >= 0) continue;
            int slot = req_insert(a, pack_key(bp.x, bp.y, bp.z));
            if (slot < 0) {
                ++lost_req;
                continue;
            }
            atomicMin(&a.req_first[slot], ((unsigned long long)i << 16) | (unsigned)s);
            mask |= (s < 63) ? (1ull << s) : (1ull << 63);
        }
        if (lost_range) atomicAdd(&ctr->dropped_range, lost_range);
        if (lost_req) atomicAdd(&ctr->dropped_requests, lost_req);
    }
    a.pix_mask[i] = mask;
}

// Re-marches a pixel whose mask is non-zero and visits every flagged step
// that is the first request of its block. Step s >= 63 shares bit 63.
template <class F>
SICP_DEV void for_owned_steps(const MapDev& m, const AllocScratch& a, const View<float>& depth,
                              int x, int y, int i, unsigned long long mask, const PoseF& pose,
                              const AllocArgs& g, F&& f) {
    MarchSetup ms;
    if (!march_setup(m, depth, x, y, pose, g.invfx, g.invfy, g.cx, g.cy, ms)) return;
    i3 prev{0, 0, 0};
    for (int s = 0; s < ms.nsteps; ++s) {
        i3 bp = voxel_to_block(floor_i3(ms.pt));
        ms.pt = add(ms.pt, ms.dir);
        if (s > 0 && bp.x == prev.x && bp.y == prev.y && bp.z == prev.z) continue;
        prev = bp;
        unsigned long long bit = (s < 63) ? (1ull << s) : (1ull << 63);
        if (!(mask & bit)) continue;
        if (!key_in_range(bp.x, bp.y, bp.z)) continue;
        uint64_t key = pack_key(bp.x, bp.y, bp.z);
        int slot = req_find(a, key);
        if (slot < 0) continue;
        if (a.req_first[slot] != (((unsigned long long)i << 16) | (unsigned)s)) continue;
        f(key, slot);
    }
}

SICP_DEV uint32_t block_exclusive_scan(uint32_t v, uint32_t* total);
SICP_DEV uint2 chunk_offset(const uint32_t* sums, uint32_t w, uint32_t n);

// The request counts and keys of the owned steps run over raster chunks of
// 256 P pixels per workgroup (P per thread), the chunks of the exclusive
// scan that ranks the requests in raster order: the count kernel leaves
// each chunk's sum in `partials`, and the emit kernel forms every pixel's
// first rank itself (the sum of the earlier chunks plus the prefix within
// the chunk) and stores the total — no separate scan launches.
template <int P>
__global__ __launch_bounds__(256) void alloc_count_kernel(MapDev m, AllocScratch a,
                                                          View<float> depth, PoseArg pa,
                                                          AllocArgs g, uint32_t* partials,
                                                          const int* gate) {
    const bool off = gate && *gate == 0;
    const int n = g.cols * g.rows;
    const int base = (int)blockIdx.x * (kScanThreads * P) + (int)threadIdx.x * P;
    const PoseF pose = load_pose(pa);
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < P; ++k) {
        const int i = base + k;
        if (i >= n) continue;
        uint32_t cnt = 0;
        const unsigned long long mask = off ? 0ull : a.pix_mask[i];
        if (mask) {
            const int y = i / g.cols, x = i - y * g.cols;
            for_owned_steps(m, a, depth, x, y, i, mask, pose, g, [&](uint64_t, int) { ++cnt; });
        }
        a.pix_count[i] = cnt;
        s += cnt;
    }
    uint32_t tot;
    block_exclusive_scan(s, &tot);
    if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

template <int P>
__global__ __launch_bounds__(256) void alloc_emit_kernel(MapDev m, MapCounters* ctr,
                                                         AllocScratch a, View<float> depth,
                                                         PoseArg pa, AllocArgs g,
                                                         const uint32_t* partials,
                                                         uint32_t nchunks, const int* gate) {
    if (gate && *gate == 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) ctr->new_total = 0u;
        return;
    }
    const int n = g.cols * g.rows;
    const int base = (int)blockIdx.x * (kScanThreads * P) + (int)threadIdx.x * P;
    uint32_t c[P], s = 0;
#pragma unroll
    for (int k = 0; k < P; ++k) {
        c[k] = base + k < n ? a.pix_count[base + k] : 0u;
        s += c[k];
    }
    const uint2 off = chunk_offset(partials, blockIdx.x, nchunks);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctr->new_total = off.y;
        // the first slot of this frame's new blocks, kept apart from
        // heap_used, which the insertion advances while others still read it
        ctr->alloc_base = ctr->heap_used;
    }
    uint32_t tot;
    uint32_t r = off.x + block_exclusive_scan(s, &tot);
    const PoseF pose = load_pose(pa);
#pragma unroll
    for (int k = 0; k < P; ++k) {
        const int i = base + k;
        if (c[k] == 0) continue;
        const unsigned long long mask = a.pix_mask[i];
        const int y = i / g.cols, x = i - y * g.cols;
        for_owned_steps(m, a, depth, x, y, i, mask, pose, g, [&](uint64_t key, int slot) {
            a.new_keys[r] = key;
            a.new_req_slot[r] = (uint32_t)slot;
            ++r;
        });
    }
}

// Clears the voxels of the blocks this frame creates (empty sdf, weight 0,
// zero semantic record: Voxel::release, Mapping/DeviceMap.h:76-87), over
// `nwg` workgroups starting at `wg`.
SICP_DEV void clear_new_blocks(const MapDev& m, const MapCounters* ctr, uint32_t wg,
                               uint32_t nwg) {
    const uint32_t total = ctr->new_total;
    const uint32_t base = ctr->alloc_base;
    uint32_t created = 0;
    if (base < (uint32_t)m.num_blocks) created = min(total, (uint32_t)m.num_blocks - base);
    const uint2 g = make_uint2(0x7fffffffu, 0u);
    const uint3 z = make_uint3(0u, 0u, 0u);
    for (uint32_t r = wg; r < created; r += nwg) {
        size_t off = (size_t)(base + r) * kBlockVoxels;
        for (int v = threadIdx.x; v < kBlockVoxels; v += blockDim.x) {
            if (m.half)
                m.geom_h[off + v] = kEmptyHalfVoxel;
            else
                ((uint2*)m.geom)[off + v] = g;
            ((uint3*)m.sem)[off + v] = z;
        }
    }
}

// Workgroups below kLinkBlocks link the new requests into their buckets'
// pending lists and write the block records; the others clear the new
// blocks' voxels (disjoint memory, so one launch serves both).
constexpr int kLinkBlocks = 512, kClearBlocks = 1024;
__global__ __launch_bounds__(256) void alloc_link_kernel(MapDev m, const MapCounters* ctr,
                                                         AllocScratch a, const int* gate) {
    if (gate && *gate == 0) return;
    if (blockIdx.x >= kLinkBlocks) {
        clear_new_blocks(m, ctr, blockIdx.x - kLinkBlocks, gridDim.x - kLinkBlocks);
        return;
    }
    const uint32_t total = ctr->new_total;
    const uint32_t base = ctr->alloc_base;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < total;
         r += kLinkBlocks * blockDim.x) {
        uint32_t slot = base + r;
        if (slot >= (uint32_t)m.num_blocks) continue;  // heap exhausted
        i3 p = unpack_key(a.new_keys[r]);
        uint32_t b = bucket_of(p.x, p.y, p.z, m);
        atomicMin(&a.bucket_leader[b], (int)r);
        int prev = atomicExch(&a.bucket_head[b], (int)r);
        a.pending_next[r] = prev;
        m.blocks[slot] = sicp_block{p.x, p.y, p.z, (int)slot};
    }
}

SICP_DEV void store_entry(sicp_hash_entry* e, uint64_t key, int slot, int next) {
    *(uint4*)e = make_uint4((uint32_t)key, (uint32_t)(key >> 32), (uint32_t)slot, (uint32_t)next);
}

// The blocks this frame creates: the heap's room bounds them; the rest are
// counted as dropped.
SICP_DEV uint32_t blocks_created(const MapDev& m, uint32_t total, uint32_t base) {
    const uint32_t room = base < (uint32_t)m.num_blocks ? (uint32_t)m.num_blocks - base : 0u;
    return min(total, room);
}

// The insertion of the frame's new blocks into the table, over the launch's
// threads (grid-stride, 256 per workgroup). Workgroup 0's first thread also
// advances the heap by the blocks created, which no thread of the launch
// reads: the ranks' slots start at alloc_base.
SICP_DEV void insert_ranks(const MapDev& m, MapCounters* ctr, const AllocScratch& a,
                           uint32_t total, uint32_t base) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const uint32_t created = blocks_created(m, total, base);
        ctr->dropped_heap += total - created;
        ctr->heap_used = base + created;
        ctr->new_blocks = created;
    }
    for (uint32_t r = blockIdx.x * 256 + threadIdx.x; r < total; r += gridDim.x * 256) {
        // Every rank clears its request-set slot for the next frame.
        uint32_t rs = a.new_req_slot[r];
        a.req_keys[rs] = kEmptyKey;
        a.req_first[rs] = kNoId;
        uint32_t slot = base + r;
        if (slot >= (uint32_t)m.num_blocks) continue;
        uint64_t key = a.new_keys[r];
        i3 p = unpack_key(key);
        uint32_t b = bucket_of(p.x, p.y, p.z, m);
        if (a.bucket_leader[b] != (int)r) continue;
        // The bucket's leader inserts its pending ranks in ascending order.
        sicp_hash_entry* e0 = m.table + 2 * (size_t)b;
        sicp_hash_entry* e1 = e0 + 1;
        int head = a.bucket_head[b];
        int last_done = -1;
        for (;;) {
            int best = 0x7fffffff;
            for (int q = head; q >= 0; q = a.pending_next[q])
                if (q > last_done && q < best) best = q;
            if (best == 0x7fffffff) break;
            last_done = best;
            uint64_t k = a.new_keys[best];
            int s = (int)(base + (uint32_t)best);
            EntryView v0 = load_entry(e0);
            EntryView v1 = load_entry(e1);
            if (v0.slot < 0) {
                store_entry(e0, k, s, v0.next);
            } else if (v1.slot < 0) {
                store_entry(e1, k, s, v1.next);
            } else {
                store_entry(m.excess + s, k, s, -1);
                if (v1.next < 0) {
                    store_entry(e1, v1.key, v1.slot, s);
                } else {
                    int n = v1.next;
                    EntryView x = load_entry(m.excess + n);
                    while (x.next >= 0) {
                        n = x.next;
                        x = load_entry(m.excess + n);
                    }
                    store_entry(m.excess + n, x.key, x.slot, s);
                }
            }
        }
        a.bucket_leader[b] = 0x7fffffff;
        a.bucket_head[b] = -1;
    }
}

// ------------------------------------------------------------------ scan --
// Exclusive scan of one value per thread over a 256-thread workgroup.
SICP_DEV uint32_t block_exclusive_scan(uint32_t v, uint32_t* total) {
    __shared__ uint32_t wave_tot[kScanThreads / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t inc = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t t = __shfl_up(inc, off, 64);
        if (lane >= off) inc += t;
    }
    if (lane == 63) wave_tot[wave] = inc;
    __syncthreads();
    uint32_t wbase = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / 64; ++w) {
        if (w < wave) wbase += wave_tot[w];
        tot += wave_tot[w];
    }
    __syncthreads();
    *total = tot;
    return wbase + inc - v;
}

// The sum of the chunk sums before chunk `w`, and of all `n` of them, formed
// by the whole workgroup (integer sums: any order gives the same values).
// Every workgroup of a pass forms its own offset, which costs a few loads
// per thread instead of a launch that scans the chunk sums.
SICP_DEV uint2 chunk_offset(const uint32_t* sums, uint32_t w, uint32_t n) {
    __shared__ uint32_t red[2][kScanThreads / 64];
    uint32_t a = 0, b = 0;
    for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) {
        const uint32_t v = sums[j];
        a += j < w ? v : 0u;
        b += v;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
    }
    __syncthreads();  // `red` may still be read by an earlier call
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = a;
        red[1][threadIdx.x >> 6] = b;
    }
    __syncthreads();
    uint2 r = make_uint2(0u, 0u);
#pragma unroll
    for (int k = 0; k < kScanThreads / 64; ++k) {
        r.x += red[0][k];
        r.y += red[1][k];
    }
    return r;
}

__global__ __launch_bounds__(kScanThreads) void scan_reduce_kernel(const uint32_t* in,
                                                                   const uint32_t* n_dev,
                                                                   uint32_t cap,
                                                                   uint32_t* partials) {
    const uint32_t n = n_dev ? min(*n_dev, cap) : cap;
    const uint32_t base = blockIdx.x * kScanChunk + threadIdx.x * kScanPerThread;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanPerThread; ++k)
        if (base + k < n) s += in[base + k];
    uint32_t tot;
    block_exclusive_scan(s, &tot);
    if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void scan_partials_kernel(uint32_t* partials, uint32_t nchunks,
                                                             uint32_t* total_out) {
    __shared__ uint32_t wave_tot[16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t v[4], s = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        uint32_t idx = threadIdx.x * 4 + k;
        v[k] = idx < nchunks ? partials[idx] : 0u;
        s += v[k];
    }
    uint32_t inc = s;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t t = __shfl_up(inc, off, 64);
        if (lane >= off) inc += t;
    }
    if (lane == 63) wave_tot[wave] = inc;
    __syncthreads();
    uint32_t wbase = 0, tot = 0;
    for (int w = 0; w < 16; ++w) {
        if (w < wave) wbase += wave_tot[w];
        tot += wave_tot[w];
    }
    uint32_t run = wbase + inc - s;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        uint32_t idx = threadIdx.x * 4 + k;
        if (idx < nchunks) partials[idx] = run;
        run += v[k];
    }
    if (threadIdx.x == 0) *total_out = tot;
}

__global__ __launch_bounds__(kScanThreads) void scan_apply_kernel(const uint32_t* in,
                                                                  uint32_t* out,
                                                                  const uint32_t* n_dev,
                                                                  uint32_t cap,
                                                                  const uint32_t* partials) {
    const uint32_t n = n_dev ? min(*n_dev, cap) : cap;
    const uint32_t base = blockIdx.x * kScanChunk + threadIdx.x * kScanPerThread;
    uint32_t v[kScanPerThread], s = 0;
#pragma unroll
    for (int k = 0; k < kScanPerThread; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t run = partials[blockIdx.x] + block_exclusive_scan(s, &tot);
#pragma unroll
    for (int k = 0; k < kScanPerThread; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
}

// ------------------------------------------------------------ visibility --
struct VisArgs {
    float fx, fy, cx, cy;
    int cols, rows;
    float dmin, dmax;
};

// Fusion::CheckVertexVisibility (FuseMap.cu:46-56).
SICP_DEV bool vertex_visible(v3 p3, const PoseF& pose, const VisArgs& g) {
    v3 pt = mv(pose.Rinv, sub(p3, pose.t));
    if (pt.z < 1e-3f) return false;
    float px = g.fx * pt.x / pt.z + g.cx;
    float py = g.fy * pt.y / pt.z + g.cy;
    return px >= 0 && py >= 0 && px < (float)g.cols && py < (float)g.rows && pt.z >= g.dmin &&
           pt.z <= g.dmax;
}

// Fusion::CheckBlockVisibility (FuseMap.cu:58-90): corners visited by the
// reference's running additions / subtractions of the block width.
SICP_DEV bool block_visible(const sicp_block& b, float s, const PoseF& pose, const VisArgs& g) {
    v3 c{(float)b.x * s, (float)b.y * s, (float)b.z * s};
    if (vertex_visible(c, pose, g)) return true;
    c.z += s;
    if (vertex_visible(c, pose, g)) return true;
    c.y += s;
    if (vertex_visible(c, pose, g)) return true;
    c.x += s;
    if (vertex_visible(c, pose, g)) return true;
    c.z -= s;
    if (vertex_visible(c, pose, g)) return true;
    c.y -= s;
    if (vertex_visible(c, pose, g)) return true;
    c.x -= s;
    c.y += s;
    if (vertex_visible(c, pose, g)) return true;
    c.x += s;
    c.y -= s;
    c.z += s;
    if (vertex_visible(c, pose, g)) return true;
    return false;
}

// P heap slots per thread (a chunk of 256 P slots per workgroup): one slot
// per thread spreads the allocated heap over four times the workgroups,
// when the partial sums of the smaller chunks fit the scan.
// The visibility flags of workgroup blockIdx.x's chunk of heap slots below
// n, and the chunk's count.
template <int P>
SICP_DEV void flag_chunk(const MapDev& m, const AllocScratch& a, const PoseF& pose,
                         const VisArgs& g, uint32_t n, uint32_t* partials) {
    const uint32_t base = blockIdx.x * (kScanThreads * P) + threadIdx.x * P;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < P; ++k) {
        uint32_t slot = base + k;
        if (slot < n) {
            uint8_t f = block_visible(m.blocks[slot], m.block_width, pose, g) ? 1 : 0;
            a.vis_flags[slot] = f;
            s += f;
        }
    }
    uint32_t tot;
    block_exclusive_scan(s, &tot);
    if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

template <int P>
__global__ __launch_bounds__(kScanThreads) void vis_flag_kernel(MapDev m, const MapCounters* ctr,
                                                                AllocScratch a, PoseArg pa,
                                                                VisArgs g, uint32_t* partials,
                                                                const int* gate) {
    const bool off = gate && *gate == 0;
    const uint32_t n = off ? 0u : ctr->heap_used;
    flag_chunk<P>(m, a, load_pose(pa), g, n, partials);
}

// The allocation's last launch: the insertion of the new blocks and, in the
// same launch, the visibility flags and chunk counts of the heap as it will
// stand (workgroups below nchunks; the heap count is formed from the
// frame's first slot and the blocks created, the value the insertion
// stores), so the visibility pass that follows only compacts.
template <int P>
__global__ __launch_bounds__(256) void alloc_insert_flag_kernel(MapDev m, MapCounters* ctr,
                                                                AllocScratch a, const int* gate,
                                                                PoseArg pa, VisArgs g,
                                                                uint32_t* partials,
                                                                uint32_t nchunks) {
    static_assert(kScanThreads == 256, "one launch shape for both parts");
    const bool off = gate && *gate == 0;
    const uint32_t total = ctr->new_total;
    const uint32_t base = ctr->alloc_base;  // stale when off: not used then
    if (off) {
        if (blockIdx.x == 0 && threadIdx.x == 0) ctr->new_blocks = 0;
    } else {
        insert_ranks(m, ctr, a, total, base);
    }
    if (blockIdx.x < nchunks) {
        const uint32_t n = off ? 0u : base + blocks_created(m, total, base);
        flag_chunk<P>(m, a, load_pose(pa), g, n, partials);
    }
}

// Also stores the visible count (the sum of every chunk's count; 0 on a lost
// frame). Workgroups past the allocated heap leave at once.
template <int P>
__global__ __launch_bounds__(kScanThreads) void vis_compact_kernel(MapDev m, MapCounters* ctr,
                                                                   AllocScratch a,
                                                                   sicp_block* visible,
                                                                   const uint32_t* partials,
                                                                   uint32_t nchunks,
                                                                   const int* gate) {
    const bool off = gate && *gate == 0;
    const uint32_t n = off ? 0u : ctr->heap_used;
    if (blockIdx.x > 0 && blockIdx.x * (uint32_t)(kScanThreads * P) >= n) return;
    const uint2 co = chunk_offset(partials, blockIdx.x, nchunks);
    if (blockIdx.x == 0 && threadIdx.x == 0) ctr->visible = off ? 0u : co.y;
    const uint32_t base = blockIdx.x * (kScanThreads * P) + threadIdx.x * P;
    uint32_t f[P], s = 0;
#pragma unroll
    for (int k = 0; k < P; ++k) {
        f[k] = (base + k < n) ? a.vis_flags[base + k] : 0u;
        s += f[k];
    }
    uint32_t tot;
    uint32_t run = co.x + block_exclusive_scan(s, &tot);
#pragma unroll
    for (int k = 0; k < P; ++k) {
        if (f[k]) visible[run++] = m.blocks[base + k];
    }
}

// ---------------------------------------------------------------- fusion --
struct FuseArgs {
    float fx, fy, cx, cy;
    int cols, rows;
    int num_detection;
};

SICP_DEV int palette_channel(int c, int L) {
    const int pal = c == 0 ? (1 << 25) - 1 : (c == 1 ? (1 << 15) - 1 : (1 << 21) - 1);
    int prod = (int)((uint32_t)pal * (uint32_t)L);
    return prod % 255;
}

SICP_DEV uint8_t u8i(float v) { return (uint8_t)(uint32_t)(int)v; }

// Fusion::createVoxels (FuseMap.cu:432-507) and, when num_detection > 0,
// Fusion::analyzeSemantics (FuseMap.cu:509-668) applied once to the same voxel.
__global__ __launch_bounds__(256) void fuse_kernel(MapDev m, const MapCounters* ctr,
                                                   const sicp_block* visible, View<float> depth,
                                                   View<uint8_t> rgb, View<float4> nmap,
                                                   View<uint8_t> mask, PoseArg pa, FuseArgs g,
                                                   uint32_t* wave_counts, const int* gate) {
    if (gate && *gate == 0) return;
    const uint32_t nvis = ctr->visible;
    const PoseF pose = load_pose(pa);
    const int lane = threadIdx.x & 63;
    const int lx = lane & 7, ly = lane >> 3;
    const uint32_t wave0 = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const uint32_t nwaves = gridDim.x * (blockDim.x >> 6);
    const float T = m.trunc;
    uint32_t updated = 0;
    // Grid-stride over the visible blocks; the next block's record is loaded
    // while the current block is fused.
    sicp_block nblk = wave0 < nvis ? visible[wave0] : sicp_block{};
    for (uint32_t vi = wave0; vi < nvis; vi += nwaves) {
        const sicp_block blk = nblk;
        if (vi + nwaves < nvis) nblk = visible[vi + nwaves];
        const size_t vbase = (size_t)blk.slot * kBlockVoxels + lane;
        const int bx = blk.x * kBlockSide + lx, by = blk.y * kBlockSide + ly;
        const int bz = blk.z * kBlockSide;
        for (int lz = 0; lz < kBlockSide; ++lz) {
            v3 pw{(float)bx * m.voxel, (float)by * m.voxel, (float)(bz + lz) * m.voxel};
            v3 pc = mv(pose.Rinv, sub(pw, pose.t));
            int u = (int)(g.fx * pc.x / pc.z + g.cx);
            int v = (int)(g.fy * pc.y / pc.z + g.cy);
            if (u < 0 || v < 0 || u >= g.cols || v >= g.rows) continue;
            float dp = depth.at(v, u);
            if (isnan_(dp) || dp > m.dmax || dp < m.dmin) continue;
            float sdf = dp - pc.z;
            if (!(sdf >= -T)) continue;
            sdf = fminf(1.0f, sdf / T);
            const uint8_t* c = &rgb.at(v, 3 * u);
            const uint8_t c0 = c[0], c1 = c[1], c2 = c[2];
            const size_t kv = vbase + (size_t)lz * 64;
            sicp_voxel_sem* sp = m.sem + kv;
            sicp_voxel_geom gv;
            sicp_voxel_sem sh;  // half form: the whole semantic record (it holds the colour)
            if (m.half) {
                const uint32_t w = m.geom_h[kv];
                sh = *sp;
                gv.sdf = h2f(w);
                gv.weight = (uint8_t)((w >> 16) & 0xffu);
                gv.r = sh.pad[0];
                gv.g = sh.pad[1];
                gv.b = sh.pad[2];
            } else {
                gv = m.geom[kv];
            }
            ++updated;
            bool new_voxel = gv.weight == 0;
            if (new_voxel) {
                gv.sdf = sdf;
                gv.weight = 1;
                gv.r = c0;
                gv.g = c1;
                gv.b = c2;
            }
            sicp_voxel_sem sv;
            const bool semantics = g.num_detection > 0;
            if (new_voxel) {
                sv = sicp_voxel_sem{};
            } else if (m.half) {
                sv = sh;
            } else if (semantics) {
                sv = *sp;
            } else {
                sv.label = 0;  // only the label is read below; loaded lazily
                sv.label = ((const uint8_t*)sp)[4];
            }
            if (!new_voxel) {
                int w = gv.weight;
                if (sv.label == 0) {
                    gv.r = u8i(((float)(gv.r * w) + (float)c0) / (float)(w + 1));
                    gv.g = u8i(((float)(gv.g * w) + (float)c1) / (float)(w + 1));
                    gv.b = u8i(((float)(gv.b * w) + (float)c2) / (float)(w + 1));
                }
                gv.sdf = (gv.sdf * (float)w + sdf) / (float)(w + 1);
                gv.weight = (uint8_t)min(m.max_weight, w + 1);
            }
            if (semantics) {
                float4 nl = nmap.at(v, u);
                if (!isnan_(nl.x)) {
                    const int L = mask.at(v, u);
                    float val[3];
                    if (L == 0) {
                        val[0] = (float)c0;
                        val[1] = (float)c1;
                        val[2] = (float)c2;
                    } else {
                        val[0] = (float)palette_channel(0, L);
                        val[1] = (float)palette_channel(1, L);
                        val[2] = (float)palette_channel(2, L);
                    }
                    if (L != 0) {
                        if (L == sv.label) {
                            sv.count = (uint16_t)min(m.max_count, (int)sv.count + 1);
                            gv.r = u8i(val[0]);
                            gv.g = u8i(val[1]);
                            gv.b = u8i(val[2]);
                        } else {
                            if (L == sv.label2) {
                                sv.count2 = (uint16_t)min(m.max_count, (int)sv.count2 + 1);
                                sv.r2 = u8i(val[0]);
                                sv.g2 = u8i(val[1]);
                                sv.b2 = u8i(val[2]);
                            } else if (sv.count2 < 7) {
                                sv.label2 = (uint8_t)L;
                                sv.count2 = 1;
                                sv.r2 = u8i(val[0]);
                                sv.g2 = u8i(val[1]);
                                sv.b2 = u8i(val[2]);
                            }
                            if (((int)sv.count < (int)sv.count2 && sv.label != 0) ||
                                ((int)sv.count < (int)sv.count2 * 2 && sv.label == 0)) {
                                uint8_t tl = sv.label;
                                sv.label = sv.label2;
                                sv.label2 = tl;
                                uint16_t tc = sv.count;
                                sv.count = sv.count2;
                                sv.count2 = tc;
                                uint8_t t0 = gv.r, t1 = gv.g, t2 = gv.b;
                                gv.r = sv.r2;
                                gv.g = sv.g2;
                                gv.b = sv.b2;
                                sv.r2 = t0;
                                sv.g2 = t1;
                                sv.b2 = t2;
                            }
                        }
                    } else {
                        // w = nl . normalised((pc, 1)) over four components (FuseMap.cu:641)
                        float n4 = sqrtf(pc.x * pc.x + pc.y * pc.y + pc.z * pc.z + 1.0f * 1.0f);
                        float q0 = pc.x / n4, q1 = pc.y / n4, q2 = pc.z / n4, q3 = 1.0f / n4;
                        float wv = nl.x * q0 + nl.y * q1 + nl.z * q2 + nl.w * q3;
                        if (sv.label != 0) {
                            if (sv.count < 7) {
                                sv.label2 = sv.label;
                                sv.r2 = gv.r;
                                sv.g2 = gv.g;
                                sv.b2 = gv.b;
                                sv.count2 = sv.count;
                                sv.label = 0;
                                sv.count = 1;
                                float aw = wv * 0.2f;
                                float o0 = (float)gv.r / 255.f, o1 = (float)gv.g / 255.f,
                                      o2 = (float)gv.b / 255.f;
                                float r0 = (aw * (val[0] / 255.f) + (1.0f - aw) * o0) * 255.f;
                                float r1 = (aw * (val[1] / 255.f) + (1.0f - aw) * o1) * 255.f;
                                float r2 = (aw * (val[2] / 255.f) + (1.0f - aw) * o2) * 255.f;
                                gv.r = u8i(r0);
                                gv.g = u8i(r1);
                                gv.b = u8i(r2);
                            }
                        } else {
                            sv.count = (uint16_t)min(m.max_count, (int)sv.count + 1);
                        }
                    }
                }
            }
            if (m.half) {
                // the sdf rounded to half on the store; the colour in the spare bytes
                sv.pad[0] = gv.r;
                sv.pad[1] = gv.g;
                sv.pad[2] = gv.b;
                *sp = sv;
                m.geom_h[kv] = f2h(gv.sdf) | ((uint32_t)gv.weight << 16);
            } else {
                if (semantics || new_voxel) *sp = sv;
                m.geom[kv] = gv;
            }
        }
    }
    // One count per wavefront into its own slot (a single device-wide counter
    // would serialise thousands of atomics on one address).
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) updated += __shfl_xor(updated, off, 64);
    if (lane == 0) wave_counts[wave0] = updated;
}

// Sums the per-wavefront counts of fuse_kernel into the map's counters; a
// gated-off frame fused nothing.
// Workgroup 0: the sum of the fusion's per-wavefront counts. With `rr.zmin`
// set, the other workgroups also clear the raycast's zRange images and
// workgroup 0 its tile and queue counters — work the prepass and the raycast
// would otherwise each start with a launch of their own.
struct RenderReset {
    View<float> zmin, zmax;
    uint32_t* rq_count;
};
__global__ __launch_bounds__(1024) void fuse_end_kernel(MapCounters* ctr,
                                                        const uint32_t* wave_counts, int n,
                                                        const int* gate, RenderReset rr) {
    if (blockIdx.x > 0) {
        const int nz = rr.zmin.cols * rr.zmin.rows;
        for (int i = (blockIdx.x - 1) * blockDim.x + threadIdx.x; i < nz;
             i += (gridDim.x - 1) * blockDim.x) {
            const int y = i / rr.zmin.cols, x = i - y * rr.zmin.cols;
            rr.zmin.at(y, x) = 100.f;
            rr.zmax.at(y, x) = 0.f;
        }
        return;
    }
    if (rr.zmin.ptr && threadIdx.x == 0) {
        ctr->render_tiles = 0u;
        *rr.rq_count = 0u;
    }
    if (gate && *gate == 0) {
        if (threadIdx.x == 0) ctr->updated_voxels = 0u;
        return;
    }
    __shared__ uint32_t part[16];
    uint32_t s = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += wave_counts[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += part[w];
        ctr->updated_voxels = t;
        ctr->updated_total += t;
        ctr->visible_total += ctr->visible;
        ctr->heap_total += ctr->heap_used;
        ctr->fusions += 1ull;
    }
}

}  // namespace

// ------------------------------------------------------------- launchers --
// The visibility pass's chunks of the heap: one slot per thread while the
// chunk sums fit the scan (returns true), else kScanPerThread.
static bool vis_chunks(const MapDev& m, uint32_t& nchunks) {
    const uint32_t nsmall = div_up(m.num_blocks, kScanThreads);
    if (nsmall <= (uint32_t)kScanMaxChunks) {
        nchunks = nsmall > 0 ? nsmall : 1;
        return true;
    }
    nchunks = div_up(m.num_blocks, kScanChunk);
    if (nchunks == 0) nchunks = 1;
    return false;
}
int launch_scan_u32(const uint32_t* in, uint32_t* out, const uint32_t* n_dev, uint32_t cap,
                    uint32_t* partials, uint32_t* total_out, hipStream_t s) {
    uint32_t nchunks = div_up(cap, kScanChunk);
    if (nchunks == 0) nchunks = 1;
    if (nchunks > kScanMaxChunks) return invalid("scan longer than 4M elements");
    scan_reduce_kernel<<<nchunks, kScanThreads, 0, s>>>(in, n_dev, cap, partials);
    SICP_LAUNCH_CHECK("scan_reduce_kernel");
    scan_partials_kernel<<<1, 1024, 0, s>>>(partials, nchunks, total_out);
    SICP_LAUNCH_CHECK("scan_partials_kernel");
    scan_apply_kernel<<<nchunks, kScanThreads, 0, s>>>(in, out, n_dev, cap, partials);
    SICP_LAUNCH_CHECK("scan_apply_kernel");
    return SICP_OK;
}

int launch_map_reset(MapDev m, MapCounters* ctr, AllocScratch a, hipStream_t s) {
    reset_table_kernel<<<2048, 256, 0, s>>>(m, a, a.req_mask + 1);
    SICP_LAUNCH_CHECK("reset_table_kernel");
    reset_counters_kernel<<<1, 1, 0, s>>>(ctr);
    SICP_LAUNCH_CHECK("reset_counters_kernel");
    return SICP_OK;
}

int launch_allocate(MapDev m, MapCounters* ctr, AllocScratch a, View<float> depth, PoseArg pose,
                    sicp_intrinsics K, VisFrustum vf, const int* gate, hipStream_t s) {
    const int cols = depth.cols, rows = depth.rows;
    if ((uint32_t)(cols * rows) > a.cap_pixels) return invalid("image larger than the map's max_cols x max_rows");
    uint32_t vchunks;
    const bool vone = vis_chunks(m, vchunks);
    if (vchunks > (uint32_t)kScanMaxChunks) return invalid("heap larger than 4M blocks");
    AllocArgs g{(float)(1.0 / (double)K.fx), (float)(1.0 / (double)K.fy), K.cx, K.cy, cols, rows};
    dim3 grid(div_up(cols, 64), div_up(rows, 4));
    alloc_mark_kernel<<<grid, 256, 0, s>>>(m, ctr, a, depth, pose, g, gate);
    SICP_LAUNCH_CHECK("alloc_mark_kernel");
    // raster chunks of 256 pixels (one per thread) while their sums fit the
    // partial scan, else of 1024
    const uint32_t npx = (uint32_t)(cols * rows);
    const bool one = div_up(npx, kScanThreads) <= (uint32_t)kScanMaxChunks;
    const uint32_t nchunks = (uint32_t)div_up(npx, one ? kScanThreads : kScanChunk);
    if (nchunks > (uint32_t)kScanMaxChunks) return invalid("image above 4M pixels");
    if (one)
        alloc_count_kernel<1><<<nchunks, kScanThreads, 0, s>>>(m, a, depth, pose, g,
                                                               a.scan_partials, gate);
    else
        alloc_count_kernel<kScanPerThread><<<nchunks, kScanThreads, 0, s>>>(
            m, a, depth, pose, g, a.scan_partials, gate);
    SICP_LAUNCH_CHECK("alloc_count_kernel");
    if (one)
        alloc_emit_kernel<1><<<nchunks, kScanThreads, 0, s>>>(m, ctr, a, depth, pose, g,
                                                              a.scan_partials, nchunks, gate);
    else
        alloc_emit_kernel<kScanPerThread><<<nchunks, kScanThreads, 0, s>>>(
            m, ctr, a, depth, pose, g, a.scan_partials, nchunks, gate);
    SICP_LAUNCH_CHECK("alloc_emit_kernel");
    alloc_link_kernel<<<kLinkBlocks + kClearBlocks, 256, 0, s>>>(m, ctr, a, gate);
    SICP_LAUNCH_CHECK("alloc_link_kernel");
    // the insertion with the visibility flags (chunks as launch_visibility
    // forms them)
    const VisArgs vg{K.fx, K.fy, K.cx, K.cy, vf.cols, vf.rows, vf.dmin, vf.dmax};
    const uint32_t nwg = vchunks > 512u ? vchunks : 512u;
    if (vone)
        alloc_insert_flag_kernel<1><<<nwg, 256, 0, s>>>(m, ctr, a, gate, pose, vg, a.scan_partials,
                                                        vchunks);
    else
        alloc_insert_flag_kernel<kScanPerThread><<<nwg, 256, 0, s>>>(
            m, ctr, a, gate, pose, vg, a.scan_partials, vchunks);
    SICP_LAUNCH_CHECK("alloc_insert_flag_kernel");
    return SICP_OK;
}

int launch_visibility(MapDev m, MapCounters* ctr, AllocScratch a, sicp_block* visible,
                      PoseArg pose, sicp_intrinsics K, int cols, int rows, float dmax,
                      float dmin, const int* gate, hipStream_t s) {
    uint32_t nchunks;
    const bool one = vis_chunks(m, nchunks);
    if (nchunks > kScanMaxChunks) return invalid("heap larger than 4M blocks");
    VisArgs g{K.fx, K.fy, K.cx, K.cy, cols, rows, dmin, dmax};
    if (one)
        vis_flag_kernel<1><<<nchunks, kScanThreads, 0, s>>>(m, ctr, a, pose, g, a.scan_partials,
                                                            gate);
    else
        vis_flag_kernel<kScanPerThread><<<nchunks, kScanThreads, 0, s>>>(m, ctr, a, pose, g,
                                                                         a.scan_partials, gate);
    SICP_LAUNCH_CHECK("vis_flag_kernel");
    return launch_visible_compact(m, ctr, a, visible, gate, s);
}

int launch_visible_compact(MapDev m, MapCounters* ctr, AllocScratch a, sicp_block* visible,
                           const int* gate, hipStream_t s) {
    uint32_t nchunks;
    const bool one = vis_chunks(m, nchunks);
    if (nchunks > kScanMaxChunks) return invalid("heap larger than 4M blocks");
    if (one)
        vis_compact_kernel<1><<<nchunks, kScanThreads, 0, s>>>(m, ctr, a, visible,
                                                               a.scan_partials, nchunks, gate);
    else
        vis_compact_kernel<kScanPerThread><<<nchunks, kScanThreads, 0, s>>>(
            m, ctr, a, visible, a.scan_partials, nchunks, gate);
    SICP_LAUNCH_CHECK("vis_compact_kernel");
    return SICP_OK;
}

int launch_fuse(MapDev m, MapCounters* ctr, const sicp_block* visible, View<float> depth,
                View<uint8_t> rgb, View<float4> nmap, View<uint8_t> mask, int num_detection,
                PoseArg pose, sicp_intrinsics K, uint32_t* wave_counts, const int* gate,
                hipStream_t s) {
    FuseArgs g{K.fx, K.fy, K.cx, K.cy, depth.cols, depth.rows, num_detection};
    fuse_kernel<<<kFuseBlocks, 256, 0, s>>>(m, ctr, visible, depth, rgb, nmap, mask, pose, g,
                                            wave_counts, gate);
    SICP_LAUNCH_CHECK("fuse_kernel");
    return SICP_OK;
}

int launch_fuse_count(MapCounters* ctr, const uint32_t* wave_counts, const int* gate,
                      hipStream_t s, View<float> zmin, View<float> zmax, uint32_t* rq_count) {
    const RenderReset rr{zmin, zmax, rq_count};
    const int nz = zmin.ptr ? zmin.cols * zmin.rows : 0;
    const int blocks = 1 + (int)div_up(nz, 1024);
    fuse_end_kernel<<<blocks, 1024, 0, s>>>(ctr, wave_counts, kFuseBlocks * 4, gate, rr);
    SICP_LAUNCH_CHECK("fuse_end_kernel");
    return SICP_OK;
}

}  // namespace sicp
