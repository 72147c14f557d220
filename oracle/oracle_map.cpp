// oracle_map.cpp — CPU restatement of the voxel-block hash, allocation,
// visibility, TSDF / colour / label fusion (Mapping/DeviceMap.cu,
// Mapping/FuseMap.cu), the rendering-block prepass and the raycast
// (Mapping/RenderScene.cu). TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Where the reference's outcome depends on thread timing, this restatement
// fixes the outcome that a serial run of the reference kernel would give, and
// the GPU implementation reproduces it:
//  * allocation: every requested block is created (no mutex give-up,
//    DeviceMap.cu:52-66); new blocks are numbered in the order a serial
//    pixel-major run of CreateBlocksKernel (FuseMap.cu:92-132) first meets
//    them, and take heap slots ascending in that order (a fresh heap hands
//    out 0, 1, 2, ..., FuseMap.cu:1275,1286 with DeviceMap.cu:14-16);
//  * table layout: buckets of two entries; a full bucket chains into one
//    overflow entry per heap slot (the slot's own), in creation order;
//  * visible blocks are listed in heap-slot order (the reference's order
//    depends on atomicAdd timing, Utility/ParallelScan.h:32);
//  * one label vote per voxel and frame (the 8x8 launch of
//    FuseMap.cu:1234-1236), applied after the TSDF update of the same voxel.
#include <omp.h>

#include <unordered_set>

#include "oracle_impl.hpp"

namespace oracle {

// IEEE binary16 of a float, rounded to nearest even (overflow to infinity,
// subnormal results rounded at their own precision, NaN kept quiet), and
// back. The precision form stores the sdf this way.
uint16_t f32_to_f16(float f) {
    uint32_t x;
    std::memcpy(&x, &f, sizeof(x));
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t e = (x >> 23) & 0xffu;
    uint32_t m = x & 0x7fffffu;
    if (e == 0xffu) return (uint16_t)(sign | 0x7c00u | (m ? 0x200u | (m >> 13) : 0u));
    const int E = (int)e - 127 + 15;
    if (E >= 31) return (uint16_t)(sign | 0x7c00u);
    if (E <= 0) {
        if (E < -10) return (uint16_t)sign;
        m |= 0x800000u;
        const int shift = 14 - E;
        uint32_t hm = m >> shift;
        const uint32_t rem = m & ((1u << shift) - 1u), halfway = 1u << (shift - 1);
        if (rem > halfway || (rem == halfway && (hm & 1u))) ++hm;
        return (uint16_t)(sign | hm);
    }
    uint32_t h = sign | ((uint32_t)E << 10) | (m >> 13);
    const uint32_t rem = m & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;
    return (uint16_t)h;
}
float f16_to_f32(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16, e = (h >> 10) & 0x1fu, mnt = h & 0x3ffu;
    uint32_t bits;
    if (e == 0x1fu)
        bits = sign | 0x7f800000u | (mnt << 13);
    else if (e != 0)
        bits = sign | ((e + 112u) << 23) | (mnt << 13);
    else if (mnt == 0)
        bits = sign;
    else {
        int sh = 0;
        uint32_t mm = mnt;
        while (!(mm & 0x400u)) {
            mm <<= 1;
            ++sh;
        }
        bits = sign | ((uint32_t)(113 - sh) << 23) | ((mm & 0x3ffu) << 13);
    }
    float f;
    std::memcpy(&f, &bits, sizeof(f));
    return f;
}
template <class R> static inline R store_sdf(const OMap<R>& m, R v) {
    if (std::is_same<R, float>::value && m.cfg.tsdf_half) return (R)f16_to_f32(f32_to_f16((float)v));
    return v;
}

static const int kBS = 8;     // DeviceMap::BlockSize (DeviceMap.h:166)
static const int kBS3 = 512;  // DeviceMap::BlockSize3

// DeviceMap::Hash (DeviceMap.cu:3-10). The products wrap in 32 bits; the XOR
// is an int, but `% NumBuckets` with the unsigned NumBuckets converts it to
// unsigned first, so the remainder is never negative.
uint32_t hash_bucket(int x, int y, int z, uint32_t nb) {
    uint32_t h = ((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349669u) ^ ((uint32_t)z * 83492791u);
    return h % nb;
}

bool key_in_range(int x, int y, int z) {
    const int lim = 1 << 20;
    return x >= -lim && x < lim && y >= -lim && y < lim && z >= -lim && z < lim;
}

uint64_t pack_key(int x, int y, int z) {
    const uint64_t m = (1ull << 21) - 1;
    return (((uint64_t)(uint32_t)x & m) << 42) | (((uint64_t)(uint32_t)y & m) << 21) |
           ((uint64_t)(uint32_t)z & m);
}

// voxelPosToBlockPos (DeviceMap.cu:169-180): negative coordinates step down
// by BlockSize-1 before the truncating division.
static inline I3 voxel_to_block(I3 v) {
    if (v.x < 0) v.x -= kBS - 1;
    if (v.y < 0) v.y -= kBS - 1;
    if (v.z < 0) v.z -= kBS - 1;
    return {v.x / kBS, v.y / kBS, v.z / kBS};
}

// voxelPosToLocalPos + localPosToLocalIdx (DeviceMap.cu:186-201).
static inline int voxel_local_idx(I3 v) {
    int x = v.x % kBS, y = v.y % kBS, z = v.z % kBS;
    if (x < 0) x += kBS;
    if (y < 0) y += kBS;
    if (z < 0) z += kBS;
    return z * kBS * kBS + y * kBS + x;
}

template <class R> void map_init(OMap<R>& m, const sicp_map_config& cfg) {
    m.cfg = cfg;
    m.voxel = (R)cfg.voxel_size;
    m.voxel_inv = (R)(1.0 / (double)cfg.voxel_size);           // DeviceMap.h:186
    m.trunc = (R)cfg.truncation;
    m.step_scale = (R)(0.5 * (double)m.trunc * (double)m.voxel_inv);  // DeviceMap.h:189
    m.block_width = m.voxel * (R)kBS;                          // DeviceMap.h:187
    m.dmin = (R)cfg.depth_min;
    m.dmax = (R)cfg.depth_max;
    m.nb = (uint32_t)cfg.num_buckets;
    m.table.resize((size_t)2 * m.nb);
    m.excess.resize(cfg.num_blocks);
    m.blocks.resize(cfg.num_blocks);
    m.sdf.resize((size_t)cfg.num_blocks * kBS3);
    m.geom.resize((size_t)cfg.num_blocks * kBS3);
    m.sem.resize((size_t)cfg.num_blocks * kBS3);
    map_reset(m);
}

// ResetMap (FuseMap.cu:1251-1304): every entry empty, heap full, voxels
// empty (sdf NaN, weight 0). The entry's chain link is cleared too (the
// reference leaves `offset` unset, DeviceMap.h:36-39).
template <class R> void map_reset(OMap<R>& m) {
    sicp_hash_entry empty{0, -1, -1};
    std::fill(m.table.begin(), m.table.end(), empty);
    std::fill(m.excess.begin(), m.excess.end(), empty);
    std::fill(m.blocks.begin(), m.blocks.end(), sicp_block{0, 0, 0, -1});
    std::fill(m.sdf.begin(), m.sdf.end(), qnan<R>());
    sicp_voxel_geom g{};
    g.sdf = std::numeric_limits<float>::quiet_NaN();
    std::fill(m.geom.begin(), m.geom.end(), g);
    std::fill(m.sem.begin(), m.sem.end(), sicp_voxel_sem{});
    m.heap_used = 0;
    m.visible.clear();
    m.stats = sicp_map_stats{};
}

// FindEntry (DeviceMap.cu:141-154) over the two-entry bucket and its chain.
template <class R> int map_lookup(const OMap<R>& m, int x, int y, int z) {
    if (!key_in_range(x, y, z)) return -1;
    uint32_t b = hash_bucket(x, y, z, m.nb);
    uint64_t key = pack_key(x, y, z);
    const sicp_hash_entry& e0 = m.table[2 * (size_t)b];
    const sicp_hash_entry& e1 = m.table[2 * (size_t)b + 1];
    if (e0.slot >= 0 && e0.key == key) return e0.slot;
    if (e1.slot >= 0 && e1.key == key) return e1.slot;
    int n = e1.next;
    while (n >= 0) {
        const sicp_hash_entry& e = m.excess[n];
        if (e.key == key) return e.slot;
        n = e.next;
    }
    return -1;
}

template <class R> static void map_insert(OMap<R>& m, I3 p, int slot) {
    uint32_t b = hash_bucket(p.x, p.y, p.z, m.nb);
    uint64_t key = pack_key(p.x, p.y, p.z);
    sicp_hash_entry& e0 = m.table[2 * (size_t)b];
    sicp_hash_entry& e1 = m.table[2 * (size_t)b + 1];
    if (e0.slot < 0) {
        e0.key = key;
        e0.slot = slot;
    } else if (e1.slot < 0) {
        e1.key = key;
        e1.slot = slot;
    } else {
        sicp_hash_entry& x = m.excess[slot];
        x.key = key;
        x.slot = slot;
        x.next = -1;
        if (e1.next < 0) {
            e1.next = slot;
        } else {
            int n = e1.next;
            while (m.excess[n].next >= 0) n = m.excess[n].next;
            m.excess[n].next = slot;
        }
    }
    m.blocks[slot] = sicp_block{p.x, p.y, p.z, slot};
    // A new block's voxels start empty (the state ResetSdfBlockKernel leaves).
    for (int i = 0; i < kBS3; ++i) {
        size_t k = (size_t)slot * kBS3 + i;
        m.sdf[k] = qnan<R>();
        m.geom[k] = sicp_voxel_geom{std::numeric_limits<float>::quiet_NaN(), 0, 0, 0, 0};
        m.sem[k] = sicp_voxel_sem{};
    }
}

// Fusion::unproject (FuseMap.cu:35-44).
template <class R>
static inline V3<R> unproject(int x, int y, R z, R invfx, R invfy, Intr<R> K, const Pose<R>& p) {
    V3<R> c;
    c.z = z;
    c.x = z * ((R)x - K.cx) * invfx;
    c.y = z * ((R)y - K.cy) * invfy;
    return add(mv(p.R_, c), p.t);
}

// Fusion::CreateBlocks (FuseMap.cu:92-132), with CreateBlock's insertion
// replaced by the serial pixel-major order described at the top of the file.
template <class R>
static void create_blocks(OMap<R>& m, const R* depth, int cols, int rows, const Pose<R>& pose,
                          Intr<R> K, R invfx, R invfy) {
    const int N = cols * rows;
    int nthreads = omp_get_max_threads();
    // Pass 1 (parallel): each thread records its pixels' requests that miss
    // the table, in pixel order, tagged with the pixel-major request id.
    std::vector<std::vector<std::pair<uint64_t, I3>>> req(nthreads);
    std::vector<uint64_t> probes(nthreads, 0), out_of_range(nthreads, 0);
#pragma omp parallel
    {
        int tid = omp_get_thread_num();
        auto& mine = req[tid];
#pragma omp for schedule(static)
        for (int i = 0; i < N; ++i) {
            int y = i / cols, x = i - y * cols;
            R z = depth[i];
            if (isnan_(z) || z < m.dmin || z > m.dmax) continue;
            R thresh = m.trunc / (R)2;
            R z_near = std::max(m.dmin, (R)(z - thresh));
            R z_far = std::min(m.dmax, (R)(z + thresh));
            if (z_near >= z_far) continue;
            V3<R> pt_near = mul(unproject(x, y, z_near, invfx, invfy, K, pose), m.voxel_inv);
            V3<R> pt_far = mul(unproject(x, y, z_far, invfx, invfy, K, pose), m.voxel_inv);
            V3<R> dir = sub(pt_far, pt_near);
            R length = norm(dir);
            int nsteps = (int)std::ceil(2.0 * (double)length);
            dir = divs(dir, (R)(nsteps - 1));
            I3 prev{0, 0, 0};
            bool have_prev = false;
            for (int s = 0; s < nsteps; ++s) {
                I3 bp = voxel_to_block(floor_i3(pt_near));
                pt_near = add(pt_near, dir);
                if (have_prev && bp.x == prev.x && bp.y == prev.y && bp.z == prev.z) continue;
                prev = bp;
                have_prev = true;
                if (!key_in_range(bp.x, bp.y, bp.z)) {
                    out_of_range[tid]++;
                    continue;
                }
                probes[tid]++;
                if (map_lookup(m, bp.x, bp.y, bp.z) >= 0) continue;
                mine.push_back({(uint64_t)i * 65536ull + (uint64_t)s, bp});
            }
        }
    }
    // Pass 2 (serial): merge in request-id order; first sighting creates.
    std::vector<std::pair<uint64_t, I3>> all;
    for (auto& v : req) all.insert(all.end(), v.begin(), v.end());
    std::stable_sort(all.begin(), all.end(),
                     [](const std::pair<uint64_t, I3>& a, const std::pair<uint64_t, I3>& b) {
                         return a.first < b.first;
                     });
    // Every recorded request missed the table as it stood before this frame,
    // so its first sighting is a new block.
    uint32_t created = 0, dropped = 0;
    std::unordered_set<uint64_t> seen;
    seen.reserve(all.size() * 2 + 16);
    for (auto& r : all) {
        I3 bp = r.second;
        if (!seen.insert(pack_key(bp.x, bp.y, bp.z)).second) continue;
        if (m.heap_used >= (uint32_t)m.cfg.num_blocks) {
            // Heap exhausted: the reference leaves the entry unallocated
            // (DeviceMap.cu:14-20); counted here instead of silently lost.
            dropped++;
            continue;
        }
        map_insert(m, bp, (int)m.heap_used);
        m.heap_used++;
        created++;
    }
    uint64_t pr = 0, oor = 0;
    for (int t = 0; t < nthreads; ++t) {
        pr += probes[t];
        oor += out_of_range[t];
    }
    m.ctr.probes = pr;
    m.stats.new_blocks = created;
    m.stats.dropped_heap += dropped;
    m.stats.dropped_range += (uint32_t)oor;
    m.stats.allocated_blocks = m.heap_used;
}

// Fusion::CheckVertexVisibility (FuseMap.cu:46-56).
template <class R>
static inline bool vertex_visible(V3<R> p3, const Pose<R>& pose, Intr<R> K, int cols, int rows,
                                  R dmin, R dmax) {
    V3<R> pt = mv(pose.Rinv, sub(p3, pose.t));
    if (pt.z < (R)1e-3f) return false;
    R px = K.fx * pt.x / pt.z + K.cx;
    R py = K.fy * pt.y / pt.z + K.cy;
    return px >= 0 && py >= 0 && px < (R)cols && py < (R)rows && pt.z >= dmin && pt.z <= dmax;
}

// Fusion::CheckBlockVisibility (FuseMap.cu:58-90): the eight corners are
// visited by the reference's running additions and subtractions of the block
// width (not by recomputing (pos + o) * width), and any visible corner makes
// the block visible.
template <class R>
static inline bool block_visible(const sicp_block& b, R scale, const Pose<R>& pose, Intr<R> K,
                                 int cols, int rows, R dmin, R dmax) {
    V3<R> c{(R)b.x * scale, (R)b.y * scale, (R)b.z * scale};
    if (vertex_visible(c, pose, K, cols, rows, dmin, dmax)) return true;
    c.z += scale;
    if (vertex_visible(c, pose, K, cols, rows, dmin, dmax)) return true;
    c.y += scale;
    if (vertex_visible(c, pose, K, cols, rows, dmin, dmax)) return true;
    c.x += scale;
    if (vertex_visible(c, pose, K, cols, rows, dmin, dmax)) return true;
    c.z -= scale;
    if (vertex_visible(c, pose, K, cols, rows, dmin, dmax)) return true;
    c.y -= scale;
    if (vertex_visible(c, pose, K, cols, rows, dmin, dmax)) return true;
    c.x -= scale;
    c.y += scale;
    if (vertex_visible(c, pose, K, cols, rows, dmin, dmax)) return true;
    c.x += scale;
    c.y -= scale;
    c.z += scale;
    if (vertex_visible(c, pose, K, cols, rows, dmin, dmax)) return true;
    return false;
}

// CheckFullVisibility (FuseMap.cu:134-168) over the allocated blocks.
template <class R>
uint32_t check_visibility(OMap<R>& m, const Pose<R>& pose, int cols, int rows, Intr<R> K,
                          R depth_max, R depth_min) {
    std::vector<uint8_t> vis(m.heap_used, 0);
#pragma omp parallel for schedule(static)
    for (int s = 0; s < (int)m.heap_used; ++s)
        vis[s] = block_visible(m.blocks[s], m.block_width, pose, K, cols, rows, depth_min,
                               depth_max)
                     ? 1
                     : 0;
    m.visible.clear();
    for (uint32_t s = 0; s < m.heap_used; ++s)
        if (vis[s]) m.visible.push_back(m.blocks[s]);
    m.stats.visible_blocks = (uint32_t)m.visible.size();
    return m.stats.visible_blocks;
}

// Palette of the label colours, Fusion::pallete (FuseMap.cu:18), and the
// colour of label L: (pal * L) % 255 in wrapping int arithmetic (FuseMap.cu:570-572).
static inline int palette_channel(int c, int L) {
    static const int pal[3] = {(1 << 25) - 1, (1 << 15) - 1, (1 << 21) - 1};
    int prod = (int)((uint32_t)pal[c] * (uint32_t)L);
    return prod % 255;
}

// Fusion::createVoxels (FuseMap.cu:432-507) followed, per voxel, by
// Fusion::analyzeSemantics (FuseMap.cu:509-668) when num_detection > 0.
template <class R>
static uint32_t fuse_voxels(OMap<R>& m, const R* depth, const uint8_t* rgb, const R* nmap,
                            const uint8_t* mask, int cols, int rows, int num_detection,
                            const Pose<R>& pose, Intr<R> K) {
    const int V = (int)m.visible.size();
    const R T = m.trunc;
    const int wmax = m.cfg.max_weight;
    const int cmax = m.cfg.max_label_count;
    uint64_t updated = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : updated)
    for (int bi = 0; bi < V; ++bi) {
        const sicp_block& blk = m.visible[bi];
        I3 base{blk.x * kBS, blk.y * kBS, blk.z * kBS};  // blockPosToVoxelPos
        for (int lz = 0; lz < kBS; ++lz)
            for (int ly = 0; ly < kBS; ++ly)
                for (int lx = 0; lx < kBS; ++lx) {
                    int loc = lz * kBS * kBS + ly * kBS + lx;
                    V3<R> pw{(R)(base.x + lx) * m.voxel, (R)(base.y + ly) * m.voxel,
                             (R)(base.z + lz) * m.voxel};
                    V3<R> pc = mv(pose.Rinv, sub(pw, pose.t));
                    int u = f2i(K.fx * pc.x / pc.z + K.cx);
                    int v = f2i(K.fy * pc.y / pc.z + K.cy);
                    if (u < 0 || v < 0 || u >= cols || v >= rows) continue;
                    size_t pix = (size_t)v * cols + u;
                    R dp = depth[pix];
                    if (isnan_(dp) || dp > m.dmax || dp < m.dmin) continue;
                    R sdf = dp - pc.z;
                    if (!(sdf >= -T)) continue;
                    sdf = std::fmin((R)1.0, sdf / T);
                    const uint8_t* c = rgb + pix * 3;
                    size_t k = (size_t)blk.slot * kBS3 + loc;
                    R& vs = m.sdf[k];
                    sicp_voxel_geom& g = m.geom[k];
                    sicp_voxel_sem& s = m.sem[k];
                    ++updated;
                    if (g.weight == 0) {
                        // Voxel(sdf, 1, rgb, 0, 0): labels and backup cleared.
                        vs = sdf;
                        g.weight = 1;
                        g.r = c[0];
                        g.g = c[1];
                        g.b = c[2];
                        s = sicp_voxel_sem{};
                    } else {
                        int w = g.weight;
                        if (s.label == 0) {
                            // (int3(color * w) + val) / (w + 1), truncated (FuseMap.cu:498-499)
                            g.r = u8_of_int(f2i(((R)(g.r * w) + (R)c[0]) / (R)(w + 1)));
                            g.g = u8_of_int(f2i(((R)(g.g * w) + (R)c[1]) / (R)(w + 1)));
                            g.b = u8_of_int(f2i(((R)(g.b * w) + (R)c[2]) / (R)(w + 1)));
                        }
                        vs = (vs * (R)w + sdf) / (R)(w + 1);
                        g.weight = (uint8_t)std::min(wmax, w + 1);
                    }
                    vs = store_sdf(m, vs);  // the precision form keeps a half
                    g.sdf = (float)vs;
                    if (num_detection <= 0) continue;
                    // ---- analyzeSemantics ----
                    const R* nl = nmap + pix * 4;
                    if (isnan_(nl[0])) continue;
                    int L = mask[pix];
                    R val[3];
                    if (L == 0) {
                        val[0] = (R)c[0];
                        val[1] = (R)c[1];
                        val[2] = (R)c[2];
                    } else {
                        for (int ch = 0; ch < 3; ++ch) val[ch] = (R)palette_channel(ch, L);
                    }
                    if (L != 0) {
                        if (L == s.label) {
                            s.count = (uint16_t)std::min(cmax, (int)s.count + 1);
                            g.r = u8_of_int(f2i(val[0]));
                            g.g = u8_of_int(f2i(val[1]));
                            g.b = u8_of_int(f2i(val[2]));
                        } else {
                            if (L == s.label2) {
                                s.count2 = (uint16_t)std::min(cmax, (int)s.count2 + 1);
                                s.r2 = u8_of_int(f2i(val[0]));
                                s.g2 = u8_of_int(f2i(val[1]));
                                s.b2 = u8_of_int(f2i(val[2]));
                            } else if (s.count2 < 7) {
                                s.label2 = (uint8_t)L;
                                s.count2 = 1;
                                s.r2 = u8_of_int(f2i(val[0]));
                                s.g2 = u8_of_int(f2i(val[1]));
                                s.b2 = u8_of_int(f2i(val[2]));
                            }
                            if (((int)s.count < (int)s.count2 && s.label != 0) ||
                                ((int)s.count < (int)s.count2 * 2 && s.label == 0)) {
                                std::swap(s.label, s.label2);
                                std::swap(s.count, s.count2);
                                uint8_t t0 = g.r, t1 = g.g, t2 = g.b;
                                g.r = s.r2;
                                g.g = s.g2;
                                g.b = s.b2;
                                s.r2 = t0;
                                s.g2 = t1;
                                s.b2 = t2;
                            }
                        }
                    } else {
                        // w = nl . normalised((pc, 1)), a four-component dot (FuseMap.cu:641)
                        V4<R> p4{pc.x, pc.y, pc.z, (R)1.0};
                        V4<R> q = normalised4(p4);
                        V4<R> n4{nl[0], nl[1], nl[2], nl[3]};
                        R w = dot4(n4, q);
                        if (s.label != 0) {
                            if (s.count < 7) {
                                s.label2 = s.label;
                                s.r2 = g.r;
                                s.g2 = g.g;
                                s.b2 = g.b;
                                s.count2 = s.count;
                                s.label = 0;
                                s.count = 1;
                                R a = w * (R)0.2f;
                                R res[3];
                                R oldc[3] = {(R)g.r / (R)255.f, (R)g.g / (R)255.f,
                                             (R)g.b / (R)255.f};
                                for (int ch = 0; ch < 3; ++ch) {
                                    R vv = val[ch] / (R)255.f;
                                    res[ch] = (a * vv + ((R)1 - a) * oldc[ch]) * (R)255.f;
                                }
                                g.r = u8_of_int(f2i(res[0]));
                                g.g = u8_of_int(f2i(res[1]));
                                g.b = u8_of_int(f2i(res[2]));
                            }
                        } else {
                            s.count = (uint16_t)std::min(cmax, (int)s.count + 1);
                        }
                    }
                }
    }
    return (uint32_t)updated;
}

// Fusion::UnifyNeighbor + UpdateNeighborLabel (FuseMap.cu:757-948) over the
// visible blocks, serially. Every voxel reads its neighbours' labels as they
// were before the pass (the reference reads them while other threads write
// them). Per column (x, y) of a block, the neighbour tallies carry over from
// one z to the next, as the reference's per-thread arrays do.
template <class R> void unify_neighbor(OMap<R>& m) {
    // the 26 neighbour steps in the order of FuseMap.cu:838-892
    static const int kN[26][3] = {
        {1, 0, 0},   {1, 1, 0},   {1, -1, 0},  {1, 0, 1},   {1, 0, -1},   {1, 1, 1},  {1, 1, -1},
        {1, -1, 1},  {1, -1, -1}, {-1, 0, 0},  {-1, 1, 0},  {-1, -1, 0},  {-1, 0, 1}, {-1, 0, -1},
        {-1, 1, 1},  {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}, {0, 1, 0},   {0, 1, 1},  {0, 1, -1},
        {0, -1, 0},  {0, -1, 1},  {0, -1, -1}, {0, 0, 1},   {0, 0, -1}};
    const std::vector<sicp_voxel_sem> sem0 = m.sem;
    const R vs = (R)m.cfg.voxel_size;
    const int mask = m.cfg.max_label_count <= 255 ? 0xff : 0xffff;
    for (const sicp_block& blk : m.visible) {
        for (int ty = 0; ty < kBS; ++ty)
            for (int tx = 0; tx < kBS; ++tx) {
                int n_label[4] = {-1, -1, -1, -1}, n_count[4] = {0, 0, 0, 0}, n_lc[4] = {0, 0, 0, 0};
                for (int z = 0; z < kBS; ++z) {
                    const size_t own = (size_t)blk.slot * kBS3 + z * 64 + ty * 8 + tx;
                    if (m.geom[own].weight == 0) continue;  // cur_label == -1
                    const int cur_label = sem0[own].label;
                    const int vx = blk.x * kBS + tx, vy = blk.y * kBS + ty, vz = blk.z * kBS + z;
                    // voxelPosToWorldPos (DeviceMap.cu:165-167) and the +-VoxelSize steps
                    const R px = (R)vx * vs, py = (R)vy * vs, pz = (R)vz * vs;
                    const R step = vs * (R)1;
                    for (int k = 0; k < 26; ++k) {
                        V3<R> q{kN[k][0] > 0 ? px + step : (kN[k][0] < 0 ? px - step : px),
                                kN[k][1] > 0 ? py + step : (kN[k][1] < 0 ? py - step : py),
                                kN[k][2] > 0 ? pz + step : (kN[k][2] < 0 ? pz - step : pz)};
                        // FindVoxel(float3, Voxel&): floor(pos / VoxelSize)
                        const I3 v = floor_i3(V3<R>{q.x / vs, q.y / vs, q.z / vs});
                        const I3 bp = voxel_to_block(v);
                        const int slot = map_lookup(m, bp.x, bp.y, bp.z);
                        if (slot < 0) continue;
                        const size_t vi = (size_t)slot * kBS3 + voxel_local_idx(v);
                        if (m.geom[vi].weight == 0) continue;
                        const int tl = sem0[vi].label, tc = sem0[vi].count;
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
                    int nl = -1, nc = 0, nlc = 0;
                    for (int j = 0; j < 4; ++j)
                        if (n_label[j] != -1 && n_lc[j] != 0 && n_count[j] > nc) {
                            nl = n_label[j];
                            nc = n_count[j];
                            nlc = n_lc[j];
                        }
                    sicp_voxel_sem& cur = m.sem[own];
                    if (cur_label != nl) {
                        const int tmp = cur_label == (int)sem0[own].label2
                                            ? (nlc > (int)sem0[own].count2 ? nlc
                                                                           : (int)sem0[own].count2 + 5)
                                            : nlc;
                        cur.label2 = (uint8_t)cur_label;
                        cur.count2 = sem0[own].count;
                        cur.label = (uint8_t)nl;
                        cur.count = (uint16_t)(tmp & mask);
                    } else {
                        cur.count = (uint16_t)(nlc & mask);
                    }
                }
            }
    }
}
template void unify_neighbor<float>(OMap<float>&);
template void unify_neighbor<double>(OMap<double>&);

template <class R>
uint32_t fuse_map_color(OMap<R>& m, const R* depth, const uint8_t* rgb, const R* nmap,
                        const uint8_t* mask, int cols, int rows, int num_detection,
                        const Pose<R>& pose, Intr<R> K) {
    // FuseMapColor passes invfx = 1.0 / fx rounded to float (FuseMap.cu:1072-1073)
    // and uses the map's own depth range (FuseMap.cu:1080-1081).
    R invfx = (R)(1.0 / (double)K.fx), invfy = (R)(1.0 / (double)K.fy);
    create_blocks(m, depth, cols, rows, pose, K, invfx, invfy);
    uint32_t nv = check_visibility(m, pose, cols, rows, K, m.dmax, m.dmin);
    m.stats.updated_voxels = 0;
    if (nv == 0) return 0;  // "no visible blocks found" (FuseMap.cu:1117-1120)
    m.stats.updated_voxels =
        fuse_voxels(m, depth, rgb, nmap, mask, cols, rows, num_detection, pose, K);
    // FuseMapColor runs UnifyNeighbor after the label update (FuseMap.cu:1140-1160)
    if (m.cfg.unify_labels && num_detection > 0 && mask) unify_neighbor(m);
    return nv;
}

// Projection::projectBlock (RenderScene.cu:54-112).
template <class R> struct RBlock {
    int ulx, uly, lrx, lry;
    R zmin, zmax;
};
template <class R>
static bool project_block(const sicp_block& b, const Pose<R>& pose, Intr<R> K, int zcols,
                          int zrows, R depth_max, R depth_min, R voxel, RBlock<R>& rb) {
    rb.ulx = zcols;
    rb.uly = zrows;
    rb.lrx = -1;
    rb.lry = -1;
    rb.zmin = depth_max;
    rb.zmax = depth_min;
    for (int corner = 0; corner < 8; ++corner) {
        int tx = b.x + ((corner & 1) ? 1 : 0);
        int ty = b.y + ((corner & 2) ? 1 : 0);
        int tz = b.z + ((corner & 4) ? 1 : 0);
        // tmp * BlockSize (int) * VoxelSize (float) (RenderScene.cu:69)
        V3<R> p{(R)(tx * kBS) * voxel, (R)(ty * kBS) * voxel, (R)(tz * kBS) * voxel};
        p = mv(pose.Rinv, sub(p, pose.t));
        if ((double)p.z < 2e-1) continue;
        R px = (K.fx * p.x / p.z + K.cx) / (R)8;
        R py = (K.fy * p.y / p.z + K.cy) / (R)8;
        R fx_ = std::floor(px), cx_ = std::ceil(px), fy_ = std::floor(py), cy_ = std::ceil(py);
        // short fields (RenderingBlock, DeviceMap.h:11-18)
        if ((R)rb.ulx > fx_) rb.ulx = (int16_t)f2i(fx_);
        if ((R)rb.lrx < cx_) rb.lrx = (int16_t)f2i(cx_);
        if ((R)rb.uly > fy_) rb.uly = (int16_t)f2i(fy_);
        if ((R)rb.lry < cy_) rb.lry = (int16_t)f2i(cy_);
        if (rb.zmin > p.z) rb.zmin = p.z;
        if (rb.zmax < p.z) rb.zmax = p.z;
    }
    if (rb.ulx < 0) rb.ulx = 0;
    if (rb.uly < 0) rb.uly = 0;
    if (rb.lrx >= zcols) rb.lrx = zcols - 1;
    if (rb.lry >= zrows) rb.lry = zrows - 1;
    if (rb.ulx > rb.lrx) return false;
    if (rb.uly > rb.lry) return false;
    if (rb.zmin < depth_min) rb.zmin = depth_min;
    if (rb.zmax < depth_min) return false;
    return true;
}

// CreateRenderingBlocks (RenderScene.cu:210-289) with the tile split of
// createRenderingBlockList (:114-132) and fillBlocks (:167-190). The tile
// loop runs x < ny and does not offset the tile's lower-right corner, so the
// covered 1/8-resolution pixels of a block box [ul, lr] are the columns
// ul.x .. min(ul.x+15, lr.x), plus column ul.x+16 when ny >= 2 and
// lr.x >= ul.x+16, times the rows ul.y .. min(ul.y+15, lr.y), plus row
// ul.y+16 on the same condition. zmin starts at 100, zmax at 0.
template <class R>
int create_rendering_blocks(OMap<R>& m, R* zmin, R* zmax, int zcols, int zrows, R depth_max,
                            R depth_min, const Pose<R>& pose, Intr<R> K) {
    for (int i = 0; i < zcols * zrows; ++i) {
        zmin[i] = (R)100.f;
        zmax[i] = (R)0;
    }
    if (m.visible.empty()) return 0;
    int touched = 0;
    for (size_t i = 0; i < m.visible.size(); ++i) {
        RBlock<R> rb;
        if (!project_block(m.visible[i], pose, K, zcols, zrows, depth_max, depth_min, m.voxel,
                           rb))
            continue;
        R dx = (R)rb.lrx - (R)rb.ulx + (R)1;
        R dy = (R)rb.lry - (R)rb.uly + (R)1;
        int nx = (int)std::ceil(dx / (R)16);
        int ny = (int)std::ceil(dy / (R)16);
        (void)nx;
        if (ny <= 0) continue;
        int xs[17], ys[17], ncx = 0, ncy = 0;
        for (int x = rb.ulx; x <= std::min(rb.ulx + 15, rb.lrx); ++x) xs[ncx++] = x;
        if (ny >= 2 && rb.lrx >= rb.ulx + 16) xs[ncx++] = rb.ulx + 16;
        for (int y = rb.uly; y <= std::min(rb.uly + 15, rb.lry); ++y) ys[ncy++] = y;
        if (ny >= 2 && rb.lry >= rb.uly + 16) ys[ncy++] = rb.uly + 16;
        for (int a = 0; a < ncy; ++a)
            for (int c = 0; c < ncx; ++c) {
                int X = xs[c], Y = ys[a];
                if (X >= zcols || Y >= zrows) continue;
                R& lo = zmin[(size_t)Y * zcols + X];
                R& hi = zmax[(size_t)Y * zcols + X];
                lo = std::fmin(rb.zmin, lo);
                hi = std::fmax(rb.zmax, hi);
                touched = 1;
            }
    }
    uint32_t tiles = 0;
    for (int i = 0; i < zcols * zrows; ++i) tiles += (zmax[i] > (R)0) ? 1 : 0;
    m.stats.render_tiles = tiles;
    return touched;
}

// DeviceMap::FindVoxel(float3, cache, valid) (DeviceMap.cu:116-133). The
// cache starts empty: the reference's default HashEntry (pos 0, ptr -1)
// would match block (0,0,0) and read voxelBlocks[-1 + idx].
template <class R> struct Cache {
    I3 pos;
    int slot;  // -1: empty
};
template <class R>
static inline int find_voxel(const OMap<R>& m, V3<R> pt, Cache<R>& cache, bool& valid,
                             uint64_t& probes) {
    I3 p = floor_i3(pt);
    I3 bp = voxel_to_block(p);
    int idx = voxel_local_idx(p);
    if (cache.slot >= 0 && bp.x == cache.pos.x && bp.y == cache.pos.y && bp.z == cache.pos.z) {
        valid = true;
        return cache.slot * kBS3 + idx;
    }
    ++probes;
    int slot = map_lookup(m, bp.x, bp.y, bp.z);
    if (slot < 0) {
        valid = false;
        return -1;
    }
    valid = true;
    cache.pos = bp;
    cache.slot = slot;
    return slot * kBS3 + idx;
}

template <class R> static inline R sdf_at(const OMap<R>& m, int k) {
    return k < 0 ? qnan<R>() : m.sdf[k];
}

// Rendering::readSdfInterped (RenderScene.cu:310-343): `valid` is that of
// the last of the eight lookups.
template <class R>
static inline R read_sdf_interped(const OMap<R>& m, V3<R> pt, Cache<R>& cache, bool& valid,
                                  uint64_t& probes) {
    V3<R> fl{std::floor(pt.x), std::floor(pt.y), std::floor(pt.z)};
    V3<R> xyz = sub(pt, fl);
    R s0, s1, r0, r1, r2, r3;
    s0 = sdf_at(m, find_voxel(m, pt, cache, valid, probes));
    s1 = sdf_at(m, find_voxel(m, add(pt, V3<R>{1, 0, 0}), cache, valid, probes));
    r0 = ((R)1.0f - xyz.x) * s0 + xyz.x * s1;
    s0 = sdf_at(m, find_voxel(m, add(pt, V3<R>{0, 1, 0}), cache, valid, probes));
    s1 = sdf_at(m, find_voxel(m, add(pt, V3<R>{1, 1, 0}), cache, valid, probes));
    r1 = ((R)1.0f - xyz.x) * s0 + xyz.x * s1;
    r2 = ((R)1.0f - xyz.y) * r0 + xyz.y * r1;
    s0 = sdf_at(m, find_voxel(m, add(pt, V3<R>{0, 0, 1}), cache, valid, probes));
    s1 = sdf_at(m, find_voxel(m, add(pt, V3<R>{1, 0, 1}), cache, valid, probes));
    r0 = ((R)1.0f - xyz.x) * s0 + xyz.x * s1;
    s0 = sdf_at(m, find_voxel(m, add(pt, V3<R>{0, 1, 1}), cache, valid, probes));
    s1 = sdf_at(m, find_voxel(m, add(pt, V3<R>{1, 1, 1}), cache, valid, probes));
    r1 = ((R)1.0f - xyz.x) * s0 + xyz.x * s1;
    r3 = ((R)1.0f - xyz.y) * r0 + xyz.y * r1;
    return ((R)1.0f - xyz.z) * r2 + xyz.z * r3;
}

// Rendering::readNormal (RenderScene.cu:345-378).
template <class R>
static inline bool read_normal(const OMap<R>& m, V3<R> pt, Cache<R>& cache, const Pose<R>& pose,
                               V3<R>& n, uint64_t& probes, int& reads) {
    bool valid;
    static const int off[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    R s[6];
    for (int i = 0; i < 6; ++i) {
        s[i] = read_sdf_interped(m, add(pt, V3<R>{(R)off[i][0], (R)off[i][1], (R)off[i][2]}),
                                 cache, valid, probes);
        ++reads;
        if (isnan_(s[i]) || s[i] == (R)1.0f || !valid) return false;
    }
    n = V3<R>{s[0] - s[1], s[2] - s[3], s[4] - s[5]};
    n = normalised(mv(pose.Rinv, n));
    return true;
}

// Rendering::operator() (RenderScene.cu:380-488).
template <class R>
void raycast(OMap<R>& m, R* vmap, R* nmap, int cols, int rows, const R* zmin, const R* zmax,
             int zcols, const Pose<R>& pose, R invfx, R invfy, R cx, R cy) {
    uint64_t probes = 0, steps = 0, c_norange = 0, c_back = 0, c_nocross = 0, c_nonormal = 0,
             c_hit = 0, c_interp = 0;
#pragma omp parallel for schedule(dynamic, 4) \
    reduction(+ : probes, steps, c_norange, c_back, c_nocross, c_nonormal, c_hit, c_interp)
    for (int y = 0; y < rows; ++y) {
        for (int x = 0; x < cols; ++x) {
            R* vo = vmap + ((size_t)y * cols + x) * 4;
            R* no = nmap + ((size_t)y * cols + x) * 4;
            for (int c = 0; c < 4; ++c) vo[c] = no[c] = qnan<R>();
            int lx = x / 8, ly = y / 8;
            R z0 = zmin[(size_t)ly * zcols + lx];
            R z1 = zmax[(size_t)ly * zcols + lx];
            if ((double)z1 < 1e-3 || (double)z0 < 1e-3 || isnan_(z0) || isnan_(z1)) {
                ++c_norange;
                continue;
            }
            R sdf = (R)1.0f;
            V3<R> p;
            p.z = z0;
            p.x = p.z * ((R)x - cx) * invfx;
            p.y = p.z * ((R)y - cy) * invfy;
            R dist_s = norm(p) * m.voxel_inv;
            V3<R> block_s = mul(add(mv(pose.R_, p), pose.t), m.voxel_inv);
            p.z = z1;
            p.x = p.z * ((R)x - cx) * invfx;
            p.y = p.z * ((R)y - cy) * invfy;
            R dist_e = norm(p) * m.voxel_inv;
            V3<R> block_e = mul(add(mv(pose.R_, p), pose.t), m.voxel_inv);
            V3<R> dir = normalised(sub(block_e, block_s));
            V3<R> result = block_s;
            bool valid_sdf = false, found_pt = false;
            R step;
            Cache<R> cache{{0, 0, 0}, -1};
            R prev_sdf;
            bool back_face = false;
            while (dist_s < dist_e) {
                ++steps;
                prev_sdf = sdf;
                // readSdf (RenderScene.cu:303-308): invalid when the block is
                // missing or the voxel has weight 0.
                int k = find_voxel(m, result, cache, valid_sdf, probes);
                if (k < 0) {
                    sdf = qnan<R>();
                } else {
                    sdf = m.sdf[k];
                    if (m.geom[k].weight == 0) valid_sdf = false;
                }
                if (!valid_sdf) {
                    step = (R)kBS;
                } else {
                    if (sdf <= (R)0.1f && sdf >= (R)-0.5f) {
                        ++c_interp;
                        sdf = read_sdf_interped(m, result, cache, valid_sdf, probes);
                    }
                    if (prev_sdf < (R)0.f && sdf >= (R)0.f) {
                        back_face = true;
                        break;
                    }
                    if (prev_sdf > (R)0.f && sdf <= (R)0.f) break;
                    step = (R)1.0;
                }
                result = add(result, mul(step, dir));
                dist_s += step;
            }
            if (back_face) {
                ++c_back;
                continue;
            }
            if (!(sdf <= (R)0.0f)) ++c_nocross;
            if (sdf <= (R)0.0f) {
                step = sdf * m.step_scale;
                result = add(result, mul(step, dir));
                sdf = read_sdf_interped(m, result, cache, valid_sdf, probes);
                ++c_interp;
                step = sdf * m.step_scale;
                result = add(result, mul(step, dir));
                found_pt = true;
            }
            if (found_pt) {
                V3<R> normal;
                int reads = 0;
                bool has_normal = read_normal(m, result, cache, pose, normal, probes, reads);
                c_interp += reads;
                if (has_normal) ++c_hit; else ++c_nonormal;
                if (has_normal) {
                    V3<R> r = mv(pose.Rinv, sub(mul(result, m.voxel), pose.t));
                    vo[0] = r.x;
                    vo[1] = r.y;
                    vo[2] = r.z;
                    vo[3] = (R)1.0;
                    no[0] = normal.x;
                    no[1] = normal.y;
                    no[2] = normal.z;
                    no[3] = (R)1.0;
                }
            }
        }
    }
    m.ctr.interp = probes;
    m.ctr.steps = steps;
    m.ctr.ray_norange = c_norange;
    m.ctr.ray_back = c_back;
    m.ctr.ray_nocross = c_nocross;
    m.ctr.ray_nonormal = c_nonormal;
    m.ctr.ray_hit = c_hit;
    m.ctr.interp_reads = c_interp;
}

template void map_init<float>(OMap<float>&, const sicp_map_config&);
template void map_init<double>(OMap<double>&, const sicp_map_config&);
template void map_reset<float>(OMap<float>&);
template void map_reset<double>(OMap<double>&);
template int map_lookup<float>(const OMap<float>&, int, int, int);
template int map_lookup<double>(const OMap<double>&, int, int, int);
template uint32_t fuse_map_color<float>(OMap<float>&, const float*, const uint8_t*, const float*,
                                        const uint8_t*, int, int, int, const Pose<float>&,
                                        Intr<float>);
template uint32_t fuse_map_color<double>(OMap<double>&, const double*, const uint8_t*,
                                         const double*, const uint8_t*, int, int, int,
                                         const Pose<double>&, Intr<double>);
template uint32_t check_visibility<float>(OMap<float>&, const Pose<float>&, int, int, Intr<float>,
                                          float, float);
template uint32_t check_visibility<double>(OMap<double>&, const Pose<double>&, int, int,
                                           Intr<double>, double, double);
template int create_rendering_blocks<float>(OMap<float>&, float*, float*, int, int, float, float,
                                            const Pose<float>&, Intr<float>);
template int create_rendering_blocks<double>(OMap<double>&, double*, double*, int, int, double,
                                             double, const Pose<double>&, Intr<double>);
template void raycast<float>(OMap<float>&, float*, float*, int, int, const float*, const float*,
                             int, const Pose<float>&, float, float, float, float);
template void raycast<double>(OMap<double>&, double*, double*, int, int, const double*,
                              const double*, int, const Pose<double>&, double, double, double,
                              double);

}  // namespace oracle

using namespace oracle;

struct OracleMapHandle {
    int fp64;
    OMap<float> f;
    OMap<double> d;
};

// ------------------------------------------------------------ meshing -----
// MeshEngine (Mapping/MeshScene.cu:4-305) run serially: every allocated
// block in heap-slot order, its voxels in index order z*64 + y*8 + x, each
// voxel's triangles in table order, written until max_tris. The tables are
// inputs, as they are arguments of the reference's MeshScene.
namespace {
struct MeshVox {
    float sdf;
    uint8_t c[3];
};
// FindVoxel(int3) (DeviceMap.cu:97-103): the default voxel (NaN sdf, black)
// when the block is missing.
MeshVox mesh_voxel(const OMap<float>& m, int x, int y, int z) {
    I3 v{x, y, z};
    I3 bp = voxel_to_block(v);
    const int slot = map_lookup(m, bp.x, bp.y, bp.z);
    if (slot < 0) return MeshVox{std::nanf(""), {0, 0, 0}};
    const size_t i = (size_t)slot * kBS3 + voxel_local_idx(v);
    return MeshVox{m.sdf[i], {m.geom[i].r, m.geom[i].g, m.geom[i].b}};
}
// readVertexAndColor's corner order (MeshScene.cu:120-152).
const int kMC[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                       {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
// buildVertexList's edges as (from, to) corners (MeshScene.cu:200-291).
const int kME[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                        {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
int mesh_corner(int ox, int oy, int oz) {
    for (int k = 0; k < 8; ++k)
        if (kMC[k][0] == ox && kMC[k][1] == oy && kMC[k][2] == oz) return k;
    return -1;
}
// MeshEngine::interp (MeshScene.cu:156-164).
float mesh_interp(float v1, float v2) {
    if (std::fabs(0 - v1) < 1e-6) return 0;
    if (std::fabs(0 - v2) < 1e-6) return 1;
    if (std::fabs(v1 - v2) < 1e-6) return 0;
    return (0 - v1) / (v2 - v1);
}
}  // namespace

uint32_t mesh_scene(const OMap<float>& m, const int32_t* edge, const int8_t* tri,
                    const int32_t* nvert, uint32_t max_tris, float* V, float* N, uint8_t* Cout,
                    uint32_t* total) {
    uint32_t nt = 0;
    for (uint32_t b = 0; b < m.heap_used; ++b) {
        const sicp_block& blk = m.blocks[b];
        for (int li = 0; li < kBS3; ++li) {
            const int lx = li & 7, ly = (li >> 3) & 7, lz = li >> 6;
            const int px = blk.x * kBS + lx, py = blk.y * kBS + ly, pz = blk.z * kBS + lz;
            // readVertexAndColor (:117-153)
            float sdf[8];
            uint8_t col[8][3];
            bool ok = true;
            for (int k = 0; k < 8 && ok; ++k) {
                MeshVox v = mesh_voxel(m, px + kMC[k][0], py + kMC[k][1], pz + kMC[k][2]);
                sdf[k] = v.sdf;
                std::memcpy(col[k], v.c, 3);
                ok = !(sdf[k] == 1.0 || std::isnan(sdf[k]));
            }
            if (!ok) continue;
            // readNormal (:39-115): central differences at every corner
            V3<float> nrm[8];
            for (int k = 0; k < 8 && ok; ++k) {
                float g[3];
                for (int a = 0; a < 3 && ok; ++a) {
                    int o[3] = {kMC[k][0], kMC[k][1], kMC[k][2]};
                    int q[3] = {o[0], o[1], o[2]};
                    q[a] = o[a] ? 2 : -1;
                    const float far = mesh_voxel(m, px + q[0], py + q[1], pz + q[2]).sdf;
                    if (std::isnan(far)) ok = false;
                    int r[3] = {o[0], o[1], o[2]};
                    r[a] = 1 - o[a];
                    const float near = sdf[mesh_corner(r[0], r[1], r[2])];
                    g[a] = o[a] ? far - near : near - far;
                }
                nrm[k] = V3<float>{g[0], g[1], g[2]};
            }
            if (!ok) continue;
            int cube = 0;
            for (int k = 0; k < 8; ++k)
                if (sdf[k] < 0) cube |= 1 << k;
            if (edge[cube] == 0 || cube == 0) continue;
            const int ntri = nvert[cube] / 3;
            for (int j = 0; j < ntri; ++j, ++nt) {
                if (nt >= max_tris) continue;
                for (int v = 0; v < 3; ++v) {
                    const int e = tri[cube * 16 + 3 * j + v];
                    const int a = kME[e][0], bb = kME[e][1];
                    const float val = mesh_interp(sdf[a], sdf[bb]);
                    const int P[3] = {px, py, pz};
                    const size_t o = ((size_t)nt * 3 + v) * 3;
                    for (int ax = 0; ax < 3; ++ax) {
                        const int ca = kMC[a][ax], cb = kMC[bb][ax];
                        const float comp = ca == cb ? (float)ca : (ca == 0 ? val : 1 - val);
                        V[o + ax] = ((float)P[ax] + comp) * m.cfg.voxel_size;
                    }
                    const V3<float> na = nrm[a], nb = nrm[bb];
                    V3<float> n{na.x + val * (nb.x - na.x), na.y + val * (nb.y - na.y),
                                na.z + val * (nb.z - na.z)};
                    const float len = std::sqrt(n.x * n.x + n.y * n.y + n.z * n.z);
                    N[o + 0] = n.x / len;
                    N[o + 1] = n.y / len;
                    N[o + 2] = n.z / len;
                    // uchar3 + (int)val * (uchar3 - uchar3), narrowing mod 256
                    const int iv = (int)val;
                    for (int ch = 0; ch < 3; ++ch) {
                        const uint8_t d = (uint8_t)(col[bb][ch] - col[a][ch]);
                        Cout[o + ch] = (uint8_t)(col[a][ch] + (uint8_t)(d * iv));
                    }
                }
            }
        }
    }
    if (total) *total = nt;
    return nt < max_tris ? nt : max_tris;
}

extern "C" {

void* oracle_map_create(const sicp_map_config* cfg, int fp64) {
    OracleMapHandle* h = new OracleMapHandle();
    h->fp64 = fp64;
    if (fp64)
        map_init(h->d, *cfg);
    else
        map_init(h->f, *cfg);
    return h;
}

void oracle_map_destroy(void* map) { delete (OracleMapHandle*)map; }

void oracle_map_reset(void* map) {
    OracleMapHandle* h = (OracleMapHandle*)map;
    if (h->fp64)
        map_reset(h->d);
    else
        map_reset(h->f);
}

uint32_t oracle_map_hash(const void* map, int x, int y, int z) {
    const OracleMapHandle* h = (const OracleMapHandle*)map;
    return hash_bucket(x, y, z, h->fp64 ? h->d.nb : h->f.nb);
}

uint32_t oracle_fuse_map_color(void* map, const float* depth, const uint8_t* rgb,
                               const float* nmap, const uint8_t* mask, int cols, int rows,
                               int num_detection, const sicp_pose_f* pose, sicp_intrinsics K) {
    OracleMapHandle* h = (OracleMapHandle*)map;
    Pose<float> p = pose_from_f<float>(pose->R, pose->Rinv, pose->t);
    return fuse_map_color(h->f, depth, rgb, nmap, mask, cols, rows, num_detection, p,
                          Intr<float>{K.fx, K.fy, K.cx, K.cy});
}

uint32_t oracle_fuse_map_color_fp64(void* map, const double* depth, const uint8_t* rgb,
                                    const double* nmap, const uint8_t* mask, int cols, int rows,
                                    int num_detection, const double* pose, const double* K4) {
    OracleMapHandle* h = (OracleMapHandle*)map;
    Pose<double> p = pose_from_d<double>(pose);
    return fuse_map_color(h->d, depth, rgb, nmap, mask, cols, rows, num_detection, p,
                          Intr<double>{K4[0], K4[1], K4[2], K4[3]});
}

uint32_t oracle_check_visibility(void* map, const sicp_pose_f* pose, int cols, int rows,
                                 sicp_intrinsics K, float depth_max, float depth_min) {
    OracleMapHandle* h = (OracleMapHandle*)map;
    Pose<float> p = pose_from_f<float>(pose->R, pose->Rinv, pose->t);
    return check_visibility(h->f, p, cols, rows, Intr<float>{K.fx, K.fy, K.cx, K.cy}, depth_max,
                            depth_min);
}

int oracle_create_rendering_blocks(void* map, float* zmin, float* zmax, int zcols, int zrows,
                                   float depth_max, float depth_min, const sicp_pose_f* pose,
                                   sicp_intrinsics K) {
    OracleMapHandle* h = (OracleMapHandle*)map;
    Pose<float> p = pose_from_f<float>(pose->R, pose->Rinv, pose->t);
    return create_rendering_blocks(h->f, zmin, zmax, zcols, zrows, depth_max, depth_min, p,
                                   Intr<float>{K.fx, K.fy, K.cx, K.cy});
}

void oracle_raycast(void* map, float* vmap, float* nmap, int cols, int rows, const float* zmin,
                    const float* zmax, int zcols, const sicp_pose_f* pose, sicp_intrinsics K) {
    OracleMapHandle* h = (OracleMapHandle*)map;
    Pose<float> p = pose_from_f<float>(pose->R, pose->Rinv, pose->t);
    // Mapping::RayTrace passes 1.0 / fx rounded to float (Mapping.cc:158-159).
    raycast(h->f, vmap, nmap, cols, rows, zmin, zmax, zcols, p, (float)(1.0 / (double)K.fx),
            (float)(1.0 / (double)K.fy), K.cx, K.cy);
}

// A new FP64 map holding exactly the state of an FP32 one (table, blocks,
// sdf widened to double, weights, colours, labels, visible list).
void* oracle_map_clone_fp64(void* src) {
    OracleMapHandle* s = (OracleMapHandle*)src;
    OracleMapHandle* h = new OracleMapHandle();
    h->fp64 = 1;
    map_init(h->d, s->f.cfg);
    h->d.table = s->f.table;
    h->d.excess = s->f.excess;
    h->d.blocks = s->f.blocks;
    for (size_t i = 0; i < s->f.sdf.size(); ++i) h->d.sdf[i] = (double)s->f.sdf[i];
    h->d.geom = s->f.geom;
    h->d.sem = s->f.sem;
    h->d.heap_used = s->f.heap_used;
    h->d.visible = s->f.visible;
    h->d.stats = s->f.stats;
    return h;
}

// The raycast of an FP64 map at a double pose (4x4 camera -> world), the
// zRange images given as floats; intrinsics as Raycast receives them.
void oracle_raycast_fp64(void* map, double* vmap, double* nmap, int cols, int rows,
                         const float* zmin, const float* zmax, int zcols, const double* pose16,
                         double invfx, double invfy, double cx, double cy) {
    OracleMapHandle* h = (OracleMapHandle*)map;
    const int zrows = (rows + 7) / 8;
    std::vector<double> z0((size_t)zcols * zrows), z1((size_t)zcols * zrows);
    for (size_t i = 0; i < z0.size(); ++i) {
        z0[i] = zmin[i];
        z1[i] = zmax[i];
    }
    Pose<double> p = pose_from_d<double>(pose16);
    raycast(h->d, vmap, nmap, cols, rows, z0.data(), z1.data(), zcols, p, invfx, invfy, cx, cy);
}

// Test hooks: one voxel's sdf (FP32 or FP64 map).
double oracle_map_get_sdf(void* map, int slot, int idx) {
    OracleMapHandle* h = (OracleMapHandle*)map;
    const size_t i = (size_t)slot * kBS3 + idx;
    return h->fp64 ? h->d.sdf[i] : (double)h->f.sdf[i];
}
void oracle_map_set_sdf(void* map, int slot, int idx, double v) {
    OracleMapHandle* h = (OracleMapHandle*)map;
    const size_t i = (size_t)slot * kBS3 + idx;
    if (h->fp64)
        h->d.sdf[i] = v;
    else
        h->f.sdf[i] = (float)v;
}

void oracle_map_get_stats(void* map, sicp_map_stats* out) {
    OracleMapHandle* h = (OracleMapHandle*)map;
    *out = h->fp64 ? h->d.stats : h->f.stats;
}

void oracle_map_export(void* map, sicp_hash_entry* table, sicp_hash_entry* excess,
                       sicp_block* blocks, sicp_voxel_geom* geom, sicp_voxel_sem* sem) {
    OracleMapHandle* h = (OracleMapHandle*)map;
    auto go = [&](auto& m) {
        if (table) std::memcpy(table, m.table.data(), m.table.size() * sizeof(sicp_hash_entry));
        if (excess)
            std::memcpy(excess, m.excess.data(), m.excess.size() * sizeof(sicp_hash_entry));
        if (blocks) std::memcpy(blocks, m.blocks.data(), m.blocks.size() * sizeof(sicp_block));
        if (geom) {
            std::memcpy(geom, m.geom.data(), m.geom.size() * sizeof(sicp_voxel_geom));
            for (size_t i = 0; i < m.sdf.size(); ++i) geom[i].sdf = (float)m.sdf[i];
        }
        if (sem) std::memcpy(sem, m.sem.data(), m.sem.size() * sizeof(sicp_voxel_sem));
    };
    if (h->fp64)
        go(h->d);
    else
        go(h->f);
}

uint16_t oracle_f32_to_f16(float f) { return f32_to_f16(f); }
float oracle_f16_to_f32(uint16_t h) { return f16_to_f32(h); }

// Test hook: sets one voxel's weight and label state (FP32 map).
void oracle_map_set_label(void* map, int slot, int idx, int weight, int label, int count,
                          int label2, int count2) {
    OracleMapHandle* h = (OracleMapHandle*)map;
    const size_t i = (size_t)slot * kBS3 + idx;
    h->f.geom[i].weight = (uint8_t)weight;
    h->f.sem[i].label = (uint8_t)label;
    h->f.sem[i].count = (uint16_t)count;
    h->f.sem[i].label2 = (uint8_t)label2;
    h->f.sem[i].count2 = (uint16_t)count2;
}

void oracle_unify_neighbor(void* map) {
    OracleMapHandle* h = (OracleMapHandle*)map;
    if (h->fp64)
        unify_neighbor(h->d);
    else
        unify_neighbor(h->f);
}

uint32_t oracle_mesh_scene(void* map, const int32_t* edge, const int8_t* tri,
                           const int32_t* nvert, uint32_t max_tris, float* vertices,
                           float* normals, uint8_t* colors, uint32_t* total) {
    OracleMapHandle* h = (OracleMapHandle*)map;
    return mesh_scene(h->f, edge, tri, nvert, max_tris, vertices, normals, colors, total);
}

uint32_t oracle_map_visible(void* map, sicp_block* out, uint32_t cap) {
    OracleMapHandle* h = (OracleMapHandle*)map;
    auto& v = h->fp64 ? h->d.visible : h->f.visible;
    uint32_t n = (uint32_t)std::min<size_t>(cap, v.size());
    if (out && n) std::memcpy(out, v.data(), n * sizeof(sicp_block));
    return (uint32_t)v.size();
}

void oracle_map_counters(void* map, uint64_t* out) {
    OracleMapHandle* h = (OracleMapHandle*)map;
    MapCounters c = h->fp64 ? h->d.ctr : h->f.ctr;
    uint64_t v[8] = {c.probes, c.interp, c.steps, c.ray_norange,
                     c.ray_back, c.ray_nocross, c.ray_nonormal, c.ray_hit};
    std::memcpy(out, v, sizeof(v));
}

}  // extern "C"
