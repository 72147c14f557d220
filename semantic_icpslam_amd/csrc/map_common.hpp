// map_common.hpp — device-side voxel-block hash shared by the allocation,
// fusion and raycast kernels.
#pragma once
#include "sicp_kernels.hpp"

namespace sicp {

// DeviceMap::Hash (Mapping/DeviceMap.cu:3-10): 32-bit wrapping products, XOR,
// and a remainder taken on the unsigned value (the int XOR converts to the
// unsigned NumBuckets type before `%`).
SICP_DEV uint32_t bucket_of(int x, int y, int z, const MapDev& m) {
    uint32_t h = ((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349669u) ^ ((uint32_t)z * 83492791u);
    return m.nb_mask ? (h & m.nb_mask) : (h % m.nb);
}

SICP_DEV bool key_in_range(int x, int y, int z) {
    const int lim = 1 << 20;
    return x >= -lim && x < lim && y >= -lim && y < lim && z >= -lim && z < lim;
}

SICP_DEV uint64_t pack_key(int x, int y, int z) {
    const uint64_t msk = (1ull << 21) - 1;
    return (((uint64_t)(uint32_t)x & msk) << 42) | (((uint64_t)(uint32_t)y & msk) << 21) |
           ((uint64_t)(uint32_t)z & msk);
}

SICP_DEV i3 unpack_key(uint64_t k) {
    auto sx = [](uint64_t v) { return (int)((int64_t)(v << 43) >> 43); };
    return i3{sx(k >> 42), sx(k >> 21), sx(k)};
}

// voxelPosToBlockPos (DeviceMap.cu:169-180).
SICP_DEV i3 voxel_to_block(i3 v) {
    if (v.x < 0) v.x -= kBlockSide - 1;
    if (v.y < 0) v.y -= kBlockSide - 1;
    if (v.z < 0) v.z -= kBlockSide - 1;
    return i3{v.x / kBlockSide, v.y / kBlockSide, v.z / kBlockSide};
}

// voxelPosToLocalIdx (DeviceMap.cu:186-201,218-220).
SICP_DEV int voxel_local_idx(i3 v) {
    int x = v.x % kBlockSide, y = v.y % kBlockSide, z = v.z % kBlockSide;
    if (x < 0) x += kBlockSide;
    if (y < 0) y += kBlockSide;
    if (z < 0) z += kBlockSide;
    return z * kBlockSide * kBlockSide + y * kBlockSide + x;
}

// ---------------------------------------------------------- voxel storage --
// The FP32 form keeps sicp_voxel_geom records; the half form one 32-bit word
// per voxel (sdf as IEEE half in bits 0-15, weight in bits 16-23) and the
// colour in the semantic record's three spare bytes. Conversions round to
// nearest even (v_cvt_f16_f32) and widen exactly.
SICP_DEV float h2f(uint32_t bits) {
    return (float)__builtin_bit_cast(_Float16, (uint16_t)(bits & 0xffffu));
}
SICP_DEV uint32_t f2h(float v) { return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)v); }
constexpr uint32_t kEmptyHalfVoxel = 0x00007fffu;  // NaN sdf, weight 0

// sdf and weight of voxel k with one load.
SICP_DEV void vox_load(const MapDev& m, size_t k, float& sdf, uint32_t& weight) {
    if (m.half) {
        const uint32_t w = m.geom_h[k];
        sdf = h2f(w);
        weight = (w >> 16) & 0xffu;
    } else {
        const uint2 w = ((const uint2*)m.geom)[k];
        sdf = __uint_as_float(w.x);
        weight = w.y & 0xffu;
    }
}
SICP_DEV float vox_sdf(const MapDev& m, size_t k) {
    return m.half ? h2f(m.geom_h[k]) : m.geom[k].sdf;
}
SICP_DEV uint32_t vox_weight(const MapDev& m, size_t k) {
    return m.half ? (m.geom_h[k] >> 16) & 0xffu : (uint32_t)m.geom[k].weight;
}
// Colour of voxel k.
SICP_DEV void vox_rgb(const MapDev& m, size_t k, uint8_t& r, uint8_t& g, uint8_t& b) {
    if (m.half) {
        const sicp_voxel_sem& s = m.sem[k];
        r = s.pad[0];
        g = s.pad[1];
        b = s.pad[2];
    } else {
        const sicp_voxel_geom& q = m.geom[k];
        r = q.r;
        g = q.g;
        b = q.b;
    }
}

struct EntryView {
    uint64_t key;
    int slot, next;
};
SICP_DEV EntryView load_entry(const sicp_hash_entry* p) {
    uint4 w = *(const uint4*)p;
    return EntryView{(uint64_t)w.x | ((uint64_t)w.y << 32), (int)w.z, (int)w.w};
}

// FindEntry (DeviceMap.cu:141-154) over the two-entry bucket (one 32-byte
// read) and its overflow chain. Returns the heap slot or -1.
SICP_DEV int table_lookup(const sicp_hash_entry* table, const sicp_hash_entry* excess,
                          uint32_t nb, uint32_t nb_mask, int x, int y, int z) {
    if (!key_in_range(x, y, z)) return -1;
    const uint32_t h =
        ((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349669u) ^ ((uint32_t)z * 83492791u);
    const uint32_t b = nb_mask ? (h & nb_mask) : (h % nb);
    const uint64_t key = pack_key(x, y, z);
    // Both entries are read and compared without a branch between them, so
    // the two 16-byte loads are in flight together (one memory round trip).
    EntryView e0 = load_entry(table + 2 * (size_t)b);
    EntryView e1 = load_entry(table + 2 * (size_t)b + 1);
    const int s1 = (e1.slot >= 0 && e1.key == key) ? e1.slot : -1;
    const int s0 = (e0.slot >= 0 && e0.key == key) ? e0.slot : s1;
    if (s0 >= 0) return s0;
    int n = e1.next;
    while (n >= 0) {
        EntryView e = load_entry(excess + n);
        if (e.key == key) return e.slot;
        n = e.next;
    }
    return -1;
}

SICP_DEV int map_lookup(const MapDev& m, int x, int y, int z) {
    return table_lookup(m.table, m.excess, m.nb, m.nb_mask, x, y, z);
}

}  // namespace sicp
