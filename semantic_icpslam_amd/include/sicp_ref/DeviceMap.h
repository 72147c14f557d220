// DeviceMap.h — the map handle and pose types the reference's mapping and
// tracking entry points take (Mapping/DeviceMap.h:191-201,
// Utility/VectorMath.h:360-380).
//
// The reference's DeviceMap is a bundle of raw device views over buffers that
// Mapping owns. Here the map is one owned object (`sicp_map`, created from a
// `sicp_map_config`) and DeviceMap is a shared handle to it;
// `visibleEntries` views the map's visible list the way the reference's
// `map.visibleEntries` does.
#pragma once

#include <hip/hip_vector_types.h>

#include <map>
#include <memory>
#include <mutex>

#include "DeviceArray.h"
#include "sicp.h"

// Row-major 3x3 float matrix (VectorMath.h:360-362).
struct Matrix3f {
    float3 rowx, rowy, rowz;
    Matrix3f() : rowx(make_float3(0, 0, 0)), rowy(make_float3(0, 0, 0)), rowz(make_float3(0, 0, 0)) {}
};

// One visible-list record (block coordinate and heap slot) stands in for the
// reference's HashEntry in visible-list arrays.
typedef sicp_block HashEntry;
// The reference's per-tile records of CreateRenderingBlocks; the tiles are
// folded straight into the zRange images here, so lists of this type are
// accepted and left untouched.
struct RenderingBlock {
    short2 upperLeft, lowerRight;
    float2 zRange;
};

// Which map a visible-list view belongs to (CreateRenderingBlocks receives
// only the list, as in the reference).
inline std::map<const void*, sicp_map*>& sicp_visible_owner() {
    static std::map<const void*, sicp_map*> m;
    return m;
}
inline std::mutex& sicp_visible_owner_lock() {
    static std::mutex m;
    return m;
}

struct DeviceMap {
    // The reference's compile-time map constants (Mapping/DeviceMap.h:166-189),
    // kept so its callers compile; here they only seed sicp_map_config.
    static constexpr unsigned BlockSize = 8;
    static constexpr unsigned BlockSize3 = 512;
    static constexpr float DepthMin = 0.1f;
    static constexpr float DepthMax = 3.0f;
    static constexpr unsigned NumBuckets = 1000000;
    static constexpr unsigned NumSdfBlocks = 50000;
    static constexpr float VoxelSize = 0.004f;
    static constexpr float TruncateDist = VoxelSize * 6;

    DeviceMap() = default;
    explicit DeviceMap(const sicp_map_config& cfg) { create(cfg); }

    void create(const sicp_map_config& cfg) {
        sicp_map* m = nullptr;
        SicpCall(sicp_map_create(&cfg, &m), "sicp_map_create");
        sicp_block* list = nullptr;
        uint32_t cap = 0;
        SicpCall(sicp_map_visible_list(m, &list, &cap), "sicp_map_visible_list");
        {
            std::lock_guard<std::mutex> g(sicp_visible_owner_lock());
            sicp_visible_owner()[list] = m;
        }
        handle = std::shared_ptr<sicp_map>(m, [list](sicp_map* q) {
            {
                std::lock_guard<std::mutex> g(sicp_visible_owner_lock());
                sicp_visible_owner().erase(list);
            }
            sicp_map_destroy(q);
        });
        visibleEntries.wrap(list, cap);
    }
    sicp_map* get() const { return handle.get(); }

    std::shared_ptr<sicp_map> handle;
    DeviceArray<HashEntry> visibleEntries;
};

// The C ABI's pose: R, its inverse as the caller computed it, t.
inline sicp_pose_f sicp_pose_of(const Matrix3f& R, const Matrix3f& Rinv, float3 t) {
    sicp_pose_f p;
    const float3 r[3] = {R.rowx, R.rowy, R.rowz};
    const float3 ri[3] = {Rinv.rowx, Rinv.rowy, Rinv.rowz};
    for (int i = 0; i < 3; ++i) {
        p.R[i * 3 + 0] = r[i].x;
        p.R[i * 3 + 1] = r[i].y;
        p.R[i * 3 + 2] = r[i].z;
        p.Rinv[i * 3 + 0] = ri[i].x;
        p.Rinv[i * 3 + 1] = ri[i].y;
        p.Rinv[i * 3 + 2] = ri[i].z;
    }
    p.t[0] = t.x;
    p.t[1] = t.y;
    p.t[2] = t.z;
    return p;
}

inline Matrix3f sicp_transpose(const Matrix3f& R) {
    Matrix3f T;
    T.rowx = make_float3(R.rowx.x, R.rowy.x, R.rowz.x);
    T.rowy = make_float3(R.rowx.y, R.rowy.y, R.rowz.y);
    T.rowz = make_float3(R.rowx.z, R.rowy.z, R.rowz.z);
    return T;
}
