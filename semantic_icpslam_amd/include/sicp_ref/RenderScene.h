// RenderScene.h — the reference's map entry points (Mapping/RenderScene.h:3-55)
// with their names, argument order and meaning, over the C ABI. Calls are
// synchronous, as in the reference. Arguments the reference passes but
// this build does not need are accepted and left untouched: `labels` (carried
// but never read by the reference's kernels, FuseMap.cu:19,1086),
// `renderingBlockList` / `noRenderingBlocks` (tiles go straight into the
// zRange images), and FuseMapColor's depthMax / depthMin (the reference uses
// the map's constants, FuseMap.cu:1080-1081; here the map's configuration).
#pragma once

#include <cstdio>

#include "DeviceArray.h"
#include "DeviceMap.h"
#include "Intrinsics.h"
#include "SafeCall.h"
#include "sicp.h"

typedef unsigned int uint;

namespace sicp_ref_detail {
inline void store_count(DeviceArray<uint>& dst, uint n) {
    if (dst.data && dst.size >= 1) dst.upload(&n, 1);
}
// "No visible blocks" is a message and an early return in the reference
// (FuseMap.cu:1033-1036,1117-1120), not a failure.
inline bool visible_status(int rc, const char* what) {
    if (rc == SICP_ERR_NO_VISIBLE) {
        std::printf("%s: no visible blocks\n", what);
        return false;
    }
    SicpCall(rc, what);
    return true;
}
}  // namespace sicp_ref_detail

// ResetMap (RenderScene.h:3; FuseMap.cu:1291-1304).
inline void ResetMap(DeviceMap map) {
    SicpCall(sicp_map_reset(map.get(), nullptr), "ResetMap");
    SafeCall(hipStreamSynchronize(nullptr));
}

// Raycast (RenderScene.h:13-18; RenderScene.cu:495-537).
inline void Raycast(DeviceMap map, DeviceArray2D<float4>& vmap, DeviceArray2D<float4>& nmap,
                    DeviceArray2D<float>& zRangeX, DeviceArray2D<float>& zRangeY, Matrix3f Rview,
                    Matrix3f RviewInv, float3 tview, float invfx, float invfy, float cx,
                    float cy) {
    const sicp_pose_f pose = sicp_pose_of(Rview, RviewInv, tview);
    SicpCall(sicp_raycast_inv(map.get(), vmap.image(), nmap.image(), zRangeX.image(),
                              zRangeY.image(), &pose, nullptr, invfx, invfy, cx, cy, nullptr),
             "Raycast");
    SafeCall(hipStreamSynchronize(nullptr));
}

// CreateRenderingBlocks (RenderScene.h:20-27; RenderScene.cu:210-289).
// `visibleBlocks` must be a map's `visibleEntries`; the map's own visible
// count of its last visibility pass is used (the value the reference's
// caller passes as noVisibleBlocks).
inline bool CreateRenderingBlocks(const DeviceArray<HashEntry>& visibleBlocks,
                                  DeviceArray2D<float>& zRangeX, DeviceArray2D<float>& zRangeY,
                                  const float& depthMax, const float& depthMin,
                                  DeviceArray<RenderingBlock>& renderingBlockList,
                                  DeviceArray<uint>& noRenderingBlocks, Matrix3f RviewInv,
                                  float3 tview, uint noVisibleBlocks, float fx, float fy,
                                  float cx, float cy) {
    (void)renderingBlockList;
    (void)noRenderingBlocks;
    (void)noVisibleBlocks;
    sicp_map* m = nullptr;
    {
        std::lock_guard<std::mutex> g(sicp_visible_owner_lock());
        auto it = sicp_visible_owner().find(visibleBlocks.data);
        if (it != sicp_visible_owner().end()) m = it->second;
    }
    if (!m) SafeCallFail(SICP_ERR_INVALID_ARGUMENT, "CreateRenderingBlocks: not a map's visibleEntries");
    const sicp_pose_f pose = sicp_pose_of(sicp_transpose(RviewInv), RviewInv, tview);
    int nonempty = 0;
    const int rc = sicp_create_rendering_blocks(m, zRangeX.image(), zRangeY.image(), depthMax,
                                                depthMin, &pose, nullptr,
                                                sicp_intrinsics{fx, fy, cx, cy}, &nonempty,
                                                nullptr);
    if (rc == SICP_ERR_NOTHING_TO_RENDER) return false;
    SicpCall(rc, "CreateRenderingBlocks");
    return nonempty != 0;
}

// CheckBlockVisibility (RenderScene.h:40-43; FuseMap.cu:991-1037).
inline void CheckBlockVisibility(DeviceMap map, DeviceArray<uint>& noVisibleBlocks, Matrix3f Rview,
                                 Matrix3f RviewInv, float3 tview, int cols, int rows, float fx,
                                 float fy, float cx, float cy, float depthMax, float depthMin,
                                 uint* host_data) {
    const sicp_pose_f pose = sicp_pose_of(Rview, RviewInv, tview);
    uint32_t n = 0;
    const int rc = sicp_check_block_visibility(map.get(), &pose, nullptr, cols, rows,
                                               sicp_intrinsics{fx, fy, cx, cy}, depthMax,
                                               depthMin, &n, nullptr);
    if (host_data) *host_data = n;
    sicp_ref_detail::store_count(noVisibleBlocks, n);
    sicp_ref_detail::visible_status(rc, "CheckBlockVisibility");
}

// MeshScene (RenderScene.h:29-38; MeshScene.cu:308-356). The three table
// arguments are accepted and not read: the library meshes with its own
// tables (sicp_mesh_tables), which give the same vertices, normals and
// colours and their own triangle connectivity. `vertex` / `normal` /
// `color` hold three entries per triangle; the smallest of them bounds the
// triangles written (the reference's MaxTriangles). `noOccupiedBlocks` gets
// the allocated blocks, `noTotalTriangles` the triangles written; no block
// list is needed (`blockPoses`). Returns the triangles written.
inline uint MeshScene(DeviceArray<uint>& noOccupiedBlocks, DeviceArray<uint>& noTotalTriangles,
                      DeviceMap map, const DeviceArray<int>& edgeTable,
                      const DeviceArray<int>& vertexTable, const DeviceArray2D<int>& triangleTable,
                      DeviceArray<float3>& normal, DeviceArray<float3>& vertex,
                      DeviceArray<uchar3>& color, DeviceArray<int3>& blockPoses) {
    (void)edgeTable;
    (void)vertexTable;
    (void)triangleTable;
    (void)blockPoses;
    size_t cap = vertex.size;
    if (normal.size < cap) cap = normal.size;
    if (color.size < cap) cap = color.size;
    const uint32_t max_tris = (uint32_t)(cap / 3);
    uint32_t n = 0;
    SicpCall(sicp_mesh_scene(map.get(), (float*)vertex.data, (float*)normal.data,
                             (uint8_t*)color.data, max_tris, &n, nullptr, nullptr),
             "MeshScene");
    sicp_map_stats st;
    SicpCall(sicp_map_get_stats(map.get(), &st, nullptr), "MeshScene");
    sicp_ref_detail::store_count(noOccupiedBlocks, st.allocated_blocks);
    sicp_ref_detail::store_count(noTotalTriangles, n);
    return n;
}

// FuseMapColor (RenderScene.h:45-55; FuseMap.cu:1039-1162).
inline void FuseMapColor(const DeviceArray2D<float>& depth, const DeviceArray2D<uchar3>& color,
                         const DeviceArray2D<float4>& nmap, const DeviceArray2D<unsigned char>& mask,
                         const DeviceArray<int>& labels, int numDetection,
                         DeviceArray<uint>& noVisibleBlocks, Matrix3f Rview, Matrix3f RviewInv,
                         float3 tview, DeviceMap map, float fx, float fy, float cx, float cy,
                         float depthMax, float depthMin, uint* host_data) {
    (void)labels;
    (void)depthMax;
    (void)depthMin;
    const sicp_pose_f pose = sicp_pose_of(Rview, RviewInv, tview);
    uint32_t n = 0;
    const int rc = sicp_fuse_map_color(map.get(), depth.image(), color.image(), nmap.image(),
                                       mask.image(), numDetection, &pose, nullptr,
                                       sicp_intrinsics{fx, fy, cx, cy}, &n, nullptr);
    if (host_data) *host_data = n;
    sicp_ref_detail::store_count(noVisibleBlocks, n);
    if (sicp_ref_detail::visible_status(rc, "FuseMapColor"))
        SafeCall(hipStreamSynchronize(nullptr));
}
