// sicp_internal.hpp — host-side objects behind the opaque C ABI handles.
#pragma once

#include <vector>

#include "sicp_kernels.hpp"

struct sicp_map {
    sicp_map_config cfg;
    sicp::MapDev dev;
    sicp::MapCounters* ctr = nullptr;  // device
    sicp::AllocScratch alloc{};
    sicp_block* visible = nullptr;     // device, num_blocks
    uint32_t* unify_snap = nullptr;    // neighbour label pass: label snapshot, num_blocks * 512
    uint32_t* unify_mark = nullptr;    // per heap slot: epoch of the last snapshot
    uint32_t unify_epoch = 0;
    std::vector<void*> owned;          // every device allocation of the map
};

namespace sicp {

int map_alloc(const sicp_map_config& cfg, sicp_map** out);
void map_free(sicp_map* m);
// The neighbour label pass over the map's visible list (unify.hip).
int launch_unify(sicp_map* map, const int* gate, hipStream_t s);
// `queue_cleared`: the hand-over queue count was cleared by launch_fuse_count
// on the same stream.
int launch_raycast_gated(MapDev m, View<float4> vmap, View<float4> nmap, View<float> zmin,
                         View<float> zmax, PoseArg pose, float invfx, float invfy, float cx,
                         float cy, const int* gate, const uint32_t* nonempty, hipStream_t s,
                         bool queue_cleared = false, RayTuning tune = RayTuning{});

// ICPStep over the four maps with `partials_dev` as the workgroups' partial
// rows (sicp_icp_step and its variants).
int icp_step(sicp_image vcurr, sicp_image vlast, sicp_image ncurr, sicp_image nlast,
             const sicp_pose_f* curr, const sicp_pose_f* last, sicp_intrinsics K,
             void* partials_dev, float* sums_dev, float* host_sums, double* A_host,
             double* b_host, float* residual_host, bool k32, void* stream);
// The largest image sicp_icp_step_static takes (the kernels index pixels
// with 32-bit integers).
constexpr int64_t kIcpStaticMaxPixels = (int64_t)1 << 30;

// A pitched device image owned by the library.
template <class T> struct DevImage {
    void* ptr = nullptr;
    size_t pitch = 0;
    int cols = 0, rows = 0;
    int create(int c, int r, size_t elem = sizeof(T)) {
        cols = c;
        rows = r;
        hipError_t e = hipMallocPitch(&ptr, &pitch, (size_t)c * elem, (size_t)(r > 0 ? r : 1));
        if (e != hipSuccess) return hip_fail(e, "hipMallocPitch");
        return SICP_OK;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
    }
    View<T> v() const { return View<T>{(char*)ptr, pitch, cols, rows}; }
    sicp_image img() const { return sicp_image{ptr, pitch, cols, rows}; }
};

}  // namespace sicp
