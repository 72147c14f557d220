// icp_static.hip — ICPStep for callers that cannot allocate device memory.
//
// sicp_icp_step takes its per-workgroup partial rows in a caller-provided
// workspace. This variant keeps them in a file-scope device array instead,
// so a call allocates nothing: it serves code under hipGraph capture, tight
// serving loops and harnesses that forbid hipMalloc. Each GPU holds its own
// copy of the array for the life of the process; its address on a device is
// looked up once per device with hipGetSymbolAddress and cached, and only
// that address reaches the kernels.
#include <mutex>

#include "sicp_internal.hpp"

namespace {

// 29 sums x 512 workgroup rows (FP64 or the compensated pair, 8 bytes each):
// the ICP kernels launch at most kIcpMaxBlocks workgroups for any image size,
// each writing its own complete row before the final sum reads it, so the
// array never needs clearing between calls.
constexpr size_t kScratchWords = (size_t)sicp::kIcpSums * sicp::kIcpMaxBlocks;
static __device__ __align__(256) double scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device.
int scratch_address(void** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return sicp::hip_fail(e, "icp_step_static: hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices)
        return sicp::invalid("icp_step_static: device index above 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return sicp::hip_fail(e, "icp_step_static: hipGetSymbolAddress");
        if (!p) return sicp::invalid("icp_step_static: no device address for the scratch");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return SICP_OK;
}

}  // namespace

extern "C" {

int sicp_icp_step_static(sicp_image vcurr, sicp_image vlast, sicp_image ncurr, sicp_image nlast,
                         const sicp_pose_f* curr, const sicp_pose_f* last, sicp_intrinsics K,
                         float* sums_dev, float* host_sums, double* A_host, double* b_host,
                         float* residual_host, void* stream) {
    if ((int64_t)vcurr.cols * (int64_t)vcurr.rows > sicp::kIcpStaticMaxPixels)
        return sicp::invalid("icp_step_static: images above 2^30 pixels are not supported");
    void* partials = nullptr;
    const int rc = scratch_address(&partials);
    if (rc) return rc;
    return sicp::icp_step(vcurr, vlast, ncurr, nlast, curr, last, K, partials, sums_dev,
                          host_sums, A_host, b_host, residual_host, false, stream);
}

}  // extern "C"
