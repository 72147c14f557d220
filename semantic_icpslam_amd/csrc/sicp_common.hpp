// sicp_common.hpp — shared definitions of the HIP implementation: device
// vector algebra with the reference's evaluation order, pitched image views,
// pose loading, status / error plumbing.
//
// The reference's vector operators (Utility/VectorMath.h) overload `*` on
// float3 as a dot product; HIP's own float3 overloads it element-wise. The
// kernels therefore use the plain struct `v3` below and named functions whose
// arithmetic order matches VectorMath.h (dot = (x*x' + y*y') + z*z',
// norm = sqrt(dot), normalised = per-component division by the norm,
// Matrix3f * v = three row dots). Everything is compiled with
// -ffp-contract=off and correctly rounded division / square root.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>
#include <string>

#include "../../include/sicp.h"
#include "sicp_expf.hpp"

#define SICP_DEV __device__ __forceinline__

namespace sicp {

constexpr int kBlockSide = 8;     // DeviceMap::BlockSize (Mapping/DeviceMap.h:166)
constexpr int kBlockVoxels = 512; // DeviceMap::BlockSize3

// ---------------------------------------------------------------- vectors --
struct v3 {
    float x, y, z;
};
struct i3 {
    int x, y, z;
};

SICP_DEV v3 mk(float x, float y, float z) { return v3{x, y, z}; }
SICP_DEV v3 add(v3 a, v3 b) { return v3{a.x + b.x, a.y + b.y, a.z + b.z}; }
SICP_DEV v3 sub(v3 a, v3 b) { return v3{a.x - b.x, a.y - b.y, a.z - b.z}; }
SICP_DEV v3 neg(v3 a) { return v3{-a.x, -a.y, -a.z}; }
SICP_DEV v3 scale(v3 a, float s) { return v3{a.x * s, a.y * s, a.z * s}; }
SICP_DEV v3 scale(float s, v3 a) { return v3{s * a.x, s * a.y, s * a.z}; }
SICP_DEV v3 divs(v3 a, float s) { return v3{a.x / s, a.y / s, a.z / s}; }
SICP_DEV float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
SICP_DEV v3 cross(v3 a, v3 b) {
    return v3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
SICP_DEV float norm(v3 a) { return sqrtf(dot(a, a)); }
SICP_DEV v3 normalised(v3 a) { return divs(a, norm(a)); }
SICP_DEV v3 xyz(float4 a) { return v3{a.x, a.y, a.z}; }

struct m3 {
    v3 r0, r1, r2;
};
SICP_DEV v3 mv(const m3& m, v3 v) { return v3{dot(m.r0, v), dot(m.r1, v), dot(m.r2, v)}; }

// make_int3(float3) of the reference (VectorMath.h:36-42): truncate, then
// step down where truncation rounded up — a floor for in-range values.
// v_cvt_i32_f32 truncates, saturates and maps NaN to 0.
SICP_DEV int f2i(float v) { return (int)v; }
SICP_DEV i3 floor_i3(v3 a) {
    i3 b{f2i(a.x), f2i(a.y), f2i(a.z)};
    b.x = ((float)b.x > a.x) ? b.x - 1 : b.x;
    b.y = ((float)b.y > a.y) ? b.y - 1 : b.y;
    b.z = ((float)b.z > a.z) ? b.z - 1 : b.z;
    return b;
}

SICP_DEV float qnanf() { return __builtin_bit_cast(float, 0x7fffffffu); }
SICP_DEV bool isnan_(float v) { return !(v == v); }

// A load issued where it stands in the source. The compiler does not move
// relaxed atomic loads past branches, so a kernel can start every
// independent load of its prologue (gates, counters, poses, first inputs)
// before the first decision that needs one of them: each is a round trip to
// memory that an earlier launch wrote, and early returns in between would
// otherwise order them one after another.
template <class T> SICP_DEV T early_load(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}
// A device flag written by an earlier launch (a gate: 0 skips the launch's
// work), or 1 when there is none.
SICP_DEV int flag_or_one(const int* flag) { return flag ? early_load(flag) : 1; }
SICP_DEV bool gate_closed(const int* gate) { return flag_or_one(gate) == 0; }

// ------------------------------------------------------------------ poses --
struct PoseF {
    m3 R, Rinv;
    v3 t;
};
// A pose argument: by value, or through a device pointer when `dev` is set.
struct PoseArg {
    sicp_pose_f val;
    const sicp_pose_f* dev;
};
SICP_DEV PoseF to_pose(const sicp_pose_f& p) {
    PoseF o;
    o.R = m3{v3{p.R[0], p.R[1], p.R[2]}, v3{p.R[3], p.R[4], p.R[5]}, v3{p.R[6], p.R[7], p.R[8]}};
    o.Rinv = m3{v3{p.Rinv[0], p.Rinv[1], p.Rinv[2]}, v3{p.Rinv[3], p.Rinv[4], p.Rinv[5]},
                v3{p.Rinv[6], p.Rinv[7], p.Rinv[8]}};
    o.t = v3{p.t[0], p.t[1], p.t[2]};
    return o;
}
SICP_DEV PoseF load_pose(const PoseArg& a) { return to_pose(a.dev ? *a.dev : a.val); }

inline PoseArg make_pose_arg(const sicp_pose_f* host, const sicp_pose_f* dev) {
    PoseArg a{};
    if (host) a.val = *host;
    a.dev = dev;
    return a;
}

// ------------------------------------------------------------ image views --
template <class T> struct View {
    char* ptr;
    size_t pitch;
    int cols, rows;
    SICP_DEV T& at(int y, int x) const { return ((T*)(ptr + (size_t)y * pitch))[x]; }
};
template <class T> inline View<T> view(const sicp_image& im) {
    return View<T>{(char*)im.ptr, im.pitch, im.cols, im.rows};
}

// ----------------------------------------------------------- error plumbing --
void set_error(const std::string& msg);
int hip_fail(hipError_t e, const char* what);
#define SICP_HIP(expr)                                        \
    do {                                                      \
        hipError_t _e = (expr);                               \
        if (_e != hipSuccess) return ::sicp::hip_fail(_e, #expr); \
    } while (0)
#define SICP_LAUNCH_CHECK(name)                               \
    do {                                                      \
        hipError_t _e = hipGetLastError();                    \
        if (_e != hipSuccess) return ::sicp::hip_fail(_e, name); \
    } while (0)
int invalid(const char* msg);
bool image_ok(const sicp_image& im, size_t elem, int cols, int rows);

inline unsigned div_up(long a, long b) { return (unsigned)((a + b - 1) / b); }

}  // namespace sicp
