// oracle_math.hpp — scalar vector algebra of the reference (Utility/VectorMath.h),
// templated on the scalar type so the same restatement serves the FP32
// same-order oracle (R = float) and the FP64 transcription (R = double).
//
// TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Every operator keeps the reference's evaluation order: a dot product is
// ((a.x*b.x + a.y*b.y) + a.z*b.z) (VectorMath.h:195-197), norm is sqrt(a*a)
// (:324-330), normalised is a per-component division by the norm (:340-346),
// Matrix3f * v is three row dot products (:368-374). Build with
// -ffp-contract=off so no multiply-add is fused.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>

namespace oracle {

template <class R> struct V3 { R x, y, z; };
template <class R> struct V4 { R x, y, z, w; };
template <class R> struct M3 { V3<R> r0, r1, r2; };
struct I3 { int x, y, z; };

template <class R> inline V3<R> v3(R x, R y, R z) { return V3<R>{x, y, z}; }
template <class R> inline V3<R> add(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> sub(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> neg(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <class R> inline V3<R> mul(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline V3<R> mul(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> inline V3<R> divs(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline R dot4(V4<R> a, V4<R> b) {
    return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w;
}
template <class R> inline V3<R> cross(V3<R> a, V3<R> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> inline R norm(V3<R> a) { return std::sqrt(dot(a, a)); }
template <class R> inline R norm4(V4<R> a) { return std::sqrt(dot4(a, a)); }
template <class R> inline V3<R> normalised(V3<R> a) { return divs(a, norm(a)); }
template <class R> inline V4<R> normalised4(V4<R> a) {
    R n = norm4(a);
    return {a.x / n, a.y / n, a.z / n, a.w / n};
}
template <class R> inline V3<R> mv(const M3<R>& m, V3<R> v) {
    return {dot(m.r0, v), dot(m.r1, v), dot(m.r2, v)};
}

// float -> int truncation toward zero with a defined result for NaN and
// out-of-range values (the reference's `(int)` casts are only ever applied to
// finite values whose result is not used otherwise; 0 for NaN, saturation).
template <class R> inline int f2i(R v) {
    if (!(v == v)) return 0;
    if (v >= (R)2147483520.0) return 2147483647;
    if (v <= (R)-2147483648.0) return (-2147483647 - 1);
    return (int)v;
}

// make_int3(float3): truncate, then step down where truncation rounded up
// (Utility/VectorMath.h:36-42) — a floor for finite in-range values.
template <class R> inline I3 floor_i3(V3<R> a) {
    I3 b{f2i(a.x), f2i(a.y), f2i(a.z)};
    b.x = ((R)b.x > a.x) ? b.x - 1 : b.x;
    b.y = ((R)b.y > a.y) ? b.y - 1 : b.y;
    b.z = ((R)b.z > a.z) ? b.z - 1 : b.z;
    return b;
}

template <class R> inline bool isnan_(R v) { return !(v == v); }

template <class R> inline R qnan() { return std::numeric_limits<R>::quiet_NaN(); }

// uchar narrowing of an int, the conversion make_uchar3((int)..) performs.
inline uint8_t u8_of_int(int v) { return (uint8_t)(uint32_t)v; }

template <class R> inline M3<R> m3_from(const float* r) {
    return {{(R)r[0], (R)r[1], (R)r[2]}, {(R)r[3], (R)r[4], (R)r[5]}, {(R)r[6], (R)r[7], (R)r[8]}};
}
template <class R> inline V3<R> v3_from(const float* t) { return {(R)t[0], (R)t[1], (R)t[2]}; }

// Pose in the precision of the restatement: rotation, its transpose, translation.
template <class R> struct Pose {
    M3<R> R_, Rinv;
    V3<R> t;
};
template <class R> inline Pose<R> pose_from_f(const float* Rm, const float* Rinv, const float* t) {
    return {m3_from<R>(Rm), m3_from<R>(Rinv), v3_from<R>(t)};
}
// From a 4x4 row-major double pose (camera -> world), rounded to R the way
// Frame::GpuRotation / GpuInvRotation / GpuTranslation do (Core/Frame.cc:626-645).
template <class R> inline Pose<R> pose_from_d(const double* T) {
    Pose<R> p;
    p.R_ = {{(R)T[0], (R)T[1], (R)T[2]}, {(R)T[4], (R)T[5], (R)T[6]}, {(R)T[8], (R)T[9], (R)T[10]}};
    p.Rinv = {{(R)T[0], (R)T[4], (R)T[8]}, {(R)T[1], (R)T[5], (R)T[9]}, {(R)T[2], (R)T[6], (R)T[10]}};
    p.t = {(R)T[3], (R)T[7], (R)T[11]};
    return p;
}

}  // namespace oracle
