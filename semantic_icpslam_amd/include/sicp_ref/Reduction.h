// Reduction.h — the reference's pyramid and ICP entry points
// (Tracking/Reduction.h:18-42,74-79) with their names, argument order and
// meaning, over the C ABI (include/sicp.h). Calls are synchronous on the
// null stream, as in the reference; failures go through SafeCall.
// Outputs that were never created are created at the size the call implies.
#pragma once

#include "DeviceArray.h"
#include "DeviceMap.h"
#include "Intrinsics.h"
#include "SafeCall.h"
#include "sicp.h"

namespace sicp_ref_detail {
template <class T> inline void ensure(DeviceArray2D<T>& a, int cols, int rows) {
    if (!a.data || a.cols != cols || a.rows != rows) a.create(cols, rows);
}
inline void sync() { SafeCall(hipStreamSynchronize(nullptr)); }
}  // namespace sicp_ref_detail

// FilterDepth (Reduction.h:18-20; Pyrdown.cu:72-121).
inline void FilterDepth(const DeviceArray2D<unsigned short>& depth, DeviceArray2D<float>& rawDepth,
                        DeviceArray2D<float>& filteredDepth, float depthScale,
                        float depthCutoff) {
    sicp_ref_detail::ensure(rawDepth, depth.cols, depth.rows);
    sicp_ref_detail::ensure(filteredDepth, depth.cols, depth.rows);
    SicpCall(sicp_filter_depth(depth.image(), rawDepth.image(), filteredDepth.image(), depthScale,
                               depthCutoff, nullptr),
             "FilterDepth");
    sicp_ref_detail::sync();
}

// ComputeVMap (Reduction.h:22-24; Pyrdown.cu:123-158).
inline void ComputeVMap(const DeviceArray2D<float>& depth, DeviceArray2D<float4>& vmap, float fx,
                        float fy, float cx, float cy, float depthCutoff) {
    sicp_ref_detail::ensure(vmap, depth.cols, depth.rows);
    SicpCall(sicp_compute_vmap(depth.image(), vmap.image(), fx, fy, cx, cy, depthCutoff, nullptr),
             "ComputeVMap");
    sicp_ref_detail::sync();
}

// ComputeNMap (Reduction.h:26-27; Pyrdown.cu:160-192).
inline void ComputeNMap(const DeviceArray2D<float4>& vmap, DeviceArray2D<float4>& nmap) {
    sicp_ref_detail::ensure(nmap, vmap.cols, vmap.rows);
    SicpCall(sicp_compute_nmap(vmap.image(), nmap.image(), nullptr), "ComputeNMap");
    sicp_ref_detail::sync();
}

// PyrDownGauss (Reduction.h:29-32; Pyrdown.cu:194-248). The destination is
// half the source, rounded down.
inline void PyrDownGauss(const DeviceArray2D<float>& src, DeviceArray2D<float>& dst) {
    sicp_ref_detail::ensure(dst, src.cols / 2, src.rows / 2);
    SicpCall(sicp_pyrdown_gauss_f32(src.image(), dst.image(), nullptr), "PyrDownGauss");
    sicp_ref_detail::sync();
}
inline void PyrDownGauss(const DeviceArray2D<unsigned char>& src,
                         DeviceArray2D<unsigned char>& dst) {
    sicp_ref_detail::ensure(dst, src.cols / 2, src.rows / 2);
    SicpCall(sicp_pyrdown_gauss_u8(src.image(), dst.image(), nullptr), "PyrDownGauss");
    sicp_ref_detail::sync();
}

// ImageToIntensity (Reduction.h:34-35; Pyrdown.cu:250-272).
inline void ImageToIntensity(const DeviceArray2D<uchar3>& rgb, DeviceArray2D<unsigned char>& image) {
    sicp_ref_detail::ensure(image, rgb.cols, rgb.rows);
    SicpCall(sicp_image_to_intensity(rgb.image(), image.image(), nullptr), "ImageToIntensity");
    sicp_ref_detail::sync();
}

// ComputeDerivativeImage (Reduction.h:37-38; Pyrdown.cu:274-316).
inline void ComputeDerivativeImage(DeviceArray2D<unsigned char>& image, DeviceArray2D<short>& dx,
                                   DeviceArray2D<short>& dy) {
    sicp_ref_detail::ensure(dx, image.cols, image.rows);
    sicp_ref_detail::ensure(dy, image.cols, image.rows);
    SicpCall(sicp_compute_derivative_image(image.image(), dx.image(), dy.image(), nullptr),
             "ComputeDerivativeImage");
    sicp_ref_detail::sync();
}

// ResizeMap (Reduction.h:40-42; Pyrdown.cu:318-358).
inline void ResizeMap(const DeviceArray2D<float4>& vsrc, const DeviceArray2D<float4>& nsrc,
                      DeviceArray2D<float4>& vdst, DeviceArray2D<float4>& ndst) {
    sicp_ref_detail::ensure(vdst, vsrc.cols / 2, vsrc.rows / 2);
    sicp_ref_detail::ensure(ndst, vsrc.cols / 2, vsrc.rows / 2);
    SicpCall(sicp_resize_map(vsrc.image(), nsrc.image(), vdst.image(), ndst.image(), nullptr),
             "ResizeMap");
    sicp_ref_detail::sync();
}

// ICPStep (Reduction.h:74-79; Reduction.cu:328-385). `sum` is the partial-row
// scratch (grown to sicp_icp_partials_size() bytes when smaller), `out` the
// 29 device sums; on the host: matrixA_host (6x6 row-major, mirrored),
// vectorB_host[6], residual = {sum r^2, count}.
inline void ICPStep(DeviceArray2D<float4>& nextVMap, DeviceArray2D<float4>& lastVMap,
                    DeviceArray2D<float4>& nextNMap, DeviceArray2D<float4>& lastNMap,
                    Matrix3f Rcurr, float3 tcurr, Matrix3f Rlast, Matrix3f RlastInv, float3 tlast,
                    Intrinsics K, DeviceArray2D<float>& sum, DeviceArray<float>& out,
                    float* residual, double* matrixA_host, double* vectorB_host) {
    const size_t need = sicp_icp_partials_size();
    if (!sum.data || sum.step * (size_t)sum.rows < need) {
        const size_t row = 29 * sizeof(float);
        sum.create(29, (int)((need + row - 1) / row));
    }
    if (!out.data || out.size < 29) out.create(29);
    const sicp_pose_f curr = sicp_pose_of(Rcurr, sicp_transpose(Rcurr), tcurr);
    const sicp_pose_f last = sicp_pose_of(Rlast, RlastInv, tlast);
    float host_sums[29];
    SicpCall(sicp_icp_step(nextVMap.image(), lastVMap.image(), nextNMap.image(), lastNMap.image(),
                           &curr, &last, K.abi(), sum.data, (float*)out.data, host_sums,
                           matrixA_host, vectorB_host, residual, nullptr),
             "ICPStep");
}

// SO3Step (Reduction.h:67-72; Reduction.cu:162-211). `sum` is the
// partial-row scratch (grown to sicp_icp_partials_size() bytes when
// smaller), `out` the 11 device sums; on the host: matrixA_host (3x3
// row-major, mirrored), vectorB_host[3], residual = {sum r^2, count}.
inline void SO3Step(const DeviceArray2D<unsigned char>& nextImage,
                    const DeviceArray2D<unsigned char>& lastImage,
                    const DeviceArray2D<short>& dIdx, const DeviceArray2D<short>& dIdy,
                    Matrix3f RcurrInv, Matrix3f Rlast, Intrinsics K, DeviceArray2D<float>& sum,
                    DeviceArray<float>& out, float* residual, double* matrixA_host,
                    double* vectorB_host) {
    const size_t need = sicp_icp_partials_size();
    if (!sum.data || sum.step * (size_t)sum.rows < need) {
        const size_t row = 11 * sizeof(float);
        sum.create(11, (int)((need + row - 1) / row));
    }
    if (!out.data || out.size < 11) out.create(11);
    const float3 zero = make_float3(0, 0, 0);
    const sicp_pose_f curr = sicp_pose_of(sicp_transpose(RcurrInv), RcurrInv, zero);
    const sicp_pose_f last = sicp_pose_of(Rlast, sicp_transpose(Rlast), zero);
    float host_sums[11];
    SicpCall(sicp_so3_step(nextImage.image(), lastImage.image(), dIdx.image(), dIdy.image(), &curr,
                           &last, K.abi(), sum.data, (float*)out.data, host_sums, matrixA_host,
                           vectorB_host, residual, nullptr),
             "SO3Step");
}

// RGBStep (Reduction.h:81-90; Reduction.cu:575-642). `sum` is the scratch
// (grown to sicp_rgb_scratch_size(cols, rows) bytes when smaller), `out` the
// 29 device sums, `outRes` the correspondence count and int sum of squared
// differences (device, 2 ints); `sumRes` is not needed and left untouched.
// On the host: matrixA_host (6x6 row-major, mirrored), vectorB_host[6],
// residual = {sum r^2, count}.
inline void RGBStep(const DeviceArray2D<unsigned char>& nextImage,
                    const DeviceArray2D<unsigned char>& lastImage,
                    const DeviceArray2D<float4>& nextVMap, const DeviceArray2D<float4>& lastVMap,
                    const DeviceArray2D<short>& dIdx, const DeviceArray2D<short>& dIdy,
                    Matrix3f Rcurr, Matrix3f RcurrInv, Matrix3f Rlast, Matrix3f RlastInv,
                    float3 tcurr, float3 tlast, Intrinsics K, DeviceArray2D<float>& sum,
                    DeviceArray<float>& out, DeviceArray2D<int>& sumRes, DeviceArray<int>& outRes,
                    float* residual, double* matrixA_host, double* vectorB_host) {
    (void)sumRes;
    const size_t need = sicp_rgb_scratch_size(nextImage.cols, nextImage.rows);
    if (!sum.data || sum.step * (size_t)sum.rows < need) {
        const size_t row = 29 * sizeof(float);
        sum.create(29, (int)((need + row - 1) / row));
    }
    if (!out.data || out.size < 29) out.create(29);
    if (!outRes.data || outRes.size < 2) outRes.create(2);
    const sicp_pose_f curr = sicp_pose_of(Rcurr, RcurrInv, tcurr);
    const sicp_pose_f last = sicp_pose_of(Rlast, RlastInv, tlast);
    float host_sums[29];
    int32_t corresp[2];
    SicpCall(sicp_rgb_step(nextImage.image(), lastImage.image(), nextVMap.image(),
                           lastVMap.image(), dIdx.image(), dIdy.image(), &curr, &last, K.abi(),
                           sum.data, sum.step * (size_t)sum.rows, (float*)out.data, host_sums,
                           corresp, nullptr, matrixA_host, vectorB_host, residual, nullptr),
             "RGBStep");
    outRes.upload(corresp, 2);
}
