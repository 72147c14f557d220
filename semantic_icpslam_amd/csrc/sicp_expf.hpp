// sicp_expf.hpp — the exponential used by the bilateral depth filter.
//
// The reference calls the platform `exp` on a float argument inside
// FilterDepthKernel (Tracking/Pyrdown.cu:100). Platform exponentials differ by
// an ulp or two between a CPU libm and the GPU device library, which would make
// every value downstream of the filtered depth (vertex / normal maps, the label
// gate of the fusion) differ between the GPU and a CPU restatement. This file
// defines one exponential, written from the mathematics, that both compile to
// the same sequence of IEEE float operations (build with -ffp-contract=off):
//
//   exp(x) = 2^k * exp(r),  k = round(x / ln 2),  r = x - k ln2_hi - k ln2_lo
//
// (Cody & Waite, "Software Manual for the Elementary Functions", 1980: the
// two-constant split keeps r exact to float precision), and exp(r) by its
// Taylor series to degree 7 in Horner form. |r| <= ln(2)/2, so the truncation
// error r^8/8! is below 6e-9, under half an ulp of 1; the result is within
// 2 ulp of the correctly rounded exponential for x in [-87, 88].
#pragma once

#if defined(__HIPCC__)
#define SICP_HD __host__ __device__
#else
#define SICP_HD
#endif

namespace sicp {

SICP_HD inline float bits_to_float(unsigned int u) {
    return __builtin_bit_cast(float, u);
}

SICP_HD inline float expf_cw(float x) {
    if (x != x) return x;                       // NaN propagates
    if (x > 88.7f) return bits_to_float(0x7f800000u);  // +inf
    if (x < -103.9f) return 0.0f;               // below the smallest denormal
    const float log2e = 1.44269504088896341f;
    const float ln2_hi = 0.693359375f;          // 0x3f318000, 9 significant bits
    const float ln2_lo = -2.12194440e-4f;       // ln 2 - ln2_hi
    float kf = x * log2e;
    int k = (int)(kf + (kf < 0.0f ? -0.5f : 0.5f));
    float fk = (float)k;
    float r = (x - fk * ln2_hi) - fk * ln2_lo;
    // Taylor series of exp(r) to r^7, Horner form.
    float p = 1.0f / 5040.0f;
    p = p * r + 1.0f / 720.0f;
    p = p * r + 1.0f / 120.0f;
    p = p * r + 1.0f / 24.0f;
    p = p * r + 1.0f / 6.0f;
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    // Scale by 2^k. For k < -126 split the power so the final product rounds
    // once into the denormal range.
    if (k > 127) {
        return (p * 2.0f) * bits_to_float((unsigned int)(k - 1 + 127) << 23);
    }
    if (k >= -126) {
        return p * bits_to_float((unsigned int)(k + 127) << 23);
    }
    float s1 = bits_to_float((unsigned int)(k + 64 + 127) << 23);
    float s2 = bits_to_float((unsigned int)(-64 + 127) << 23);
    return (p * s1) * s2;
}

}  // namespace sicp
