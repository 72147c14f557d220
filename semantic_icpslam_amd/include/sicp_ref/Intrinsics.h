// Intrinsics.h — camera intrinsics of the reference (Utility/Intrinsics.h:3-18):
// operator()(scale) gives the level-`scale` pyramid intrinsics, K / 2^scale
// in float (cx / 2^s, not (cx + 0.5) / 2^s - 0.5).
#pragma once

#include "sicp.h"

struct Intrinsics {
    Intrinsics() : fx(0), fy(0), cx(0), cy(0) {}
    Intrinsics(float fx_, float fy_, float cx_, float cy_) : fx(fx_), fy(fy_), cx(cx_), cy(cy_) {}
    Intrinsics operator()(int scale) const {
        const int i = 1 << scale;
        return Intrinsics(fx / i, fy / i, cx / i, cy / i);
    }
    sicp_intrinsics abi() const { return sicp_intrinsics{fx, fy, cx, cy}; }
    float fx, fy, cx, cy;
};
