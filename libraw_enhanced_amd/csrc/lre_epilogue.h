// Per-pixel output epilogue shared by the epilogue kernel and the stages that
// fuse it into their store: colour matrix (core/cpu_accelerator.cpp:2874-2897)
// then colour-space gamma (:2903-2976, colour-space codes
// core/metal/constants.h:28-38).
#pragma once
#include "lre_common.h"

namespace lre {

struct Epilogue {
    float t[12];     // row-major 3x4 colour matrix
    int use_matrix;  // 0: no matrix
    float gp, gs;    // gamma power and slope of the custom curve
    int cs;          // output colour space; < 0: no gamma
    bool active() const { return use_matrix || cs >= 0; }
};

// powf for v > 0 (every caller's branch), evaluated in double as
// exp2(p * log2 v) and rounded once.  The double path carries a relative error
// near 1e-14, far below half a float ulp, so the result is the correctly
// rounded powf except where the exact value lies that close to a rounding tie.
// pow(1, p) and pow(v, 0) are exactly 1 (the product form would give 0 * inf).
LRE_DEV float pow_f(float v, float p) {
    if (v == 1.0f || p == 0.0f) return 1.0f;
    return (float)exp2((double)p * log2((double)v));
}

LRE_DEV float gamma_value(float v, float gp, float gs, int cs) {
    switch (cs) {
    case 1:  // sRGB
    case 7:  // P3-D65
        return (v <= 0.0031308f) ? 12.92f * v : 1.055f * pow_f(v, 1.0f / 2.4f) - 0.055f;
    case 2:  // Adobe RGB
    case 3:  // wide-gamut RGB
        return v <= 0.0f ? 0.0f : pow_f(v, 1.0f / 2.222f);
    case 4:  // ProPhoto RGB
        return v <= 0.0f ? 0.0f : pow_f(v, 1.0f / 1.8f);
    case 8: {  // Rec.2020
        const float a = 1.09929682680944f, b = 0.018053968510807f;
        return (v < b) ? 4.5f * v : a * pow_f(v, 0.45f) - (a - 1.0f);
    }
    default:
        if (v <= 0.0f) return 0.0f;
        if (gs <= 0.f) return pow_f(v, 1.0f / gp);
        return (v < 1.0f / gs) ? v * gs : pow_f(v, 1.0f / gp);
    }
}

// matrix (when set), then gamma (when cs >= 0), on one pixel
LRE_DEV void epilogue_rgb(float &r, float &g, float &b, const Epilogue &E) {
    if (E.use_matrix) {
        const float *t = E.t;
        const float r2 = t[0] * r + t[1] * g + t[2] * b + t[3];
        const float g2 = t[4] * r + t[5] * g + t[6] * b + t[7];
        const float b2 = t[8] * r + t[9] * g + t[10] * b + t[11];
        r = r2;
        g = g2;
        b = b2;
    }
    if (E.cs >= 0) {
        r = gamma_value(r, E.gp, E.gs, E.cs);
        g = gamma_value(g, E.gp, E.gs, E.cs);
        b = gamma_value(b, E.gp, E.gs, E.cs);
    }
}

}  // namespace lre
