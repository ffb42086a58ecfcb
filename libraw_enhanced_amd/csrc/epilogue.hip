// Output epilogue for gfx950: colour matrix (core/cpu_accelerator.cpp:2874-2897),
// colour-space gamma (:2903-2976, colour-space codes core/metal/constants.h:28-38)
// and the 8/16-bit output of the processed image (core/python_bindings.cpp:235-257).
// Pointwise, float4-free AoS RGB; the fused form reads 12 B/px once and writes
// the final samples (3, 6 or 12 B/px).
#include "lre_common.h"
#include "lre_internal.h"

namespace {

constexpr int BLOCK = 256;

struct Mat {
    float t[12];
};

// powf evaluated in double and rounded once: the correctly rounded float
// result wherever the double result is not within an ulp of a rounding tie
LRE_DEV float pow_f(float v, float p) { return (float)pow((double)v, (double)p); }

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

LRE_DEV float clamp01(float v) { return v < 0.f ? 0.f : (1.f < v ? 1.f : v); }

// one pixel: optional matrix, optional gamma, then float or quantized store
__global__ __launch_bounds__(BLOCK) void epilogue_kernel(const float *in, void *out, size_t npx,
                                                         Mat M, int use_matrix, float gp, float gs,
                                                         int cs, int bits) {
    for (size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x; i < npx;
         i += (size_t)gridDim.x * BLOCK) {
        float r = in[3 * i], g = in[3 * i + 1], b = in[3 * i + 2];
        if (use_matrix) {
            const float *t = M.t;
            const float r2 = t[0] * r + t[1] * g + t[2] * b + t[3];
            const float g2 = t[4] * r + t[5] * g + t[6] * b + t[7];
            const float b2 = t[8] * r + t[9] * g + t[10] * b + t[11];
            r = r2; g = g2; b = b2;
        }
        if (cs >= 0) {
            r = gamma_value(r, gp, gs, cs);
            g = gamma_value(g, gp, gs, cs);
            b = gamma_value(b, gp, gs, cs);
        }
        if (bits == 8) {
            uint8_t *o = static_cast<uint8_t *>(out) + 3 * i;
            o[0] = (uint8_t)(clamp01(r) * 255.0f + 0.5f);
            o[1] = (uint8_t)(clamp01(g) * 255.0f + 0.5f);
            o[2] = (uint8_t)(clamp01(b) * 255.0f + 0.5f);
        } else if (bits == 16) {
            uint16_t *o = static_cast<uint16_t *>(out) + 3 * i;
            o[0] = (uint16_t)(clamp01(r) * 65535.0f + 0.5f);
            o[1] = (uint16_t)(clamp01(g) * 65535.0f + 0.5f);
            o[2] = (uint16_t)(clamp01(b) * 65535.0f + 0.5f);
        } else {
            float *o = static_cast<float *>(out) + 3 * i;
            o[0] = r;
            o[1] = g;
            o[2] = b;
        }
    }
}

int grid_for(size_t n) {
    size_t b = (n + BLOCK - 1) / BLOCK;
    if (b > 65536) b = 65536;
    return (int)(b > 0 ? b : 1);
}

}  // namespace

int lre_launch_epilogue(const float *d_in, void *d_out, size_t npx, const float *transform12,
                        float gamma_power, float gamma_slope, int output_color_space, int bits,
                        hipStream_t s) {
    if (bits != 8 && bits != 16 && bits != 32) {
        lre_set_error("output bits must be 8, 16 or 32 (got %d)", bits);
        return LRE_ERR_ARG;
    }
    Mat M{};
    if (transform12)
        for (int k = 0; k < 12; k++) M.t[k] = transform12[k];
    hipLaunchKernelGGL(epilogue_kernel, dim3(grid_for(npx)), dim3(BLOCK), 0, s, d_in, d_out, npx, M,
                       transform12 ? 1 : 0, gamma_power, gamma_slope, output_color_space, bits);
    LRE_HIP_CHECK(hipGetLastError());
    return LRE_OK;
}
