// Output epilogue for gfx950: colour matrix (core/cpu_accelerator.cpp:2874-2897),
// colour-space gamma (:2903-2976, colour-space codes core/metal/constants.h:28-38)
// and the 8/16-bit output of the processed image (core/python_bindings.cpp:235-257).
// Pointwise, float4-free AoS RGB; the fused form reads 12 B/px once and writes
// the final samples (3, 6 or 12 B/px).
#include "lre_common.h"
#include "lre_epilogue.h"
#include "lre_internal.h"

namespace {

constexpr int BLOCK = 256;

LRE_DEV float clamp01(float v) { return v < 0.f ? 0.f : (1.f < v ? 1.f : v); }

// one pixel: optional matrix, optional gamma, then float or quantized store
__global__ __launch_bounds__(BLOCK) void epilogue_kernel(const float *in, void *out, size_t npx,
                                                         lre::Epilogue E, int bits) {
    for (size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x; i < npx;
         i += (size_t)gridDim.x * BLOCK) {
        float r = in[3 * i], g = in[3 * i + 1], b = in[3 * i + 2];
        lre::epilogue_rgb(r, g, b, E);
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
    const lre::Epilogue E = lre_make_epilogue(transform12, gamma_power, gamma_slope, output_color_space);
    hipLaunchKernelGGL(epilogue_kernel, dim3(grid_for(npx)), dim3(BLOCK), 0, s, d_in, d_out, npx, E, bits);
    LRE_HIP_CHECK(hipGetLastError());
    return LRE_OK;
}

lre::Epilogue lre_make_epilogue(const float *transform12, float gamma_power, float gamma_slope,
                                int output_color_space) {
    lre::Epilogue E{};
    if (transform12)
        for (int k = 0; k < 12; k++) E.t[k] = transform12[k];
    E.use_matrix = transform12 ? 1 : 0;
    E.gp = gamma_power;
    E.gs = gamma_slope;
    E.cs = output_color_space;
    return E;
}
