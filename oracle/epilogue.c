/*
 * TEST INFRASTRUCTURE ONLY — CPU restatement of the output epilogue of the
 * develop pipeline.  Never called by the product.
 *
 *  oracle_convert_color_space  convert_color_space, core/cpu_accelerator.cpp:2874-2897:
 *                              out[c] = t[c][0]*r + t[c][1]*g + t[c][2]*b + t[c][3],
 *                              summed left to right
 *  oracle_gamma_correct        gamma_correct, core/cpu_accelerator.cpp:2903-2976: per
 *                              channel by output colour space (constants
 *                              core/metal/constants.h:28-38): sRGB / P3-D65 piecewise
 *                              sRGB curve, Adobe / wide-gamut pure power 2.222,
 *                              ProPhoto 1.8, Rec.2020 piecewise curve, otherwise power
 *                              gamma_power with linear slope gamma_slope; powf is the
 *                              C library's, as std::pow(float, float) in the reference
 *  oracle_quantize             the 8/16-bit output of core/python_bindings.cpp:235-257:
 *                              clamp(v, 0, 1) * 255 (or 65535) + 0.5, truncated
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>

#include "oracle.h"

int oracle_convert_color_space(const float *in3, float *out3, int w, int h, const float *t) {
    const size_t n = (size_t)w * h;
    for (size_t i = 0; i < n; i++) {
        const float r = in3[3 * i], g = in3[3 * i + 1], b = in3[3 * i + 2];
        out3[3 * i] = t[0] * r + t[1] * g + t[2] * b + t[3];
        out3[3 * i + 1] = t[4] * r + t[5] * g + t[6] * b + t[7];
        out3[3 * i + 2] = t[8] * r + t[9] * g + t[10] * b + t[11];
    }
    return 0;
}

static float srgb_encode(float v) {
    return (v <= 0.0031308f) ? 12.92f * v : 1.055f * powf(v, 1.0f / 2.4f) - 0.055f;
}

static float rec2020_encode(float v) {
    const float a = 1.09929682680944f, b = 0.018053968510807f;
    return (v < b) ? 4.5f * v : a * powf(v, 0.45f) - (a - 1.0f);
}

static float pure_power(float v, float p) {
    if (v <= 0.0f) return 0.0f;
    return powf(v, 1.0f / p);
}

static float power_with_slope(float v, float p, float s) {
    if (v <= 0.0f) return 0.0f;
    if (s <= 0.f) return powf(v, 1.0f / p);
    return (v < 1.0f / s) ? v * s : powf(v, 1.0f / p);
}

float oracle_gamma_value(float v, float gamma_power, float gamma_slope, int cs) {
    switch (cs) {
    case 1: /* sRGB */
    case 7: /* P3-D65 */
        return srgb_encode(v);
    case 2: /* Adobe RGB */
    case 3: /* wide-gamut RGB */
        return pure_power(v, 2.222f);
    case 4: /* ProPhoto RGB */
        return pure_power(v, 1.8f);
    case 8: /* Rec.2020 */
        return rec2020_encode(v);
    default:
        return power_with_slope(v, gamma_power, gamma_slope);
    }
}

int oracle_gamma_correct(const float *in3, float *out3, int w, int h, float gamma_power,
                         float gamma_slope, int output_color_space) {
    const size_t n = (size_t)w * h * 3;
    for (size_t i = 0; i < n; i++)
        out3[i] = oracle_gamma_value(in3[i], gamma_power, gamma_slope, output_color_space);
    return 0;
}

int oracle_quantize(const float *in, void *out, size_t n, int bits) {
    if (bits == 8) {
        uint8_t *o = (uint8_t *)out;
        for (size_t i = 0; i < n; i++) {
            const float v = in[i] < 0.f ? 0.f : (1.f < in[i] ? 1.f : in[i]);
            o[i] = (uint8_t)(v * 255.0f + 0.5f);
        }
        return 0;
    }
    if (bits == 16) {
        uint16_t *o = (uint16_t *)out;
        for (size_t i = 0; i < n; i++) {
            const float v = in[i] < 0.f ? 0.f : (1.f < in[i] ? 1.f : in[i]);
            o[i] = (uint16_t)(v * 65535.0f + 0.5f);
        }
        return 0;
    }
    return -1;
}
