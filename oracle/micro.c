/*
 * TEST INFRASTRUCTURE ONLY — CPU restatement of the micro-contrast
 * enhancement.  Never called by the product.
 *
 *  oracle_micro_contrast  core/cpu_accelerator.cpp:3135-3319
 *                         (enhance_micro_contrast), with its helpers
 *                         create_gaussian_kernel :3077-3099 and
 *                         apply_gaussian_blur :3107-3128
 *  oracle_mc_kernel       the normalised 9x9 Gaussian of :3165-3169
 *
 * The reference's own code for this block is written against Apple's
 * <simd/simd.h> float3 type and is not part of the reference build here.  Its
 * float3 operations are componentwise; a float3 divided by a float divides
 * every component.  The kernel weights take exp of a float argument in float
 * (the overload `exp(float)` resolves to on the reference's platform).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define MC_K 9 /* kernel_size (:3167) */

static inline float maxf_std(float a, float b) { return (a < b) ? b : a; } /* std::max */
static inline float clampf_std(float v, float lo, float hi) {               /* std::clamp */
    return v < lo ? lo : (hi < v ? hi : v);
}
static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (hi < v ? hi : v); }

/* create_gaussian_kernel(9, 5/3.7) (:3077-3099): weights summed in row order,
 * then each divided by the sum; k[ky * 9 + kx] */
void oracle_mc_kernel(float k[MC_K * MC_K]) {
    const float sigma = 5.f / 3.7f;
    const int center = MC_K / 2;
    float sum = 0.0f;
    for (int i = 0; i < MC_K; i++)
        for (int j = 0; j < MC_K; j++) {
            const float x = (float)(i - center), y = (float)(j - center);
            k[i * MC_K + j] = expf(-(x * x + y * y) / (2.0f * sigma * sigma));
            sum += k[i * MC_K + j];
        }
    for (int i = 0; i < MC_K * MC_K; i++) k[i] /= sum;
}

/* apply_gaussian_blur (:3107-3128): per pixel, taps in (ky, kx) raster order,
 * coordinates clamped at the frame; sq selects the blur of I * I */
static void blur3(const float *in3, float *out3, int w, int h, const float *k, int sq) {
    const int c0 = MC_K / 2;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            float acc[3] = {0.f, 0.f, 0.f};
            for (int ky = 0; ky < MC_K; ky++)
                for (int kx = 0; kx < MC_K; kx++) {
                    const int px = clampi(x + kx - c0, 0, w - 1), py = clampi(y + ky - c0, 0, h - 1);
                    const float wgt = k[ky * MC_K + kx];
                    const float *p = in3 + ((size_t)py * w + px) * 3;
                    for (int c = 0; c < 3; c++) {
                        const float v = sq ? p[c] * p[c] : p[c];
                        acc[c] += v * wgt;
                    }
                }
            float *o = out3 + ((size_t)y * w + x) * 3;
            for (int c = 0; c < 3; c++) o[c] = acc[c];
        }
}

int oracle_micro_contrast(const float *in3, float *out3, int w, int h, float threshold,
                          float strength, float target_contrast, const float *mask) {
    if (!in3 || !out3 || w <= 0 || h <= 0) return -1;
    const size_t n = (size_t)w * h;
    float k[MC_K * MC_K];
    oracle_mc_kernel(k);
    float *mean = malloc(n * 3 * sizeof(float)), *sd = malloc(n * 3 * sizeof(float));
    if (!mean || !sd) {
        free(mean);
        free(sd);
        return -2;
    }
    blur3(in3, mean, w, h, k, 0); /* local_mean (:3174) */
    blur3(in3, sd, w, h, k, 1);   /* blur of I * I (:3187) */
    /* local variance, then its non-negative square root (:3188-3205) */
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n * 3; i++) {
        const float var = sd[i] - mean[i] * mean[i];
        sd[i] = sqrtf(maxf_std(var, 0.f));
    }
    float max_std = 0.f; /* :3207-3210 */
    for (size_t i = 0; i < n; i++)
        max_std = maxf_std(max_std, maxf_std(maxf_std(sd[3 * i], sd[3 * i + 1]), sd[3 * i + 2]));

#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) {
        float e[3]; /* local_mean + enhanced high frequency (:3212-3261) */
        for (int c = 0; c < 3; c++) {
            const float cm = sd[3 * i + c] / max_std;
            const float f = (cm < target_contrast) ? strength * (target_contrast - cm) / target_contrast : 0.f;
            const float hf = in3[3 * i + c] - mean[3 * i + c];
            e[c] = mean[3 * i + c] + hf * (1.f + f);
        }
        const float *s = in3 + 3 * i;
        float *o = out3 + 3 * i;
        const int enhance = mask != NULL || s[1] >= threshold; /* :3263-3318 */
        if (enhance) {
            float r = e[0], g = e[1], b = e[2];
            const float maxc = maxf_std(r, maxf_std(g, b));
            if (maxc > 1.f) {
                const float scale = 1.f / maxc;
                r *= scale;
                g *= scale;
                b *= scale;
            }
            if (mask != NULL) {
                const float wt = clampf_std(mask[i], 0.f, 1.f);
                const float ir = s[0], ig = s[1], ib = s[2];
                o[0] = ir + (r - ir) * wt;
                o[1] = ig + (g - ig) * wt;
                o[2] = ib + (b - ib) * wt;
            } else {
                o[0] = r;
                o[1] = g;
                o[2] = b;
            }
        } else {
            o[0] = s[0];
            o[1] = s[1];
            o[2] = s[2];
        }
    }
    free(mean);
    free(sd);
    return 0;
}
