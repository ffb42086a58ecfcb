/*
 * TEST INFRASTRUCTURE ONLY — CPU restatement of the pointwise stages and the
 * detail-preserving tone map.  Never called by the product.
 *
 *  oracle_tone_mapping    core/cpu_accelerator.cpp:3034-3068 (ACES x after_scale)
 *  oracle_detail_tonemap  core/libraw_wrapper.cpp:1389-1471, with the sliding
 *                         box filter of core/libraw_wrapper.cpp:45-116
 *  oracle_white_balance   core/cpu_accelerator.cpp:2789-2869
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static inline float aces(float x) { /* cpu_accelerator.cpp:3047-3055, libraw_wrapper.cpp:1389-1396 */
    const float a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f;
    return (x * (a * x + b)) / (x * (c * x + d) + e);
}

int oracle_tone_mapping(const float *in, float *out, int w, int h, float after_scale) {
    if (!in || !out || w <= 0 || h <= 0) return -1;
    size_t n = (size_t)w * h * 3;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) out[i] = aces(in[i]) * after_scale;
    return 0;
}

static inline float maxf_std(float a, float b) { return (a < b) ? b : a; } /* std::max */
static inline float clampf_std(float v, float lo, float hi) {               /* std::clamp */
    return v < lo ? lo : (hi < v ? hi : v);
}

/* Running-sum box mean with a window clipped at the frame (true-area divisor):
 * rows first into tmp, then columns into dst.  libraw_wrapper.cpp:45-116. */
static void box_mean_sliding(const float *src, float *dst, float *tmp, int W, int H, int r) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++) {
        const float *s = src + (size_t)y * W;
        float *t = tmp + (size_t)y * W;
        float acc = 0.0f;
        int first = W - 1 < r ? W - 1 : r;
        for (int x = 0; x <= first; x++) acc += s[x];
        for (int x = 0; x < W; x++) {
            int lo = x > r ? x - r : 0;
            int hi = x + r < W - 1 ? x + r : W - 1;
            t[x] = acc / (float)(hi - lo + 1);
            int rem = x >= r ? x - r : W;
            int add = x + r + 1;
            if (rem < W) acc -= s[rem];
            if (add < W) acc += s[add];
        }
    }
#pragma omp parallel for schedule(static)
    for (int x = 0; x < W; x++) {
        float acc = 0.0f;
        int first = H - 1 < r ? H - 1 : r;
        for (int y = 0; y <= first; y++) acc += tmp[(size_t)y * W + x];
        for (int y = 0; y < H; y++) {
            int lo = y > r ? y - r : 0;
            int hi = y + r < H - 1 ? y + r : H - 1;
            dst[(size_t)y * W + x] = acc / (float)(hi - lo + 1);
            int rem = y >= r ? y - r : H;
            int add = y + r + 1;
            if (rem < H) acc -= tmp[(size_t)rem * W + x];
            if (add < H) acc += tmp[(size_t)add * W + x];
        }
    }
}

int oracle_detail_tonemap(float *rgb, int w, int h) {
    if (!rgb || w <= 0 || h <= 0) return -1;
    const size_t n = (size_t)w * h;
    const int radius = 12;
    const float eps = 0.0025f, edge0 = 0.55f, edge1 = 0.90f;
    float *guide = malloc(n * sizeof(float)), *mean = malloc(n * sizeof(float));
    float *work = malloc(n * sizeof(float)), *b = malloc(n * sizeof(float));
    float *tmp = malloc(n * sizeof(float));
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) {
        const float *p = rgb + 3 * i;
        guide[i] = maxf_std(0.0f, maxf_std(p[0], maxf_std(p[1], p[2])));
        work[i] = guide[i] * guide[i];
    }
    box_mean_sliding(guide, mean, tmp, w, h, radius);
    box_mean_sliding(work, work, tmp, w, h, radius);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) {
        float var = maxf_std(0.0f, work[i] - mean[i] * mean[i]);
        float a = var / (var + eps);
        work[i] = a;
        b[i] = mean[i] * (1.0f - a);
    }
    box_mean_sliding(work, work, tmp, w, h, radius);
    box_mean_sliding(b, b, tmp, w, h, radius);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) {
        float base = maxf_std(1e-6f, work[i] * guide[i] + b[i]);
        float gain = clampf_std(aces(base) / base, 0.0f, 1.0f);
        float t = clampf_std((base - edge0) / (edge1 - edge0), 0.0f, 1.0f);
        t = t * t * (3.0f - 2.0f * t);
        if (t <= 0.0f) continue;
        gain = 1.0f + (gain - 1.0f) * t;
        rgb[3 * i + 0] *= gain;
        rgb[3 * i + 1] *= gain;
        rgb[3 * i + 2] *= gain;
    }
    free(guide); free(mean); free(work); free(b); free(tmp);
    return 0;
}

int oracle_white_balance(const uint16_t *img4, float *rgb3, int w, int h, const float *wb4,
                         uint32_t filters, const signed char *xtrans36) {
    if (!img4 || !rgb3 || w <= 0 || h <= 0) return -1;
#pragma omp parallel for schedule(static)
    for (int row = 0; row < h; row++)
        for (int col = 0; col < w; col++) {
            size_t i = (size_t)row * w + col;
            unsigned c;
            if (filters == 9) /* FILTERS_XTRANS, accelerator.h:179 */
                c = (unsigned)xtrans36[((row + 6) % 6) * 6 + (col + 6) % 6];
            else
                c = filters >> ((((row << 1) & 14) | (col & 1)) << 1) & 3;
            float v = (float)img4[4 * i + c] * wb4[c];
            if (c == 3) c = 1;
            rgb3[3 * i + c] = v;
        }
    return 0;
}
