/*
 * Seeded synthetic scene and CFA mosaic generator (host C).
 *
 * The develop path has no camera files to run on, so every workload is a
 * deterministic synthetic scene sampled through a colour filter array:
 *   - a smooth per-channel colour gradient  0.15 + 0.6 (a_c x/W + b_c y/H)
 *   - 200 hard-edged polygons (3..8 vertices, random orientation, radius
 *     U(0.01,0.15) * min(H,W), uniform RGB fill U(0.05,0.95)), painted in order
 *   - a centred sinusoidal zone plate 0.5 + 0.5 cos(pi f r^2 / R) whose local
 *     frequency reaches f = 0.45 cycles/px at the rim r = R, blended 50 %
 *   - clip to [0,1]; optional radial lateral-CA pre-warp of R and B
 *   - CFA sampling (Bayer via the LibRaw `filters` word, or a 6x6 X-Trans table)
 *   - Poisson-Gaussian noise y = Poisson(4000 x)/4000 + N(0, 0.002^2)
 *
 * All randomness comes from this file's own generators (SplitMix64 seeding a
 * PCG32 stream), so the bytes are identical on every machine that runs this
 * code.  Per-row streams make the OpenMP loops order-independent.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define LRE_PI 3.14159265358979323846

typedef struct { uint64_t state, inc; } pcg32_t;

static uint64_t splitmix64(uint64_t *x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static uint32_t pcg32_next(pcg32_t *g) {
    uint64_t old = g->state;
    g->state = old * 6364136223846793005ull + g->inc;
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((-rot) & 31));
}

static void pcg32_seed(pcg32_t *g, uint64_t seed, uint64_t stream) {
    uint64_t sm = seed ^ (stream * 0xD1B54A32D192ED03ull);
    g->state = 0;
    g->inc = (splitmix64(&sm) << 1u) | 1u;
    pcg32_next(g);
    g->state += splitmix64(&sm);
    pcg32_next(g);
}

/* uniform double in [0,1) with 53 random bits */
static double u01(pcg32_t *g) {
    uint64_t a = pcg32_next(g) >> 5, b = pcg32_next(g) >> 6;
    return ((double)a * 67108864.0 + (double)b) * (1.0 / 9007199254740992.0);
}

static double std_normal(pcg32_t *g) {
    /* Box-Muller, one value per call (the partner is discarded for simplicity) */
    double u1 = u01(g), u2 = u01(g);
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(2.0 * LRE_PI * u2);
}

/* Poisson sampler: Knuth's product method for small means, Hormann's PTRS
 * transformed rejection (W. Hormann, "The transformed rejection method for
 * generating Poisson random variables", 1993) for mean >= 10. */
static double log_factorial(double k) { int sg; return lgamma_r(k + 1.0, &sg); }

static uint32_t poisson(pcg32_t *g, double lam) {
    if (lam <= 0.0) return 0;
    if (lam < 10.0) {
        double L = exp(-lam), p = 1.0;
        uint32_t k = 0;
        do { k++; p *= u01(g); } while (p > L);
        return k - 1;
    }
    double slam = sqrt(lam), loglam = log(lam);
    double b = 0.931 + 2.53 * slam;
    double a = -0.059 + 0.02483 * b;
    double invalpha = 1.1239 + 1.1328 / (b - 3.4);
    double vr = 0.9277 - 3.6224 / (b - 2.0);
    for (;;) {
        double U = u01(g) - 0.5, V = u01(g);
        double us = 0.5 - fabs(U);
        double k = floor((2.0 * a / us + b) * U + lam + 0.43);
        if (us >= 0.07 && V <= vr) return (uint32_t)k;
        if (k < 0.0 || (us < 0.013 && V > us)) continue;
        if (log(V) + log(invalpha) - log(a / (us * us) + b) <=
            -lam + k * loglam - log_factorial(k))
            return (uint32_t)k;
    }
}

typedef struct {
    int n;
    double vx[8], vy[8];
    double ymin, ymax;
    float rgb[3];
} polygon_t;

#define NPOLY 200

/*
 * Render the clean scene (H x W x 3 interleaved float32, [0,1]).
 */
void lre_scene_render(float *rgb, int W, int H, uint64_t seed) {
    pcg32_t g;
    pcg32_seed(&g, seed, 0);
    double ga[3], gb[3];
    for (int c = 0; c < 3; c++) { ga[c] = u01(&g); gb[c] = u01(&g); }
    polygon_t *poly = (polygon_t *)malloc(sizeof(polygon_t) * NPOLY);
    double mind = (double)(H < W ? H : W);
    for (int p = 0; p < NPOLY; p++) {
        polygon_t *P = &poly[p];
        P->n = 3 + (int)(pcg32_next(&g) % 6u);
        double cx = u01(&g) * W, cy = u01(&g) * H;
        double rad = (0.01 + 0.14 * u01(&g)) * mind;
        double th0 = 2.0 * LRE_PI * u01(&g);
        P->ymin = 1e300; P->ymax = -1e300;
        for (int k = 0; k < P->n; k++) {
            double jitter = (u01(&g) - 0.5) * 0.6 * (LRE_PI / P->n);
            double th = th0 + 2.0 * LRE_PI * k / P->n + jitter;
            double rr = rad * (0.6 + 0.4 * u01(&g));
            P->vx[k] = cx + rr * cos(th);
            P->vy[k] = cy + rr * sin(th);
            if (P->vy[k] < P->ymin) P->ymin = P->vy[k];
            if (P->vy[k] > P->ymax) P->ymax = P->vy[k];
        }
        for (int c = 0; c < 3; c++) P->rgb[c] = (float)(0.05 + 0.9 * u01(&g));
    }
    const double R = 0.45 * mind;            /* zone plate disk radius */
    const double fmax = 0.45;                /* cycles/px at r = R */
    const double zx = 0.5 * W, zy = 0.5 * H;

#pragma omp parallel for schedule(dynamic, 16)
    for (int y = 0; y < H; y++) {
        double py = y + 0.5;
        float *row = rgb + (size_t)y * W * 3;
        for (int x = 0; x < W; x++)
            for (int c = 0; c < 3; c++)
                row[x * 3 + c] = (float)(0.15 + 0.6 * (ga[c] * x / W + gb[c] * y / H));
        /* scanline fill of each polygon, in paint order, even-odd rule */
        for (int p = 0; p < NPOLY; p++) {
            const polygon_t *P = &poly[p];
            if (py < P->ymin || py > P->ymax) continue;
            double xs[8];
            int nx = 0;
            for (int k = 0; k < P->n; k++) {
                int k2 = (k + 1) % P->n;
                double y0 = P->vy[k], y1 = P->vy[k2];
                if ((y0 <= py && y1 > py) || (y1 <= py && y0 > py)) {
                    double t = (py - y0) / (y1 - y0);
                    xs[nx++] = P->vx[k] + t * (P->vx[k2] - P->vx[k]);
                }
            }
            for (int i = 1; i < nx; i++)   /* insertion sort */
                for (int j = i; j > 0 && xs[j] < xs[j - 1]; j--) {
                    double t = xs[j]; xs[j] = xs[j - 1]; xs[j - 1] = t;
                }
            for (int i = 0; i + 1 < nx; i += 2) {
                int x0 = (int)ceil(xs[i] - 0.5), x1 = (int)ceil(xs[i + 1] - 0.5);
                if (x0 < 0) x0 = 0;
                if (x1 > W) x1 = W;
                for (int x = x0; x < x1; x++)
                    for (int c = 0; c < 3; c++) row[x * 3 + c] = P->rgb[c];
            }
        }
        double dy = py - zy;
        for (int x = 0; x < W; x++) {
            double dx = x + 0.5 - zx;
            double r2 = dx * dx + dy * dy;
            if (r2 < R * R) {
                float zp = (float)(0.5 + 0.5 * cos(LRE_PI * fmax * r2 / R));
                for (int c = 0; c < 3; c++) row[x * 3 + c] = 0.5f * row[x * 3 + c] + 0.5f * zp;
            }
            for (int c = 0; c < 3; c++) {
                float v = row[x * 3 + c];
                row[x * 3 + c] = v < 0.f ? 0.f : (v > 1.f ? 1.f : v);
            }
        }
    }
    free(poly);
}

/* Bilinear sample of one channel of an interleaved RGB image, clamped. */
static float sample_bilinear(const float *rgb, int W, int H, int c, double x, double y) {
    if (x < 0) x = 0;
    if (y < 0) y = 0;
    if (x > W - 1) x = W - 1;
    if (y > H - 1) y = H - 1;
    int x0 = (int)x, y0 = (int)y;
    int x1 = x0 + 1 < W ? x0 + 1 : W - 1, y1 = y0 + 1 < H ? y0 + 1 : H - 1;
    double fx = x - x0, fy = y - y0;
    double v00 = rgb[((size_t)y0 * W + x0) * 3 + c], v01 = rgb[((size_t)y0 * W + x1) * 3 + c];
    double v10 = rgb[((size_t)y1 * W + x0) * 3 + c], v11 = rgb[((size_t)y1 * W + x1) * 3 + c];
    return (float)((v00 * (1 - fx) + v01 * fx) * (1 - fy) + (v10 * (1 - fx) + v11 * fx) * fy);
}

/*
 * Radial lateral-CA pre-warp: channel c of the output at p samples the input
 * at centre + (p - centre) * scale_c (R uses scale_r, B uses scale_b, G is
 * copied).  Writes a new interleaved image.
 */
void lre_scene_lateral_warp(const float *src, float *dst, int W, int H,
                            double scale_r, double scale_b) {
    const double cx = 0.5 * (W - 1), cy = 0.5 * (H - 1);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            size_t i = ((size_t)y * W + x) * 3;
            dst[i + 0] = sample_bilinear(src, W, H, 0, cx + (x - cx) * scale_r, cy + (y - cy) * scale_r);
            dst[i + 1] = src[i + 1];
            dst[i + 2] = sample_bilinear(src, W, H, 2, cx + (x - cx) * scale_b, cy + (y - cy) * scale_b);
        }
}

static inline int bayer_color3(uint32_t filters, int row, int col) {
    int c = (int)(filters >> ((((row << 1) & 14) | (col & 1)) << 1) & 3u);
    return c == 3 ? 1 : c;
}

/*
 * Sample the clean scene through the CFA and add Poisson-Gaussian noise.
 * cfa: H x W float plane (one sample per site).  xtrans: 36 entries or NULL
 * (then `filters` selects the Bayer pattern).  gain_inv = 4000 photons per
 * unit, read_sigma = 0.002 in the synthetic workloads; noise is skipped when
 * gain_inv <= 0.
 */
void lre_scene_mosaic(const float *rgb, float *cfa, int W, int H, uint32_t filters,
                      const signed char *xtrans, double gain_inv, double read_sigma,
                      uint64_t seed) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++) {
        pcg32_t g;
        pcg32_seed(&g, seed, 1000003ull + (uint64_t)y);
        for (int x = 0; x < W; x++) {
            int c = xtrans ? xtrans[(y % 6) * 6 + (x % 6)] : bayer_color3(filters, y, x);
            double v = rgb[((size_t)y * W + x) * 3 + c];
            if (gain_inv > 0) {
                v = poisson(&g, v * gain_inv) / gain_inv;
                v += read_sigma * std_normal(&g);
            }
            cfa[(size_t)y * W + x] = (float)v;
        }
    }
}

/*
 * Quantise a noisy CFA plane to 14-bit raw counts:
 *   q = clamp(round(y / wb_c * (white - black)) + black, 0, white)
 * wb has 3 entries (R, G, B); the result is a H x W uint16 plane.
 */
void lre_scene_quantize_u16(const float *cfa, uint16_t *out, int W, int H, uint32_t filters,
                            const signed char *xtrans, const float *wb, int black, int white) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            int c = xtrans ? xtrans[(y % 6) * 6 + (x % 6)] : bayer_color3(filters, y, x);
            double q = nearbyint((double)cfa[(size_t)y * W + x] / wb[c] * (white - black)) + black;
            if (q < 0) q = 0;
            if (q > white) q = white;
            out[(size_t)y * W + x] = (uint16_t)q;
        }
}

/* Expand a CFA plane into the 3-channel demosaic input layout
 * (sample at its native channel, the other two channels 0). */
void lre_scene_cfa_to_rgb3(const float *cfa, float *rgb3, int W, int H, uint32_t filters,
                           const signed char *xtrans) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            int c = xtrans ? xtrans[(y % 6) * 6 + (x % 6)] : bayer_color3(filters, y, x);
            float *o = rgb3 + ((size_t)y * W + x) * 3;
            o[0] = o[1] = o[2] = 0.f;
            o[c] = cfa[(size_t)y * W + x];
        }
}
