/*
 * TEST INFRASTRUCTURE ONLY — CPU restatement of the two chromatic-aberration
 * stages of the reference.  Never called by the product.
 *
 *  oracle_axial_ca    core/cpu_accelerator.cpp:3799-3928, with the FP64
 *                     summed-area tables and box means of :3689-3716 and the
 *                     shared-guide guided filter of :3725-3777
 *  oracle_lateral_ca  core/cpu_accelerator.cpp:3496-3673, with the 2x2 box
 *                     pyramid (:3369-3384), per-cell Lucas-Kanade (:3390-3471),
 *                     3x3 cell smoothing (:3474-3492) and the bilinear sampler
 *                     (:3328-3346)
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static inline float maxf_(float a, float b) { return (a < b) ? b : a; }
static inline float clampf_(float v, float lo, float hi) { return v < lo ? lo : (hi < v ? hi : v); }
static inline int maxi_(int a, int b) { return a < b ? b : a; }
static inline int mini_(int a, int b) { return b < a ? b : a; }

/* ---------------- axial CA ---------------- */

static double *build_sat(const float *plane, int W, int H) { /* :3689-3701 */
    const size_t SW = (size_t)W + 1, SH = (size_t)H + 1;
    double *sat = calloc(SW * SH, sizeof(double));
    for (int y = 0; y < H; y++) {
        double row = 0.0;
        for (int x = 0; x < W; x++) {
            row += (double)plane[(size_t)y * W + x];
            sat[(size_t)(y + 1) * SW + x + 1] = sat[(size_t)y * SW + x + 1] + row;
        }
    }
    return sat;
}

static inline double box_mean(const double *sat, int W, int H, int x, int y, int r) { /* :3703-3716 */
    const int xl = maxi_(0, x - r), yl = maxi_(0, y - r);
    const int xr = mini_(W - 1, x + r), yb = mini_(H - 1, y + r);
    const size_t SW = (size_t)W + 1;
    const double s = sat[(size_t)(yb + 1) * SW + xr + 1] - sat[(size_t)yl * SW + xr + 1] -
                     sat[(size_t)(yb + 1) * SW + xl] + sat[(size_t)yl * SW + xl];
    const int area = (xr - xl + 1) * (yb - yl + 1);
    return area > 0 ? s / (double)area : 0.0;
}

static void guided(const float *I, const float *p, float *out, int W, int H, int r, float eps,
                   const double *satI, const double *satII) { /* :3725-3777 */
    const size_t N = (size_t)W * H;
    float *Ip = malloc(N * sizeof(float)), *ap = malloc(N * sizeof(float)), *bp = malloc(N * sizeof(float));
    for (size_t i = 0; i < N; i++) Ip[i] = I[i] * p[i];
    double *satp = build_sat(p, W, H), *satIp = build_sat(Ip, W, H);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const double mI = box_mean(satI, W, H, x, y, r);
            const double mp = box_mean(satp, W, H, x, y, r);
            const double mIp = box_mean(satIp, W, H, x, y, r);
            const double mII = box_mean(satII, W, H, x, y, r);
            const double var = mII - mI * mI;
            const double cov = mIp - mI * mp;
            const double a = cov / (var + (double)eps);
            const double b = mp - a * mI;
            ap[(size_t)y * W + x] = (float)a;
            bp[(size_t)y * W + x] = (float)b;
        }
    double *sata = build_sat(ap, W, H), *satb = build_sat(bp, W, H);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const double ma = box_mean(sata, W, H, x, y, r);
            const double mb = box_mean(satb, W, H, x, y, r);
            const size_t i = (size_t)y * W + x;
            out[i] = (float)(ma * I[i] + mb);
        }
    free(Ip); free(ap); free(bp); free(satp); free(satIp); free(sata); free(satb);
}

int oracle_axial_ca(const float *src, float *dst, int W, int H, int radius, float epsilon,
                    float strength) {
    if (!src || !dst || W <= 0 || H <= 0) return -1;
    if (radius < 1) radius = 1;
    strength = clampf_(strength, 0.f, 1.f);
    const size_t N = (size_t)W * H;
    float max_g = 1e-6f;
    for (size_t i = 0; i < N; i++) {
        const float g = src[3 * i + 1];
        if (g > max_g) max_g = g;
    }
    const float norm_ref = maxf_(max_g, 1e-6f);
    const float inv_norm = 1.f / norm_ref;
    float *R = malloc(N * sizeof(float)), *G = malloc(N * sizeof(float)), *B = malloc(N * sizeof(float)), *II = malloc(N * sizeof(float));
    float *Rf = malloc(N * sizeof(float)), *Bf = malloc(N * sizeof(float)), *eraw = calloc(N, sizeof(float)), *ew = malloc(N * sizeof(float));
    for (size_t i = 0; i < N; i++) {
        R[i] = src[3 * i] * inv_norm;
        G[i] = src[3 * i + 1] * inv_norm;
        B[i] = src[3 * i + 2] * inv_norm;
    }
    for (size_t i = 0; i < N; i++) II[i] = G[i] * G[i];
    double *satI = build_sat(G, W, H), *satII = build_sat(II, W, H);
    guided(G, R, Rf, W, H, radius, epsilon, satI, satII);
    guided(G, B, Bf, W, H, radius, epsilon, satI, satII);
    for (int y = 1; y < H - 1; y++)
        for (int x = 1; x < W - 1; x++) {
            const size_t i = (size_t)y * W + x;
            const float gx = G[i + 1] - G[i - 1];
            const float gy = G[i + W] - G[i - W];
            eraw[i] = sqrtf(gx * gx + gy * gy);
        }
    double *sate = build_sat(eraw, W, H);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const double m = box_mean(sate, W, H, x, y, radius);
            const float t = clampf_(((float)m - 0.01f) / 0.05f, 0.f, 1.f);
            ew[(size_t)y * W + x] = t * t * (3.f - 2.f * t);
        }
    for (size_t i = 0; i < N; i++) {
        const float w = strength * ew[i];
        const float r_out = R[i] * (1.f - w) + Rf[i] * w;
        const float b_out = B[i] * (1.f - w) + Bf[i] * w;
        const float g_in = src[3 * i + 1];
        dst[3 * i] = r_out * norm_ref;
        dst[3 * i + 1] = g_in;
        dst[3 * i + 2] = b_out * norm_ref;
    }
    free(R); free(G); free(B); free(II); free(Rf); free(Bf); free(eraw); free(ew);
    free(satI); free(satII); free(sate);
    return 0;
}

/* ---------------- lateral CA ---------------- */

static float bilinear(const float *plane, int W, int H, float x, float y) { /* :3328-3346 */
    if (x < 0.f) x = 0.f;
    if (y < 0.f) y = 0.f;
    if (x > (float)(W - 1)) x = (float)(W - 1);
    if (y > (float)(H - 1)) y = (float)(H - 1);
    const size_t x0 = (size_t)x, y0 = (size_t)y;
    const size_t x1 = x0 + 1 < (size_t)W - 1 ? x0 + 1 : (size_t)W - 1;
    const size_t y1 = y0 + 1 < (size_t)H - 1 ? y0 + 1 : (size_t)H - 1;
    const float fx = x - (float)x0, fy = y - (float)y0;
    const float v00 = plane[y0 * W + x0], v01 = plane[y0 * W + x1];
    const float v10 = plane[y1 * W + x0], v11 = plane[y1 * W + x1];
    return (v00 * (1.f - fx) + v01 * fx) * (1.f - fy) + (v10 * (1.f - fx) + v11 * fx) * fy;
}

static float *downsample(const float *in, int W, int H, int *W2, int *H2) { /* :3369-3384 */
    *W2 = W / 2;
    *H2 = H / 2;
    float *out = calloc((size_t)*W2 * *H2 + 1, sizeof(float));
    for (int y = 0; y < *H2; y++) {
        const int y0 = 2 * y, y1 = mini_(y0 + 1, H - 1);
        for (int x = 0; x < *W2; x++) {
            const int x0 = 2 * x, x1 = mini_(x0 + 1, W - 1);
            out[(size_t)y * *W2 + x] = 0.25f * (in[(size_t)y0 * W + x0] + in[(size_t)y0 * W + x1] +
                                                 in[(size_t)y1 * W + x0] + in[(size_t)y1 * W + x1]);
        }
    }
    return out;
}

static void lk_level(const float *G, const float *T, int W, int H, int mw, int mh, int iters,
                     float max_shift, float *dxm, float *dym, float *conf) { /* :3390-3471 */
    const float cw = (float)W / (float)mw, ch = (float)H / (float)mh;
#pragma omp parallel for schedule(dynamic)
    for (int c = 0; c < mw * mh; c++) {
        const int cy = c / mw, cx = c % mw;
        const size_t x0 = (size_t)(cx * cw), y0 = (size_t)(cy * ch);
        size_t x1 = (size_t)((cx + 1) * cw), y1 = (size_t)((cy + 1) * ch);
        if (x1 > (size_t)W) x1 = W;
        if (y1 > (size_t)H) y1 = H;
        if (x1 <= x0 + 4 || y1 <= y0 + 4) continue;
        float adx = dxm[c], ady = dym[c], min_eig = 0.f;
        for (int it = 0; it < iters; it++) {
            double A00 = 0, A01 = 0, A11 = 0, bx = 0, by = 0;
            size_t valid = 0;
            for (size_t y = y0 + 1; y + 1 < y1; y++)
                for (size_t x = x0 + 1; x + 1 < x1; x++) {
                    const float gx = G[y * W + x + 1] - G[y * W + x - 1];
                    const float gy = G[(y + 1) * W + x] - G[(y - 1) * W + x];
                    const float tw = bilinear(T, W, H, (float)x + adx, (float)y + ady);
                    const float err = tw - G[y * W + x];
                    A00 += (double)gx * gx;
                    A01 += (double)gx * gy;
                    A11 += (double)gy * gy;
                    bx -= (double)gx * err;
                    by -= (double)gy * err;
                    valid++;
                }
            if (valid < 32) break;
            const double det = A00 * A11 - A01 * A01;
            const double tr = A00 + A11;
            const double disc = (tr * tr - 4.0 * det) > 0.0 ? (tr * tr - 4.0 * det) : 0.0; /* std::max(0.0, .) */
            min_eig = (float)(0.5 * (tr - sqrt(disc)));
            if (det <= 1e-12) break;
            const double inv = 1.0 / det;
            const double sx = inv * (A11 * bx - A01 * by);
            const double sy = inv * (-A01 * bx + A00 * by);
            adx += (float)sx;
            ady += (float)sy;
            if (adx > max_shift) adx = max_shift;
            if (adx < -max_shift) adx = -max_shift;
            if (ady > max_shift) ady = max_shift;
            if (ady < -max_shift) ady = -max_shift;
            if (fabs(sx) < 0.005 && fabs(sy) < 0.005) break;
        }
        dxm[c] = adx;
        dym[c] = ady;
        conf[c] = maxf_(0.f, min_eig);
    }
}

static void smooth3(float *m, int mw, int mh) { /* :3474-3492 */
    float *o = calloc((size_t)mw * mh, sizeof(float));
    for (int cy = 0; cy < mh; cy++)
        for (int cx = 0; cx < mw; cx++) {
            float acc = 0.f;
            int n = 0;
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    const int yy = cy + dy, xx = cx + dx;
                    if (yy < 0 || yy >= mh || xx < 0 || xx >= mw) continue;
                    acc += m[yy * mw + xx];
                    n++;
                }
            o[cy * mw + cx] = n > 0 ? acc / n : 0.f;
        }
    memcpy(m, o, (size_t)mw * mh * sizeof(float));
    free(o);
}

typedef struct {
    int mw, mh, cell;
    float clamp_shift;
} shiftmap_t;

static void sample_shift(const shiftmap_t *S, const float *dxm, const float *dym, float px,
                         float py, float *odx, float *ody) { /* :3623-3649 */
    const float gxf = (px / (float)S->cell) - 0.5f, gyf = (py / (float)S->cell) - 0.5f;
    int gx0 = (int)floorf(gxf), gy0 = (int)floorf(gyf);
    if (gx0 < 0) gx0 = 0;
    if (gy0 < 0) gy0 = 0;
    if (gx0 >= S->mw - 1) gx0 = S->mw - 2;
    if (gy0 >= S->mh - 1) gy0 = S->mh - 2;
    if (gx0 < 0) gx0 = 0;
    if (gy0 < 0) gy0 = 0;
    const float fx = clampf_(gxf - gx0, 0.f, 1.f), fy = clampf_(gyf - gy0, 0.f, 1.f);
    const int gx1 = mini_(gx0 + 1, S->mw - 1), gy1 = mini_(gy0 + 1, S->mh - 1);
#define M(m, xi, yi) (m)[(size_t)(yi) * S->mw + (xi)]
    const float dx = (M(dxm, gx0, gy0) * (1.f - fx) + M(dxm, gx1, gy0) * fx) * (1.f - fy) +
                     (M(dxm, gx0, gy1) * (1.f - fx) + M(dxm, gx1, gy1) * fx) * fy;
    const float dy = (M(dym, gx0, gy0) * (1.f - fx) + M(dym, gx1, gy0) * fx) * (1.f - fy) +
                     (M(dym, gx0, gy1) * (1.f - fx) + M(dym, gx1, gy1) * fx) * fy;
#undef M
    *odx = clampf_(dx, -S->clamp_shift, S->clamp_shift);
    *ody = clampf_(dy, -S->clamp_shift, S->clamp_shift);
}

int oracle_lateral_ca(const float *src, float *dst, int W0, int H0, int cell_size, int max_iters,
                      float max_shift, float min_conf, int levels) {
    if (!src || !dst || W0 <= 0 || H0 <= 0) return -1;
    if (cell_size < 16) cell_size = 16;
    if (max_iters < 1) max_iters = 1;
    if (levels < 1) levels = 1;
    const size_t N0 = (size_t)W0 * H0;
    int L = levels;
    while (L > 1 && ((W0 >> (L - 1)) < 32 || (H0 >> (L - 1)) < 32)) L--;
    float *Rl[16], *Gl[16], *Bl[16];
    int Wl[16], Hl[16];
    Rl[0] = malloc(N0 * sizeof(float)); Gl[0] = malloc(N0 * sizeof(float)); Bl[0] = malloc(N0 * sizeof(float));
    for (size_t i = 0; i < N0; i++) {
        Rl[0][i] = src[3 * i];
        Gl[0][i] = src[3 * i + 1];
        Bl[0][i] = src[3 * i + 2];
    }
    Wl[0] = W0;
    Hl[0] = H0;
    for (int l = 1; l < L; l++) {
        Rl[l] = downsample(Rl[l - 1], Wl[l - 1], Hl[l - 1], &Wl[l], &Hl[l]);
        Gl[l] = downsample(Gl[l - 1], Wl[l - 1], Hl[l - 1], &Wl[l], &Hl[l]);
        Bl[l] = downsample(Bl[l - 1], Wl[l - 1], Hl[l - 1], &Wl[l], &Hl[l]);
    }
    const int mw = maxi_(2, (W0 + cell_size - 1) / cell_size);
    const int mh = maxi_(2, (H0 + cell_size - 1) / cell_size);
    const int nc = mw * mh;
    float *dxr = calloc(nc, sizeof(float)), *dyr = calloc(nc, sizeof(float)), *cr = calloc(nc, sizeof(float));
    float *dxb = calloc(nc, sizeof(float)), *dyb = calloc(nc, sizeof(float)), *cb = calloc(nc, sizeof(float));
    for (int l = L - 1; l >= 0; l--) {
        const float down = 1.f / (float)(1 << l);
        for (int i = 0; i < nc; i++) { dxr[i] *= down; dyr[i] *= down; dxb[i] *= down; dyb[i] *= down; }
        const float lms = max_shift * down + (float)(1 << l) * 0.5f;
        lk_level(Gl[l], Rl[l], Wl[l], Hl[l], mw, mh, max_iters, lms, dxr, dyr, cr);
        lk_level(Gl[l], Bl[l], Wl[l], Hl[l], mw, mh, max_iters, lms, dxb, dyb, cb);
        const float up = (float)(1 << l);
        for (int i = 0; i < nc; i++) { dxr[i] *= up; dyr[i] *= up; dxb[i] *= up; dyb[i] *= up; }
    }
    float *maps[2][3] = {{dxr, dyr, cr}, {dxb, dyb, cb}};
    for (int k = 0; k < 2; k++) { /* :3584-3596 */
        float maxc = 0.f;
        for (int i = 0; i < nc; i++) maxc = maxf_(maxc, maps[k][2][i]);
        if (maxc <= 0.f) continue;
        const float thr = maxc * min_conf;
        for (int i = 0; i < nc; i++)
            if (maps[k][2][i] < thr) { maps[k][0][i] = 0.f; maps[k][1][i] = 0.f; }
    }
    smooth3(dxr, mw, mh);
    smooth3(dyr, mw, mh);
    smooth3(dxb, mw, mh);
    smooth3(dyb, mw, mh);
    shiftmap_t S = {mw, mh, cell_size, max_shift * (float)(1 << (L - 1))};
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H0; y++)
        for (int x = 0; x < W0; x++) {
            const size_t i = (size_t)y * W0 + x;
            float drx, dry, dbx, dby;
            sample_shift(&S, dxr, dyr, (float)x, (float)y, &drx, &dry);
            sample_shift(&S, dxb, dyb, (float)x, (float)y, &dbx, &dby);
            const float rn = bilinear(Rl[0], W0, H0, (float)x + drx, (float)y + dry);
            const float bn = bilinear(Bl[0], W0, H0, (float)x + dbx, (float)y + dby);
            dst[3 * i] = rn;
            dst[3 * i + 1] = Gl[0][i];
            dst[3 * i + 2] = bn;
        }
    for (int l = 0; l < L; l++) { free(Rl[l]); free(Gl[l]); free(Bl[l]); }
    free(dxr); free(dyr); free(cr); free(dxb); free(dyb); free(cb);
    return 0;
}
