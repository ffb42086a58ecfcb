/*
 * TEST INFRASTRUCTURE ONLY — CPU restatement of the reference AMaZE demosaic
 * (core/cpu_accelerator.cpp:1145-1630, entry :1635-1653, post-pass
 * border_interpolate :2984-3028).  Used as the parity checker of the HIP
 * kernel; never called by the product.
 *
 * Bit-exactness with the reference needs three things reproduced literally:
 *  1. the same 160x160 tiles (step 128, 16-px reflect-101 border, corners of
 *     the border left unfilled: :1245-1326);
 *  2. the same per-tile scratch arena and its overlays (:1220-1240 and the
 *     re-used regions named at :1444,1445,1484,1485,1505-1507,1535,1571), since
 *     later stages read locations that earlier stages wrote under another
 *     name (for example the diagonal weights written over the horizontal
 *     direction weights that stage 9 still reads, :1507 vs :1554);
 *  3. the in-place, raster-order loops whose outputs feed later iterations of
 *     the same loop (the colour-difference variance test at :1377-1397, the
 *     hvwt smoothing at :1446-1448, the pmwt smoothing at :1537-1539).
 * Every arithmetic expression keeps the reference's operand order, and the
 * file is compiled with -ffp-contract=off.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

enum { TS = 160, TSH = 80 };
#define N2 (TS * TS)

static const float EPS = 1e-5f, EPSSQ = 1e-10f, ARTHRESH = 0.75f, NYQTHRESH = 0.5f;
/* Weight tables, :1155-1161 */
static const float GODD[4] = {0.14659727707323927f, 0.103592713382435f, 0.0732036125103057f,
                              0.0365543548389495f};
static const float GEVEN[2] = {0.13719494435797422f, 0.05640252782101291f};
static const float GQUINC[4] = {0.169917f, 0.108947f, 0.069855f, 0.0287182f};
static float ggrad(int k) {
    static const float base[6] = {0.07384411893421103f, 0.06207511968171489f, 0.0521818194747806f,
                                  0.03687419286733595f, 0.03099732204057846f, 0.018413194161458882f};
    return NYQTHRESH * base[k];
}

static inline float sq(float x) { return x * x; }
static inline float half(float x) { return x * 0.5f; }
static inline float twice(float x) { return x + x; }
static inline float fmin_std(float a, float b) { return (b < a) ? b : a; } /* std::min */
static inline float med3(float a, float b, float c) {                    /* shader_common.h:55 */
    float t;
    if (a > b) { t = a; a = b; b = t; }
    if (b > c) { t = b; b = c; c = t; }
    if (a > b) { t = a; a = b; b = t; }
    return b;
}
static inline unsigned bayer3(uint32_t filters, unsigned r, unsigned c) {
    unsigned n = filters >> ((((r << 1) & 14) | (c & 1)) << 1) & 3; /* shader_common.h:25-31 */
    return n == 3 ? 1 : n;
}

/* The per-tile arena: same sizes and offsets as the reference so that its
 * overlays line up (byte offsets relative to a 64-byte aligned base). */
typedef struct {
    char *base;
    float *green, *dsq, *wv, *wh, *vcd, *hcd, *vcdalt, *hcdalt, *cdsq, *hvwt, *dgv, *dgh, *q1m,
        *q1p, *cfa, *nyqtest;
    unsigned char *nyq, *nyq2;
    size_t bytes;
} arena_t;

static int arena_init(arena_t *a, void *mem) {
    a->bytes = 15 * sizeof(float) * N2 + TS * TSH + 256;
    a->base = (char *)(((uintptr_t)mem + 63) & ~(uintptr_t)63);
    float *f = (float *)a->base;
    a->green = f;
    a->dsq = a->green + N2;
    a->wv = a->dsq + N2;
    a->wh = a->wv + N2;
    a->vcd = a->wh + N2;
    a->hcd = a->vcd + N2;
    a->vcdalt = a->hcd + N2;
    a->hcdalt = a->vcdalt + N2;
    a->cdsq = a->hcdalt + N2;
    a->hvwt = a->cdsq + N2;
    a->dgv = a->hvwt + TS * TSH;
    a->dgh = a->dgv + N2;
    a->q1m = a->dgh + N2;
    a->q1p = a->q1m + TS * TSH;
    a->cfa = a->q1p + TS * TSH;
    a->nyq = (unsigned char *)(a->cfa + N2);
    a->nyq2 = (unsigned char *)a->cdsq;
    a->nyqtest = (float *)(a->nyq + TS * TSH);
    return 0;
}

typedef struct {
    const float *raw;
    float *out;
    int W, H;
    unsigned cf[2][2];
    int ex, ey;
    float clip_pt, clip_pt8, scale;
    int minimal_clear; /* clear only the cells a tile may read before writing */
} amaze_t;

static inline unsigned fct(const amaze_t *P, int r, int c) { return P->cf[r & 1][c & 1]; }

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

static inline float sample(const amaze_t *P, int row, int col) {
    return P->raw[((size_t)row * P->W + col) * 3 + fct(P, row, col)] * P->scale;
}

/* one tile, :1247-1625 */
static void amaze_tile(const amaze_t *P, arena_t *A, int top, int left) {
    const int v1 = TS, v2 = 2 * TS, v3 = 3 * TS;
    const int p1 = -TS + 1, p2 = -2 * TS + 2, p3 = -3 * TS + 3;
    const int m1 = TS + 1, m2 = 2 * TS + 2, m3 = 3 * TS + 3;
    const int W = P->W, H = P->H;
    float *cfa = A->cfa, *green = A->green;
    const int bottom = top + TS < H + 16 ? top + TS : H + 16;
    const int right = left + TS < W + 16 ? left + TS : W + 16;
    const int rr1 = bottom - top, cc1 = right - left;
    const int rrmin = top < 0 ? 16 : 0, ccmin = left < 0 ? 16 : 0;
    const int rrmax = bottom > H ? H - top : rr1;
    const int ccmax = right > W ? W - left : cc1;

    if (P->minimal_clear) {
        /* The cells of a tile whose value before this tile's writes reaches its
         * output: the CFA plane (the fill leaves the corners unwritten) and
         * the colour-difference planes outside [4, rr1-4) x [4, cc1-4) (the
         * stage-3 scans start from, and stage 4-5 reads, cells next to that
         * range; the Dgrb0/Dgrb1 overlays read the planes' first half).  The
         * Nyquist flags are cleared after stage 5 as in the reference.  The
         * unfilled green corners and the Dgrb2 pairs outside their written
         * range are read only for sites whose results never reach the output. */
        memset(A->cfa, 0, sizeof(float) * N2);
        float *cd[4] = {A->vcd, A->hcd, A->vcdalt, A->hcdalt};
        for (int k = 0; k < 4; k++)
            for (int rr = 0; rr < TS; rr++)
                for (int cc = 0; cc < TS; cc++)
                    if (rr < 4 || rr >= rr1 - 4 || cc < 4 || cc >= cc1 - 4) cd[k][rr * TS + cc] = 0.f;
    }

    /* tile fill with reflect-101 borders; corners stay as they were (:1326) */
    if (rrmin > 0)
        for (int rr = 0; rr < 16; rr++)
            for (int cc = ccmin; cc < ccmax; cc++) {
                float v = sample(P, clampi(32 - rr + top, 0, H - 1), clampi(cc + left, 0, W - 1));
                cfa[rr * TS + cc] = v;
                green[rr * TS + cc] = v;
            }
    for (int rr = rrmin; rr < rrmax; rr++)
        for (int cc = ccmin; cc < ccmax; cc++) {
            float v = sample(P, rr + top, cc + left);
            cfa[rr * TS + cc] = v;
            green[rr * TS + cc] = v;
        }
    if (rrmax < rr1)
        for (int rr = 0; rr < 16; rr++)
            for (int cc = ccmin; cc < ccmax; cc++) {
                float v = sample(P, clampi(H - rr - 2, 0, H - 1), clampi(left + cc, 0, W - 1));
                cfa[(rrmax + rr) * TS + cc] = v;
                green[(rrmax + rr) * TS + cc] = v;
            }
    if (ccmin > 0)
        for (int rr = rrmin; rr < rrmax; rr++)
            for (int cc = 0; cc < 16; cc++) {
                float v = sample(P, clampi(rr + top, 0, H - 1), clampi(32 - cc + left, 0, W - 1));
                cfa[rr * TS + cc] = v;
                green[rr * TS + cc] = v;
            }
    if (ccmax < cc1)
        for (int rr = rrmin; rr < rrmax; rr++)
            for (int cc = 0; cc < 16; cc++) {
                float v = sample(P, clampi(top + rr, 0, H - 1), clampi(W - cc - 2, 0, W - 1));
                cfa[rr * TS + ccmax + cc] = v;
                green[rr * TS + ccmax + cc] = v;
            }

    float *wv = A->wv, *wh = A->wh, *dsq = A->dsq;
    float *vcd = A->vcd, *hcd = A->hcd, *vcdalt = A->vcdalt, *hcdalt = A->hcdalt;
    float *dgv = A->dgv, *dgh = A->dgh, *cdsq = A->cdsq, *hvwt = A->hvwt;

    /* stage 1: gradients and direction weights, :1331-1339 */
    for (int rr = 2; rr < rr1 - 2; rr++)
        for (int cc = 2; cc < cc1 - 2; cc++) {
            int i = rr * TS + cc;
            float dh = fabsf(cfa[i + 1] - cfa[i - 1]);
            float dv = fabsf(cfa[i + v1] - cfa[i - v1]);
            wv[i] = EPS + fabsf(cfa[i + v2] - cfa[i]) + fabsf(cfa[i] - cfa[i - v2]) + dv;
            wh[i] = EPS + fabsf(cfa[i + 2] - cfa[i]) + fabsf(cfa[i] - cfa[i - 2]) + dh;
            dsq[i] = sq(dh) + sq(dv);
        }

    /* stage 2: vertical/horizontal colour differences, :1342-1373 */
    for (int rr = 4; rr < rr1 - 4; rr++)
        for (int cc = 4; cc < cc1 - 4; cc++) {
            int i = rr * TS + cc;
            float sgn = (fct(P, rr, cc) & 1) ? -1.f : 1.f;
            float c0 = cfa[i];
            float ru = cfa[i - v1] * (wv[i - v2] + wv[i]) / (wv[i - v2] * (EPS + c0) + wv[i] * (EPS + cfa[i - v2]));
            float rd = cfa[i + v1] * (wv[i + v2] + wv[i]) / (wv[i + v2] * (EPS + c0) + wv[i] * (EPS + cfa[i + v2]));
            float rl = cfa[i - 1] * (wh[i - 2] + wh[i]) / (wh[i - 2] * (EPS + c0) + wh[i] * (EPS + cfa[i - 2]));
            float rr_ = cfa[i + 1] * (wh[i + 2] + wh[i]) / (wh[i + 2] * (EPS + c0) + wh[i] * (EPS + cfa[i + 2]));
            float gu_h = cfa[i - v1] + half(c0 - cfa[i - v2]);
            float gd_h = cfa[i + v1] + half(c0 - cfa[i + v2]);
            float gl_h = cfa[i - 1] + half(c0 - cfa[i - 2]);
            float gr_h = cfa[i + 1] + half(c0 - cfa[i + 2]);
            float gu_a = fabsf(1.f - ru) < ARTHRESH ? c0 * ru : gu_h;
            float gd_a = fabsf(1.f - rd) < ARTHRESH ? c0 * rd : gd_h;
            float gl_a = fabsf(1.f - rl) < ARTHRESH ? c0 * rl : gl_h;
            float gr_a = fabsf(1.f - rr_) < ARTHRESH ? c0 * rr_ : gr_h;
            float hw = wh[i - 1] / (wh[i - 1] + wh[i + 1]);
            float vw = wv[i - v1] / (wv[i + v1] + wv[i - v1]);
            float gv_h = vw * gd_h + (1.f - vw) * gu_h;
            float gh_h = hw * gr_h + (1.f - hw) * gl_h;
            vcdalt[i] = sgn * (gv_h - c0);
            hcdalt[i] = sgn * (gh_h - c0);
            if (c0 > P->clip_pt8 || gv_h > P->clip_pt8 || gh_h > P->clip_pt8) {
                vcd[i] = vcdalt[i];
                hcd[i] = hcdalt[i];
                gu_a = gu_h; gd_a = gd_h; gl_a = gl_h; gr_a = gr_h;
            } else {
                vcd[i] = sgn * (vw * gd_a + (1.f - vw) * gu_a - c0);
                hcd[i] = sgn * (hw * gr_a + (1.f - hw) * gl_a - c0);
            }
            dgv[i] = fmin_std(sq(gu_h - gd_h), sq(gu_a - gd_a));
            dgh[i] = fmin_std(sq(gl_h - gr_h), sq(gl_a - gr_a));
        }

    /* stage 3: variance test and bounding, in place in raster order, :1376-1397 */
    for (int rr = 4; rr < rr1 - 4; rr++)
        for (int cc = 4; cc < cc1 - 4; cc++) {
            int i = rr * TS + cc;
            float c0 = cfa[i];
            float hv = 3.f * (sq(hcd[i - 2]) + sq(hcd[i]) + sq(hcd[i + 2])) - sq(hcd[i - 2] + hcd[i] + hcd[i + 2]);
            float hav = 3.f * (sq(hcdalt[i - 2]) + sq(hcdalt[i]) + sq(hcdalt[i + 2])) - sq(hcdalt[i - 2] + hcdalt[i] + hcdalt[i + 2]);
            float vv = 3.f * (sq(vcd[i - v2]) + sq(vcd[i]) + sq(vcd[i + v2])) - sq(vcd[i - v2] + vcd[i] + vcd[i + v2]);
            float vav = 3.f * (sq(vcdalt[i - v2]) + sq(vcdalt[i]) + sq(vcdalt[i + v2])) - sq(vcdalt[i - v2] + vcdalt[i] + vcdalt[i + v2]);
            if (hav < hv) hcd[i] = hcdalt[i];
            if (vav < vv) vcd[i] = vcdalt[i];
            float gh, gv;
            int is_green = fct(P, rr, cc) & 1;
            if (is_green) {
                gh = -hcd[i] + c0;
                gv = -vcd[i] + c0;
                if (hcd[i] > 0) {
                    if (3.f * hcd[i] > (gh + c0))
                        hcd[i] = -med3(gh, cfa[i - 1], cfa[i + 1]) + c0;
                    else {
                        float w2 = 1.f - 3.f * hcd[i] / (EPS + gh + c0);
                        hcd[i] = w2 * hcd[i] + (1.f - w2) * (-med3(gh, cfa[i - 1], cfa[i + 1]) + c0);
                    }
                }
                if (vcd[i] > 0) {
                    if (3.f * vcd[i] > (gv + c0))
                        vcd[i] = -med3(gv, cfa[i - v1], cfa[i + v1]) + c0;
                    else {
                        float w2 = 1.f - 3.f * vcd[i] / (EPS + gv + c0);
                        vcd[i] = w2 * vcd[i] + (1.f - w2) * (-med3(gv, cfa[i - v1], cfa[i + v1]) + c0);
                    }
                }
            } else {
                gh = hcd[i] + c0;
                gv = vcd[i] + c0;
                if (hcd[i] < 0) {
                    if (3.f * hcd[i] < -(gh + c0))
                        hcd[i] = med3(gh, cfa[i - 1], cfa[i + 1]) - c0;
                    else {
                        float w2 = 1.f + 3.f * hcd[i] / (EPS + gh + c0);
                        hcd[i] = w2 * hcd[i] + (1.f - w2) * (med3(gh, cfa[i - 1], cfa[i + 1]) - c0);
                    }
                }
                if (vcd[i] < 0) {
                    if (3.f * vcd[i] < -(gv + c0))
                        vcd[i] = med3(gv, cfa[i - v1], cfa[i + v1]) - c0;
                    else {
                        float w2 = 1.f + 3.f * vcd[i] / (EPS + gv + c0);
                        vcd[i] = w2 * vcd[i] + (1.f - w2) * (med3(gv, cfa[i - v1], cfa[i + v1]) - c0);
                    }
                }
            }
            float s = is_green ? -1.f : 1.f;
            if (gh > P->clip_pt) hcd[i] = s * (med3(gh, cfa[i - 1], cfa[i + 1]) - c0);
            if (gv > P->clip_pt) vcd[i] = s * (med3(gv, cfa[i - v1], cfa[i + v1]) - c0);
            cdsq[i] = sq(vcd[i] - hcd[i]);
        }

    /* stages 4-5: directional variances -> hvwt; Nyquist texture test, :1400-1433 */
    for (int rr = 6; rr < rr1 - 6; rr++)
        for (int cc = 6 + (fct(P, rr, 2) & 1); cc < cc1 - 6; cc += 2) {
            int i = rr * TS + cc;
            float uave = vcd[i] + vcd[i - v1] + vcd[i - v2] + vcd[i - v3];
            float dave = vcd[i] + vcd[i + v1] + vcd[i + v2] + vcd[i + v3];
            float lave = hcd[i] + hcd[i - 1] + hcd[i - 2] + hcd[i - 3];
            float rave = hcd[i] + hcd[i + 1] + hcd[i + 2] + hcd[i + 3];
            float vu = sq(vcd[i] - uave) + sq(vcd[i - v1] - uave) + sq(vcd[i - v2] - uave) + sq(vcd[i - v3] - uave);
            float vd = sq(vcd[i] - dave) + sq(vcd[i + v1] - dave) + sq(vcd[i + v2] - dave) + sq(vcd[i + v3] - dave);
            float hl = sq(hcd[i] - lave) + sq(hcd[i - 1] - lave) + sq(hcd[i - 2] - lave) + sq(hcd[i - 3] - lave);
            float hr = sq(hcd[i] - rave) + sq(hcd[i + 1] - rave) + sq(hcd[i + 2] - rave) + sq(hcd[i + 3] - rave);
            float hw = wh[i - 1] / (wh[i - 1] + wh[i + 1]);
            float vw = wv[i - v1] / (wv[i + v1] + wv[i - v1]);
            float vvar = EPSSQ + vw * vd + (1.f - vw) * vu;
            float hvar = EPSSQ + hw * hr + (1.f - hw) * hl;
            float vu2 = dgv[i] + dgv[i - v1] + dgv[i - v2];
            float vd2 = dgv[i] + dgv[i + v1] + dgv[i + v2];
            float hl2 = dgh[i] + dgh[i - 1] + dgh[i - 2];
            float hr2 = dgh[i] + dgh[i + 1] + dgh[i + 2];
            float vvar1 = EPSSQ + vw * vd2 + (1.f - vw) * vu2;
            float hvar1 = EPSSQ + hw * hr2 + (1.f - hw) * hl2;
            float varwt = hvar / (vvar + hvar);
            float diffwt = hvar1 / (vvar1 + hvar1);
            hvwt[i >> 1] = ((0.5f - varwt) * (0.5f - diffwt) > 0.f && fabsf(0.5f - diffwt) < fabsf(0.5f - varwt)) ? varwt : diffwt;
        }
    for (int rr = 6; rr < rr1 - 6; rr++)
        for (int cc = 6 + (fct(P, rr, 2) & 1); cc < cc1 - 6; cc += 2) {
            int i = rr * TS + cc;
            const float *d = cdsq, *g = dsq;
            A->nyqtest[i >> 1] =
                (GODD[0] * d[i] + GODD[1] * (d[i - m1] + d[i + p1] + d[i - p1] + d[i + m1]) +
                 GODD[2] * (d[i - v2] + d[i - 2] + d[i + 2] + d[i + v2]) +
                 GODD[3] * (d[i - m2] + d[i + p2] + d[i - p2] + d[i + m2])) -
                (ggrad(0) * g[i] + ggrad(1) * (g[i - v1] + g[i + 1] + g[i - 1] + g[i + v1]) +
                 ggrad(2) * (g[i - m1] + g[i + p1] + g[i - p1] + g[i + m1]) +
                 ggrad(3) * (g[i - v2] + g[i - 2] + g[i + 2] + g[i + v2]) +
                 ggrad(4) * (g[i - v2 - 1] + g[i - v2 + 1] + g[i - TS - 2] + g[i - TS + 2] + g[i + TS - 2] + g[i + TS + 2] + g[i + v2 - 1] + g[i + v2 + 1]) +
                 ggrad(5) * (g[i - m2] + g[i + p2] + g[i - p2] + g[i + m2]));
        }
    int any_nyq = 0;
    memset(A->nyq, 0, TS * TSH);
    for (int rr = 6; rr < rr1 - 6; rr++)
        for (int cc = 6 + (fct(P, rr, 2) & 1); cc < cc1 - 6; cc += 2) {
            int i = rr * TS + cc;
            if (A->nyqtest[i >> 1] > 0.f) { A->nyq[i >> 1] = 1; any_nyq = 1; }
        }

    /* stage 6: Nyquist refinement and green at R/B sites, :1436-1481 */
    unsigned char *nyq = A->nyq, *nyq2 = A->nyq2;
    if (any_nyq) {
        memset(nyq2, 0, TS * TSH);
        for (int rr = 8; rr < rr1 - 8; rr++)
            for (int cc = 8 + (fct(P, rr, 2) & 1); cc < cc1 - 8; cc += 2) {
                int i = rr * TS + cc;
                unsigned s = nyq[(i - v2) >> 1] + nyq[(i - m1) >> 1] + nyq[(i + p1) >> 1] + nyq[(i - 2) >> 1] +
                             nyq[(i + 2) >> 1] + nyq[(i - p1) >> 1] + nyq[(i + m1) >> 1] + nyq[(i + v2) >> 1];
                nyq2[i >> 1] = s > 4 ? 1 : (s < 4 ? 0 : nyq[i >> 1]);
            }
        for (int rr = 8; rr < rr1 - 8; rr++)
            for (int cc = 8 + (fct(P, rr, 2) & 1); cc < cc1 - 8; cc += 2) {
                int i = rr * TS + cc;
                if (!nyq2[i >> 1]) continue;
                float scfa = 0, sh = 0, sv = 0, sqh = 0, sqv = 0, area = 0;
                for (int a = -6; a < 7; a += 2)
                    for (int b = -6; b < 7; b += 2) {
                        int j = i + a * TS + b;
                        if (j >= 0 && j < N2 && nyq2[j >> 1]) {
                            float t = cfa[j];
                            scfa += t;
                            sh += (cfa[j - 1] + cfa[j + 1]);
                            sv += (cfa[j - v1] + cfa[j + v1]);
                            sqh += sq(t - cfa[j - 1]) + sq(t - cfa[j + 1]);
                            sqv += sq(t - cfa[j - v1]) + sq(t - cfa[j + v1]);
                            area += 1;
                        }
                    }
                sh = scfa - half(sh);
                sv = scfa - half(sv);
                area = half(area);
                float hv2 = EPSSQ + fabsf(area * sqh - sh * sh), vv2 = EPSSQ + fabsf(area * sqv - sv * sv);
                hvwt[i >> 1] = hv2 / (vv2 + hv2);
            }
    }
    float *dg0 = vcdalt;      /* Dgrb0 lives over vcdalt, half-index */
    float *d2 = A->q1m;       /* Dgrb2 {h,v} pairs over Dgrbsq1m/Dgrbsq1p */
    for (int rr = 8; rr < rr1 - 8; rr++)
        for (int cc = 8 + (fct(P, rr, 2) & 1); cc < cc1 - 8; cc += 2) {
            int i = rr * TS + cc, k = i >> 1;
            float alt = 0.25f * (hvwt[(i - m1) >> 1] + hvwt[(i + p1) >> 1] + hvwt[(i - p1) >> 1] + hvwt[(i + m1) >> 1]);
            hvwt[k] = fabsf(0.5f - hvwt[k]) < fabsf(0.5f - alt) ? alt : hvwt[k];
            dg0[k] = hvwt[k] * vcd[i] + (1.f - hvwt[k]) * hcd[i];
            green[i] = cfa[i] + dg0[k];
            d2[2 * k] = nyq2[k] ? sq(green[i] - half(green[i - 1] + green[i + 1])) : 0.f;
            d2[2 * k + 1] = nyq2[k] ? sq(green[i] - half(green[i - v1] + green[i + v1])) : 0.f;
        }
    if (any_nyq) {
        for (int rr = 8; rr < rr1 - 8; rr++)
            for (int cc = 8 + (fct(P, rr, 2) & 1); cc < cc1 - 8; cc += 2) {
                int i = rr * TS + cc, k = i >> 1;
                if (!nyq2[k]) continue;
#define D2H(o) d2[2 * ((i + (o)) >> 1)]
#define D2V(o) d2[2 * ((i + (o)) >> 1) + 1]
                float gvh = EPSSQ + (GQUINC[0] * D2H(0) + GQUINC[1] * (D2H(-m1) + D2H(p1) + D2H(-p1) + D2H(m1)) +
                                     GQUINC[2] * (D2H(-v2) + D2H(-2) + D2H(2) + D2H(v2)) +
                                     GQUINC[3] * (D2H(-m2) + D2H(p2) + D2H(-p2) + D2H(m2)));
                float gvv = EPSSQ + (GQUINC[0] * D2V(0) + GQUINC[1] * (D2V(-m1) + D2V(p1) + D2V(-p1) + D2V(m1)) +
                                     GQUINC[2] * (D2V(-v2) + D2V(-2) + D2V(2) + D2V(v2)) +
                                     GQUINC[3] * (D2V(-m2) + D2V(p2) + D2V(-p2) + D2V(m2)));
#undef D2H
#undef D2V
                dg0[k] = (hcd[i] * gvv + vcd[i] * gvh) / (gvv + gvh);
                green[i] = cfa[i] + dg0[k];
            }
    }

    /* stage 7: diagonal gradients, :1484-1502 */
    float *delp = cdsq, *delm = cdsq + TS * TSH, *q1p = A->q1p, *q1m = A->q1m;
    for (int rr = 6; rr < rr1 - 6; rr++) {
        if ((fct(P, rr, 2) & 1) == 0) {
            for (int cc = 6; cc < cc1 - 6; cc += 2) {
                int i = rr * TS + cc;
                delp[i >> 1] = fabsf(cfa[i + p1] - cfa[i - p1]);
                delm[i >> 1] = fabsf(cfa[i + m1] - cfa[i - m1]);
                q1p[i >> 1] = sq(cfa[i + 1] - cfa[i + 1 - p1]) + sq(cfa[i + 1] - cfa[i + 1 + p1]);
                q1m[i >> 1] = sq(cfa[i + 1] - cfa[i + 1 - m1]) + sq(cfa[i + 1] - cfa[i + 1 + m1]);
            }
        } else {
            for (int cc = 6; cc < cc1 - 6; cc += 2) {
                int i = rr * TS + cc;
                q1p[i >> 1] = sq(cfa[i] - cfa[i - p1]) + sq(cfa[i] - cfa[i + p1]);
                q1m[i >> 1] = sq(cfa[i] - cfa[i - m1]) + sq(cfa[i] - cfa[i + m1]);
                delp[i >> 1] = fabsf(cfa[i + 1 + p1] - cfa[i + 1 - p1]);
                delm[i >> 1] = fabsf(cfa[i + 1 + m1] - cfa[i + 1 - m1]);
            }
        }
    }

    /* stage 8: R/B along the diagonals, :1505-1532 */
    float *rbm = vcd, *rbp = hcdalt, *pmwt = wh;
    for (int rr = 8; rr < rr1 - 8; rr++)
        for (int cc = 8 + (fct(P, rr, 2) & 1); cc < cc1 - 8; cc += 2) {
            int i = rr * TS + cc, k = i >> 1;
            float c0 = cfa[i];
            float rse = twice(cfa[i + m1]) / (EPS + c0 + cfa[i + m2]);
            float rnw = twice(cfa[i - m1]) / (EPS + c0 + cfa[i - m2]);
            float rne = twice(cfa[i + p1]) / (EPS + c0 + cfa[i + p2]);
            float rsw = twice(cfa[i - p1]) / (EPS + c0 + cfa[i - p2]);
            float bse = fabsf(1.f - rse) < ARTHRESH ? c0 * rse : (cfa[i + m1]) + half(c0 - cfa[i + m2]);
            float bnw = fabsf(1.f - rnw) < ARTHRESH ? c0 * rnw : (cfa[i - m1]) + half(c0 - cfa[i - m2]);
            float bne = fabsf(1.f - rne) < ARTHRESH ? c0 * rne : (cfa[i + p1]) + half(c0 - cfa[i + p2]);
            float bsw = fabsf(1.f - rsw) < ARTHRESH ? c0 * rsw : (cfa[i - p1]) + half(c0 - cfa[i - p2]);
            float wse = EPS + delm[k] + delm[(i + m1) >> 1] + delm[(i + m2) >> 1];
            float wnw = EPS + delm[k] + delm[(i - m1) >> 1] + delm[(i - m2) >> 1];
            float wne = EPS + delp[k] + delp[(i + p1) >> 1] + delp[(i + p2) >> 1];
            float wsw = EPS + delp[k] + delp[(i - p1) >> 1] + delp[(i - p2) >> 1];
            rbm[k] = (wse * bnw + wnw * bse) / (wse + wnw);
            rbp[k] = (wne * bsw + wsw * bne) / (wne + wsw);
            if (rbp[k] < c0) {
                if (twice(rbp[k]) < c0)
                    rbp[k] = med3(rbp[k], cfa[i - p1], cfa[i + p1]);
                else {
                    float w = twice(c0 - rbp[k]) / (EPS + rbp[k] + c0);
                    rbp[k] = w * rbp[k] + (1.f - w) * med3(rbp[k], cfa[i - p1], cfa[i + p1]);
                }
            }
            if (rbm[k] < c0) {
                if (twice(rbm[k]) < c0)
                    rbm[k] = med3(rbm[k], cfa[i - m1], cfa[i + m1]);
                else {
                    float w = twice(c0 - rbm[k]) / (EPS + rbm[k] + c0);
                    rbm[k] = w * rbm[k] + (1.f - w) * med3(rbm[k], cfa[i - m1], cfa[i + m1]);
                }
            }
            if (rbp[k] > P->clip_pt) rbp[k] = med3(rbp[k], cfa[i - p1], cfa[i + p1]);
            if (rbm[k] > P->clip_pt) rbm[k] = med3(rbm[k], cfa[i - m1], cfa[i + m1]);
#define QS(a, o) a[(i + (o)) >> 1]
            float varm = EPSSQ + (GEVEN[0] * (QS(q1m, -v1) + QS(q1m, -1) + QS(q1m, 1) + QS(q1m, v1)) +
                                  GEVEN[1] * (QS(q1m, -v2 - 1) + QS(q1m, -v2 + 1) + QS(q1m, -2 - v1) + QS(q1m, 2 - v1) +
                                              QS(q1m, -2 + v1) + QS(q1m, 2 + v1) + QS(q1m, v2 - 1) + QS(q1m, v2 + 1)));
            float varp = EPSSQ + (GEVEN[0] * (QS(q1p, -v1) + QS(q1p, -1) + QS(q1p, 1) + QS(q1p, v1)) +
                                  GEVEN[1] * (QS(q1p, -v2 - 1) + QS(q1p, -v2 + 1) + QS(q1p, -2 - v1) + QS(q1p, 2 - v1) +
                                              QS(q1p, -2 + v1) + QS(q1p, 2 + v1) + QS(q1p, v2 - 1) + QS(q1p, v2 + 1)));
#undef QS
            pmwt[k] = varm / (varp + varm);
        }

    /* stage 9: R+B average, green re-interpolation, :1535-1592 */
    float *rbint = dsq;
    for (int rr = 10; rr < rr1 - 10; rr++)
        for (int cc = 10 + (fct(P, rr, 2) & 1); cc < cc1 - 10; cc += 2) {
            int i = rr * TS + cc, k = i >> 1;
            float alt = 0.25f * (pmwt[(i - m1) >> 1] + pmwt[(i + p1) >> 1] + pmwt[(i - p1) >> 1] + pmwt[(i + m1) >> 1]);
            if (fabsf(0.5f - pmwt[k]) < fabsf(0.5f - alt)) pmwt[k] = alt;
            rbint[k] = half(cfa[i] + rbm[k] * (1.f - pmwt[k]) + rbp[k] * pmwt[k]);
        }
    for (int rr = 12; rr < rr1 - 12; rr++)
        for (int cc = 12 + (fct(P, rr, 2) & 1); cc < cc1 - 12; cc += 2) {
            int i = rr * TS + cc, k = i >> 1;
            if (fabsf(0.5f - pmwt[k]) >= fabsf(0.5f - hvwt[k])) continue;
            float rb = rbint[k];
            float ru = twice(cfa[i - v1]) / (EPS + rb + rbint[k - v1]);
            float rd = twice(cfa[i + v1]) / (EPS + rb + rbint[k + v1]);
            float rl = twice(cfa[i - 1]) / (EPS + rb + rbint[k - 1]);
            float rr_ = twice(cfa[i + 1]) / (EPS + rb + rbint[k + 1]);
            float gu = fabsf(1.f - ru) < ARTHRESH ? rb * ru : cfa[i - v1] + half(rb - rbint[k - v1]);
            float gd = fabsf(1.f - rd) < ARTHRESH ? rb * rd : cfa[i + v1] + half(rb - rbint[k + v1]);
            float gl = fabsf(1.f - rl) < ARTHRESH ? rb * rl : cfa[i - 1] + half(rb - rbint[k - 1]);
            float gr = fabsf(1.f - rr_) < ARTHRESH ? rb * rr_ : cfa[i + 1] + half(rb - rbint[k + 1]);
            float gv = (wv[i - v1] * gd + wv[i + v1] * gu) / (wv[i + v1] + wv[i - v1]);
            float gh = (wh[i - 1] * gr + wh[i + 1] * gl) / (wh[i - 1] + wh[i + 1]);
            if (gv < rb) {
                if (twice(gv) < rb)
                    gv = med3(gv, cfa[i - v1], cfa[i + v1]);
                else {
                    float w = twice(rb - gv) / (EPS + gv + rb);
                    gv = w * gv + (1.f - w) * med3(gv, cfa[i - v1], cfa[i + v1]);
                }
            }
            if (gh < rb) {
                if (twice(gh) < rb)
                    gh = med3(gh, cfa[i - 1], cfa[i + 1]);
                else {
                    float w = twice(rb - gh) / (EPS + gh + rb);
                    gh = w * gh + (1.f - w) * med3(gh, cfa[i - 1], cfa[i + 1]);
                }
            }
            if (gh > P->clip_pt) gh = med3(gh, cfa[i - 1], cfa[i + 1]);
            if (gv > P->clip_pt) gv = med3(gv, cfa[i - v1], cfa[i + v1]);
            green[i] = gh * (1.f - hvwt[k]) + gv * hvwt[k];
            dg0[k] = green[i] - cfa[i];
        }

    /* colour differences of the second chroma, :1596-1625 */
    float *dg1 = hcd; /* Dgrb1 lives over hcd, half-index */
    for (int rr = 13 - P->ey; rr < rr1 - 12; rr += 2)
        for (int k = (rr * TS + 13 - P->ex) >> 1; k < (rr * TS + cc1 - 12) >> 1; k++) {
            dg1[k] = dg0[k];
            dg0[k] = 0;
        }
    for (int rr = 14; rr < rr1 - 14; rr++)
        for (int cc = 14 + (fct(P, rr, 2) & 1); cc < cc1 - 14; cc += 2) {
            int i = rr * TS + cc;
            float *D = (1 - (int)fct(P, rr, cc) / 2) ? dg1 : dg0;
#define DD(o) D[(i + (o)) >> 1]
            float wnw = 1.f / (EPS + fabsf(DD(-m1) - DD(m1)) + fabsf(DD(-m1) - DD(-m3)) + fabsf(DD(m1) - DD(-m3)));
            float wne = 1.f / (EPS + fabsf(DD(p1) - DD(-p1)) + fabsf(DD(p1) - DD(p3)) + fabsf(DD(-p1) - DD(p3)));
            float wsw = 1.f / (EPS + fabsf(DD(-p1) - DD(p1)) + fabsf(DD(-p1) - DD(m3)) + fabsf(DD(p1) - DD(-p3)));
            float wse = 1.f / (EPS + fabsf(DD(m1) - DD(-m1)) + fabsf(DD(m1) - DD(-p3)) + fabsf(DD(-m1) - DD(m3)));
            D[i >> 1] = (wnw * (1.325f * DD(-m1) - 0.175f * DD(-m3) - 0.075f * (DD(-m1 - 2) + DD(-m1 - v2))) +
                         wne * (1.325f * DD(p1) - 0.175f * DD(p3) - 0.075f * (DD(p1 + 2) + DD(p1 + v2))) +
                         wsw * (1.325f * DD(-p1) - 0.175f * DD(-p3) - 0.075f * (DD(-p1 - 2) + DD(-p1 - v2))) +
                         wse * (1.325f * DD(m1) - 0.175f * DD(m3) - 0.075f * (DD(m1 + 2) + DD(m1 + v2)))) /
                        (wnw + wne + wsw + wse);
#undef DD
        }

    /* stage 10: output, :1597-1625 */
    for (int rr = 16; rr < rr1 - 16; rr++) {
        int row = rr + top;
        if (row < 0 || row >= H) continue;
        for (int cc = 16; cc < cc1 - 16; cc++) {
            int col = cc + left;
            if (col < 0 || col >= W) continue;
            int t = rr * TS + cc;
            float r, g, b;
            if (fct(P, row, col) == 1) {
                g = cfa[t];
                float hu = hvwt[(t - v1) >> 1], hr = hvwt[(t + 1) >> 1], hl = hvwt[(t - 1) >> 1], hd = hvwt[(t + v1) >> 1];
                float inv = 1.0f / (hu + 2.f - hr - hl + hd);
                float rd = (hu * dg0[(t - v1) >> 1] + (1.f - hr) * dg0[(t + 1) >> 1] + (1.f - hl) * dg0[(t - 1) >> 1] + hd * dg0[(t + v1) >> 1]) * inv;
                float bd = (hu * dg1[(t - v1) >> 1] + (1.f - hr) * dg1[(t + 1) >> 1] + (1.f - hl) * dg1[(t - 1) >> 1] + hd * dg1[(t + v1) >> 1]) * inv;
                r = g - rd;
                b = g - bd;
            } else {
                g = green[t];
                r = g - dg0[t >> 1];
                b = g - dg1[t >> 1];
            }
            float *o = P->out + ((size_t)row * W + col) * 3;
            o[0] = (0.f < r) ? r : 0.f;
            o[1] = (0.f < g) ? g : 0.f;
            o[2] = (0.f < b) ? b : 0.f;
        }
    }
}

/* border_interpolate, core/cpu_accelerator.cpp:2984-3028 */
void oracle_border_interpolate(float *img, int w, int h, uint32_t filters, int border) {
#pragma omp parallel for schedule(static)
    for (int row = 0; row < h; row++)
        for (int col = 0; col < w; col++) {
            if (col >= border && col < w - border && row >= border && row < h - border) continue;
            float sum[8] = {0};
            for (int y = row > 0 ? row - 1 : 0; y < row + 2 && y < h; y++)
                for (int x = col > 0 ? col - 1 : 0; x < col + 2 && x < w; x++) {
                    unsigned f = bayer3(filters, (unsigned)y, (unsigned)x);
                    sum[f] += img[((size_t)y * w + x) * 3 + f];
                    sum[f + 4]++;
                }
            unsigned f = bayer3(filters, (unsigned)row, (unsigned)col);
            for (unsigned c = 0; c < 3; c++)
                if (c != f && sum[c + 4] > 0) img[((size_t)row * w + col) * 3 + c] = sum[c] / sum[c + 4];
        }
}

/* demosaic_bayer_amaze, core/cpu_accelerator.cpp:1635-1653 (constructor :1184-1206) */
int oracle_amaze(const float *raw3, float *rgb3, int w, int h, uint32_t filters,
                 const float *cam_mul4, float maximum_value, int tile_mode) {
    if (!raw3 || !rgb3 || w <= 0 || h <= 0) return -1;
    amaze_t P;
    P.raw = raw3;
    P.out = rgb3;
    P.W = w;
    P.H = h;
    float mx = cam_mul4[0] < cam_mul4[2] ? cam_mul4[2] : cam_mul4[0];   /* std::max({R,B}) */
    float mn = cam_mul4[2] < cam_mul4[0] ? cam_mul4[2] : cam_mul4[0];   /* std::min({R,B}) */
    float gain = (mn > 1e-6f) ? (mx / mn) : 1.0f;
    P.clip_pt = 1.0f / gain;
    P.clip_pt8 = 0.8f / gain;
    P.scale = 1.0f / maximum_value;
    P.minimal_clear = tile_mode == ORACLE_TILES_POISON_NAN || tile_mode == ORACLE_TILES_POISON_BIG;
    for (int r = 0; r < 2; r++)
        for (int c = 0; c < 2; c++) P.cf[r][c] = bayer3(filters, (unsigned)r, (unsigned)c);
    if (P.cf[0][0] == 1) {
        if (P.cf[0][1] == 0) { P.ey = 0; P.ex = 1; } else { P.ey = 1; P.ex = 0; }
    } else {
        if (P.cf[0][0] == 0) { P.ey = 0; P.ex = 0; } else { P.ey = 1; P.ex = 1; }
    }
    const size_t arena_bytes = 15 * sizeof(float) * N2 + TS * TSH + 256;
    if (tile_mode == ORACLE_TILES_SEQUENTIAL) {
        void *mem = calloc(1, arena_bytes);
        arena_t A;
        arena_init(&A, mem);
        for (int top = -16; top < h; top += TS - 32)
            for (int left = -16; left < w; left += TS - 32) amaze_tile(&P, &A, top, left);
        free(mem);
    } else {
        int ntr = 0, ntc = 0;
        for (int top = -16; top < h; top += TS - 32) ntr++;
        for (int left = -16; left < w; left += TS - 32) ntc++;
#pragma omp parallel
        {
            void *mem = malloc(arena_bytes);
            arena_t A;
            arena_init(&A, mem);
#pragma omp for schedule(dynamic)
            for (int t = 0; t < ntr * ntc; t++) {
                if (P.minimal_clear) {
                    /* stale scratch: NaN, or a large finite value */
                    const uint32_t pat = tile_mode == ORACLE_TILES_POISON_NAN ? 0x7fc0a5a5u : 0x4b4b4b4bu;
                    uint32_t *u = (uint32_t *)mem;
                    for (size_t k = 0; k < arena_bytes / 4; k++) u[k] = pat;
                } else {
                    memset(mem, 0, arena_bytes);
                }
                amaze_tile(&P, &A, -16 + (t / ntc) * (TS - 32), -16 + (t % ntc) * (TS - 32));
            }
            free(mem);
        }
    }
    oracle_border_interpolate(rgb3, w, h, filters, 4);
    return 0;
}
