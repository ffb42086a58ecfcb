/*
 * TEST INFRASTRUCTURE ONLY — CPU restatement of the reference's X-Trans 3-pass
 * (Markesteijn) demosaic, XTrans_Processor::run_3pass
 * (core/cpu_accelerator.cpp:1870-2676) and xtrans_border_interpolate
 * (:1806-1861).  Never called by the product.
 *
 * The reference works on 114x114 tiles (step 98, starting at (3,3)) with
 * per-thread buffers that are allocated once and reused; a few reads see cells
 * not written for the current tile (rows of the homogeneity map at the
 * bottom/right of the last tile row/column, :2497-2559).  ORACLE_TILES_SEQUENTIAL
 * reproduces the single-thread reference exactly (one set of buffers, raster
 * tile order); ORACLE_TILES_CLEAN clears the buffers for every tile.
 * The CIELab branch of the reference is disabled (`if (false)`, :2322), so the
 * YPbPr path is the one restated; the NEON blocks are not compiled on x86.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

enum { TS = 114 };
typedef struct { float min, max; } mm_t;

typedef struct {
    const float *raw;
    float *out;
    int W, H;
    float maxv;
    signed char xt[6][6];
    short allhex[2][3][3][8];
    int sgrow, sgcol;
    int right_shift[3];
} xt_t;

typedef struct {
    float *rgb_base;          /* [8][TS][TS][3] */
    float *lab;               /* [3][TS-8][TS-8] */
    float *drv;               /* [8][TS-10][TS-10] */
    unsigned char *homo;      /* [8][TS][TS] */
    unsigned char *homosum;   /* [8][TS][TS] */
    unsigned char *homosummax;/* [TS][TS] */
    mm_t *gmm;                /* [TS][TS/2] */
} xbuf_t;

#define RGBP(base, d, r, c) ((base) + ((((size_t)(d) * TS + (r)) * TS + (c)) * 3))

static inline int fcolx(const xt_t *P, int row, int col) { return P->xt[(row + 6) % 6][(col + 6) % 6]; }
static inline int isgreen(const xt_t *P, int row, int col) { return P->xt[row % 3][col % 3] & 1; }
static inline float rawv(const xt_t *P, int row, int col) {
    return P->raw[((size_t)row * P->W + col) * 3 + fcolx(P, row, col)] / P->maxv;
}
static inline float rawhex(const xt_t *P, int row, int col, int hex) {
    /* :1691-1695 — the offset is applied to the linear index */
    unsigned pos = (unsigned)(row * P->W + col + hex);
    int f = fcolx(P, (int)(pos / (unsigned)P->W), (int)(pos % (unsigned)P->W));
    return P->raw[(size_t)pos * 3 + f] / P->maxv;
}
static inline float limf(float v, float lo, float hi) { /* LIM, shader_common.h:50-53 */
    float m = (hi < v) ? hi : v;
    return (lo < m) ? m : lo;
}
static inline float sq(float x) { return x * x; }

/* :1806-1861 */
static void border8(const xt_t *P, int border) {
    static const float wt[3][3] = {{0.25f, 0.5f, 0.25f}, {0.5f, 0.f, 0.5f}, {0.25f, 0.5f, 0.25f}};
    const int W = P->W, H = P->H;
#pragma omp parallel for schedule(static)
    for (int row = 0; row < H; row++)
        for (int col = 0; col < W; col++) {
            if (col >= border && col < W - border && row >= border && row < H - border) continue;
            float sum[6] = {0};
            for (int y = row > 0 ? row - 1 : 0, v = row == 0 ? 0 : -1; y < row + 2 && y < H; y++, v++)
                for (int x = col > 0 ? col - 1 : 0, h = col == 0 ? 0 : -1; x < col + 2 && x < W; x++, h++) {
                    int f = fcolx(P, y, x);
                    sum[f] += rawv(P, y, x) * wt[v + 1][h + 1];
                    sum[f + 3] += wt[v + 1][h + 1];
                }
            float *o = P->out + ((size_t)row * W + col) * 3;
            switch (fcolx(P, row, col)) {
            case 0:
                o[0] = rawv(P, row, col);
                o[1] = sum[1] / sum[4];
                o[2] = sum[2] / sum[5];
                break;
            case 1:
                if (sum[3] == 0.f) {
                    float v = rawv(P, row, col);
                    o[0] = o[1] = o[2] = v;
                } else {
                    o[0] = sum[0] / sum[3];
                    o[1] = rawv(P, row, col);
                    o[2] = sum[2] / sum[5];
                }
                break;
            case 2:
                o[0] = sum[0] / sum[3];
                o[1] = sum[1] / sum[4];
                o[2] = rawv(P, row, col);
            }
        }
}

/* green min/max of the hexagon around a non-green site (:1975-2059) */
static void minmax_hex(const xt_t *P, int row, int col, const short *hex, float *mn, float *mx) {
    float lo = FLT_MAX, hi = 0.f;
    for (int c = 0; c < 6; c++) {
        float v = rawhex(P, row, col, hex[c]);
        lo = lo < v ? lo : v;
        hi = hi > v ? hi : v;
    }
    *mn = lo;
    *mx = hi;
}

static void interp_green(const xt_t *P, float *rgb, int row, int col, int top, int left,
                         const short *hex, int swap, const mm_t *gm) { /* :2075-2128 */
    float color[4];
    color[0] = 0.6796875f * (rawhex(P, row, col, hex[1]) + rawhex(P, row, col, hex[0])) -
               0.1796875f * (rawhex(P, row, col, 2 * hex[1]) + rawhex(P, row, col, 2 * hex[0]));
    color[1] = 0.87109375f * rawhex(P, row, col, hex[3]) + rawhex(P, row, col, hex[2]) * 0.12890625f +
               0.359375f * (rawhex(P, row, col, 0) - rawhex(P, row, col, -hex[2]));
    for (int c = 0; c < 2; c++)
        color[2 + c] = 0.640625f * rawhex(P, row, col, hex[4 + c]) + 0.359375f * rawhex(P, row, col, -2 * hex[4 + c]) +
                       0.12890625f * (2.f * rawhex(P, row, col, 0) - rawhex(P, row, col, 3 * hex[4 + c]) -
                                      rawhex(P, row, col, -3 * hex[4 + c]));
    const mm_t m = gm[(row - top) * (TS / 2) + ((col - left) >> 1)];
    for (int c = 0; c < 4; c++) RGBP(rgb, c ^ swap, row - top, col - left)[1] = limf(color[c], m.min, m.max);
}

static void xtrans_tile(const xt_t *P, xbuf_t *B, int top, int left) {
    const int W = P->W, H = P->H;
    int mrow = top + TS < H - 3 ? top + TS : H - 3;
    int mcol = left + TS < W - 3 ? left + TS : W - 3;
    mm_t *gm = B->gmm;
#define GM(r, c) gm[(size_t)(r) * (TS / 2) + (c)]
    const int ndir = 8;

    /* green min/max, :1973-2059 */
    for (int row = top; row < mrow; row++) {
        int ls = left;
        for (; ls < mcol; ls++)
            if (!isgreen(P, row, ls)) break;
        int coff = P->right_shift[row % 3] == 1 ? 3 : 1 + (fcolx(P, row, ls + 1) & 1);
        float mn, mx;
        if (coff == 3) {
            const short *hex = P->allhex[0][row % 3][ls % 3];
            for (int col = ls; col < mcol; col += coff) {
                minmax_hex(P, row, col, hex, &mn, &mx);
                GM(row - top, (col - left) >> 1).min = mn;
                GM(row - top, (col - left) >> 1).max = mx;
            }
        } else {
            int col = ls;
            if (coff == 2) {
                minmax_hex(P, row, col, P->allhex[0][row % 3][col % 3], &mn, &mx);
                GM(row - top, (col - left) >> 1).min = mn;
                GM(row - top, (col - left) >> 1).max = mx;
                col += 2;
            }
            const short *hex = P->allhex[0][row % 3][col % 3];
            for (; col < mcol - 1; col += 3) {
                minmax_hex(P, row, col, hex, &mn, &mx);
                GM(row - top, (col - left) >> 1).min = mn;
                GM(row - top, (col - left) >> 1).max = mx;
                GM(row - top, (col + 1 - left) >> 1).min = mn;
                GM(row - top, (col + 1 - left) >> 1).max = mx;
            }
            if (col < mcol) {
                minmax_hex(P, row, col, hex, &mn, &mx);
                GM(row - top, (col - left) >> 1).min = mn;
                GM(row - top, (col - left) >> 1).max = mx;
            }
        }
    }

    float *rgb = B->rgb_base;
    const size_t dirsz = (size_t)TS * TS * 3; /* floats per direction buffer */
    memset(rgb, 0, dirsz * sizeof(float));
    for (int row = top; row < mrow; row++)
        for (int col = left; col < mcol; col++) RGBP(rgb, 0, row - top, col - left)[fcolx(P, row, col)] = rawv(P, row, col);
    for (int c = 0; c < 3; c++) memcpy(rgb + (c + 1) * dirsz, rgb, dirsz * sizeof(float) - sizeof(float));

    /* green along 4 directions, :2075-2128 */
    for (int row = top; row < mrow; row++) {
        int ls = left;
        for (; ls < mcol; ls++)
            if (!isgreen(P, row, ls)) break;
        int coff = P->right_shift[row % 3] == 1 ? 3 : 1 + (fcolx(P, row, ls + 1) & 1);
        if (coff == 3) {
            const short *hex = P->allhex[0][row % 3][ls % 3];
            for (int col = ls; col < mcol; col += coff) interp_green(P, rgb, row, col, top, left, hex, 0, gm);
        } else {
            const short *hm[2] = {P->allhex[0][row % 3][ls % 3], P->allhex[0][row % 3][(ls + coff) % 3]};
            for (int col = ls, hi = 0; col < mcol; col += coff, coff ^= 3, hi ^= 1)
                interp_green(P, rgb, row, col, top, left, hm[hi], 1, gm);
        }
    }

    for (int pass = 0; pass < 3; pass++) {
        if (pass == 1) {
            /* directions 0-3 copied into 4-7 (one float short), later passes refine 4-7 */
            memcpy(rgb + 4 * dirsz, B->rgb_base, 4 * dirsz * sizeof(float) - sizeof(float));
            rgb += 4 * dirsz;
        }
        if (pass) { /* :2141-2182 */
            for (int row = top + 2; row < mrow - 2; row++) {
                int ls = left + 2;
                for (; ls < mcol - 2; ls++)
                    if (!isgreen(P, row, ls)) break;
                int coff = P->right_shift[row % 3] == 1 ? 3 : 1 + (fcolx(P, row, ls + 1) & 1);
                if (coff == 3) {
                    int f = fcolx(P, row, ls);
                    const short *hex = P->allhex[1][row % 3][ls % 3];
                    for (int col = ls; col < mcol - 2; col += coff, f ^= 2)
                        for (int d = 3; d < 6; d++) {
                            float *rix = RGBP(rgb, d - 2, row - top, col - left);
                            float val = 0.33333333f * (rix[3 * (-2 * hex[d]) + 1] + 2 * (rix[3 * hex[d] + 1] - rix[3 * hex[d] + f]) -
                                                       rix[3 * (-2 * hex[d]) + f]) + rix[f];
                            const mm_t m = GM(row - top, (col - left) >> 1);
                            rix[1] = limf(val, m.min, m.max);
                        }
                } else {
                    int f = fcolx(P, row, ls);
                    const short *hm[2] = {P->allhex[1][row % 3][ls % 3], P->allhex[1][row % 3][(ls + coff) % 3]};
                    for (int col = ls, hi = 0; col < mcol - 2; col += coff, coff ^= 3, f = f ^ (coff & 2), hi ^= 1) {
                        const short *hex = hm[hi];
                        for (int d = 3; d < 6; d++) {
                            float *rix = RGBP(rgb, (d - 2) ^ 1, row - top, col - left);
                            float val = 0.33333333f * (rix[3 * (-2 * hex[d]) + 1] + 2 * (rix[3 * hex[d] + 1] - rix[3 * hex[d] + f]) -
                                                       rix[3 * (-2 * hex[d]) + f]) + rix[f];
                            const mm_t m = GM(row - top, (col - left) >> 1);
                            rix[1] = limf(val, m.min, m.max);
                        }
                    }
                }
            }
        }

        /* R and B at solitary green sites, :2184-2222 */
        {
            const int sgstart = (left - P->sgcol + 4) / 3 * 3 + P->sgcol;
            float color[3][6];
            for (int row = (top - P->sgrow + 4) / 3 * 3 + P->sgrow; row < mrow - 2; row += 3)
                for (int col = sgstart, h = fcolx(P, row, col + 1); col < mcol - 2; col += 3, h ^= 2) {
                    float *rix = RGBP(rgb, 0, row - top, col - left);
                    float diff[6] = {0.f};
                    for (int i = 1, d = 0; d < 6; d++, i ^= TS ^ 1, h ^= 2) {
                        for (int c = 0; c < 2; c++, h ^= 2) {
                            const int o = i << c;
                            float g = rix[1] + rix[1] - rix[3 * o + 1] - rix[3 * -o + 1];
                            color[h][d] = g + rix[3 * o + h] + rix[3 * -o + h];
                            if (d > 1)
                                diff[d] += sq(rix[3 * o + 1] - rix[3 * -o + 1] - rix[3 * o + h] + rix[3 * -o + h]) + sq(g);
                        }
                        if (d > 2 && (d & 1))
                            if (diff[d - 1] < diff[d])
                                for (int c = 0; c < 2; c++) color[c * 2][d] = color[c * 2][d - 1];
                        if ((d & 1) || d < 2) {
                            for (int c = 0; c < 2; c++) rix[c * 2] = 0.5f * color[c * 2][d];
                            rix += dirsz;
                        }
                    }
                }
        }

        /* R at B sites and B at R sites, :2224-2266 */
        for (int row = top + 3; row < mrow - 3; row++) {
            int ls = left + 3;
            for (; ls < mcol - 1; ls++)
                if (!isgreen(P, row, ls)) break;
            int coff = P->right_shift[row % 3] == 1 ? 3 : 1;
            const int c = ((row - P->sgrow) % 3) ? TS : 1;
            const int h = 3 * (c ^ TS ^ 1);
            if (coff == 3) {
                int f = 2 - fcolx(P, row, ls);
                for (int col = ls; col < mcol - 3; col += coff, f ^= 2) {
                    float *rix = RGBP(rgb, 0, row - top, col - left);
                    for (int d = 0; d < 4; d++, rix += dirsz) {
                        int i = d > 1 || ((d ^ c) & 1) ||
                                        ((fabsf(rix[1] - rix[3 * c + 1]) + fabsf(rix[1] - rix[3 * -c + 1])) <
                                         2.f * (fabsf(rix[1] - rix[3 * h + 1]) + fabsf(rix[1] - rix[3 * -h + 1])))
                                    ? c : h;
                        rix[f] = rix[1] + 0.5f * (rix[3 * i + f] + rix[3 * -i + f] - rix[3 * i + 1] - rix[3 * -i + 1]);
                    }
                }
            } else {
                coff = fcolx(P, row, ls + 1) == 1 ? 2 : 1;
                int f = 2 - fcolx(P, row, ls);
                for (int col = ls; col < mcol - 3; col += coff, coff ^= 3, f = f ^ (coff & 2)) {
                    float *rix = RGBP(rgb, 0, row - top, col - left);
                    for (int d = 0; d < 4; d++, rix += dirsz) {
                        int i = d > 1 || ((d ^ c) & 1) ||
                                        ((fabsf(rix[1] - rix[3 * c + 1]) + fabsf(rix[1] - rix[3 * -c + 1])) <
                                         2.f * (fabsf(rix[1] - rix[3 * h + 1]) + fabsf(rix[1] - rix[3 * -h + 1])))
                                    ? c : h;
                        rix[f] = rix[1] + 0.5f * (rix[3 * i + f] + rix[3 * -i + f] - rix[3 * i + 1] - rix[3 * -i + 1]);
                    }
                }
            }
        }

        /* R and B for 2x2 green blocks, :2268-2314 */
        {
            int topstart = top + 2;
            for (; topstart < mrow - 2; topstart++)
                if ((topstart - P->sgrow) % 3) break;
            int ls = left + 2;
            for (; ls < mcol - 2; ls++)
                if ((ls - P->sgcol) % 3) break;
            const int coffstart = 2 - (fcolx(P, topstart, ls + 1) & 1);
            for (int row = topstart; row < mrow - 2; row++) {
                if (!((row - P->sgrow) % 3)) continue;
                const short *hm[2] = {P->allhex[1][row % 3][ls % 3], P->allhex[1][row % 3][(ls + coffstart) % 3]};
                for (int col = ls, coff = coffstart, hi = 0; col < mcol - 2; col += coff, coff ^= 3, hi ^= 1) {
                    float *rix = RGBP(rgb, 0, row - top, col - left);
                    const short *hex = hm[hi];
                    for (int d = 0; d < ndir; d += 2, rix += dirsz) {
                        if (hex[d] + hex[d + 1]) {
                            float g = 3 * rix[1] - 2 * rix[3 * hex[d] + 1] - rix[3 * hex[d + 1] + 1];
                            for (int c = 0; c < 4; c += 2)
                                rix[c] = (g + 2 * rix[3 * hex[d] + c] + rix[3 * hex[d + 1] + c]) * 0.33333333f;
                        } else {
                            float g = 2 * rix[1] - rix[3 * hex[d] + 1] - rix[3 * hex[d + 1] + 1];
                            for (int c = 0; c < 4; c += 2)
                                rix[c] = (g + rix[3 * hex[d] + c] + rix[3 * hex[d + 1] + c]) * 0.5f;
                        }
                    }
                }
            }
        }
    }

    rgb = B->rgb_base;
    mrow -= top;
    mcol -= left;

    /* YPbPr and directional derivatives, :2346-2419 */
    const int LW = TS - 8, DW = TS - 10;
    float *yuv = B->lab;
    static const short dirs4[4] = {1, TS, TS + 1, TS - 1};
    for (int d = 0; d < ndir; d++) {
        for (int row = 4; row < mrow - 4; row++)
            for (int col = 4; col < mcol - 4; col++) {
                const float *p = RGBP(rgb, d, row, col);
                float y = 0.2627f * p[0] + 0.6780f * p[1] + 0.0593f * p[2];
                yuv[(size_t)(row - 4) * LW + col - 4] = y;
                yuv[(size_t)LW * LW + (size_t)(row - 4) * LW + col - 4] = (p[2] - y) * 0.56433f;
                yuv[(size_t)2 * LW * LW + (size_t)(row - 4) * LW + col - 4] = (p[0] - y) * 0.67815f;
            }
        int f = dirs4[d & 3];
        f = f == 1 ? 1 : f - 8;
        for (int row = 5; row < mrow - 5; row++)
            for (int col = 5; col < mcol - 5; col++) {
                const float *y = yuv + (size_t)(row - 4) * LW + col - 4;
                const float *u = y + (size_t)LW * LW;
                const float *v = u + (size_t)LW * LW;
                B->drv[(size_t)d * DW * DW + (size_t)(row - 5) * DW + col - 5] =
                    sq(2 * y[0] - y[f] - y[-f]) + sq(2 * u[0] - u[f] - u[-f]) + sq(2 * v[0] - v[f] - v[-f]);
            }
    }

    /* homogeneity maps, :2421-2495 */
#define DRV(d, r, c) B->drv[(size_t)(d) * DW * DW + (size_t)(r) * DW + (c)]
#define HOMO(d, r, c) B->homo[((size_t)(d) * TS + (r)) * TS + (c)]
#define HSUM(d, r, c) B->homosum[((size_t)(d) * TS + (r)) * TS + (c)]
    for (int row = 6; row < mrow - 6; row++)
        for (int col = 6; col < mcol - 6; col++) {
            float tr = DRV(0, row - 5, col - 5) < DRV(1, row - 5, col - 5) ? DRV(0, row - 5, col - 5) : DRV(1, row - 5, col - 5);
            for (int d = 2; d < ndir; d++) tr = DRV(d, row - 5, col - 5) < tr ? DRV(d, row - 5, col - 5) : tr;
            tr *= 8;
            for (int d = 0; d < ndir; d++) {
                unsigned char t = 0;
                for (int v = -1; v <= 1; v++)
                    for (int h = -1; h <= 1; h++) t += (DRV(d, row + v - 5, col + h - 5) <= tr ? 1 : 0);
                HOMO(d, row, col) = t;
            }
        }

    if (H - top < TS + 4) mrow = H - top + 2;
    if (W - left < TS + 4) mcol = W - left + 2;

    /* 5x5 sums of the homogeneity maps (running column sums), :2505-2559 */
    const int startcol = left < 8 ? left : 8;
    for (int d = 0; d < ndir; d++)
        for (int row = top < 8 ? top : 8; row < mrow - 8; row++) {
            int col = startcol;
            if (col < mcol - 8) {
                int v5[5] = {0};
                for (int v = -2; v <= 2; v++)
                    for (int h = -2; h <= 2; h++) v5[2 + h] += HOMO(d, row + v, col + h);
                int bs = v5[0] + v5[1] + v5[2] + v5[3] + v5[4];
                HSUM(d, row, col) = (unsigned char)bs;
                col++;
                for (int vo = 0; col < mcol - 8; col++, vo++) {
                    int cs = HOMO(d, row - 2, col + 2) + HOMO(d, row - 1, col + 2) + HOMO(d, row, col + 2) +
                             HOMO(d, row + 1, col + 2) + HOMO(d, row + 2, col + 2);
                    vo = vo == 5 ? 0 : vo;
                    bs -= v5[vo];
                    bs += cs;
                    v5[vo] = cs;
                    HSUM(d, row, col) = (unsigned char)bs;
                }
            }
        }
    for (int row = top < 8 ? top : 8; row < mrow - 8; row++)
        for (int col = startcol; col < mcol - 8; col++) {
            unsigned char m = HSUM(0, row, col);
            for (int d = 1; d < ndir; d++) m = m < HSUM(d, row, col) ? HSUM(d, row, col) : m;
            m -= m >> 3;
            B->homosummax[(size_t)row * TS + col] = m;
        }

    /* average of the most homogeneous directions, :2630-2670 */
    unsigned char hm[8] = {0};
    for (int row = top < 8 ? top : 8; row < mrow - 8; row++)
        for (int col = left < 8 ? left : 8; col < mcol - 8; col++) {
            for (int d = 0; d < 4; d++) hm[d] = HSUM(d, row, col);
            for (int d = 4; d < ndir; d++) {
                hm[d] = HSUM(d, row, col);
                if (hm[d - 4] < hm[d]) hm[d - 4] = 0;
                else if (hm[d - 4] > hm[d]) hm[d] = 0;
            }
            float avg[4] = {0.f, 0.f, 0.f, 0.f};
            const unsigned char mx = B->homosummax[(size_t)row * TS + col];
            for (int d = 0; d < ndir; d++)
                if (hm[d] >= mx) {
                    const float *p = RGBP(rgb, d, row, col);
                    for (int c = 0; c < 3; c++) avg[c] += p[c];
                    avg[3] += 1.f;
                }
            float *o = P->out + ((size_t)(row + top) * W + col + left) * 3;
            float a0 = avg[0] / avg[3], a1 = avg[1] / avg[3], a2 = avg[2] / avg[3];
            o[0] = (0.f < a0) ? a0 : 0.f;
            o[1] = (0.f < a1) ? a1 : 0.f;
            o[2] = (0.f < a2) ? a2 : 0.f;
        }
#undef DRV
#undef HOMO
#undef HSUM
#undef GM
}

static size_t xbuf_bytes(void) {
    return sizeof(float) * 8 * TS * TS * 3 + sizeof(float) * 3 * (TS - 8) * (TS - 8) +
           sizeof(float) * 8 * (TS - 10) * (TS - 10) + 8 * TS * TS * 2 + TS * TS + sizeof(mm_t) * TS * (TS / 2) + 256;
}

static void xbuf_init(xbuf_t *B, char *mem) {
    B->rgb_base = (float *)mem;
    B->lab = B->rgb_base + 8 * TS * TS * 3;
    B->drv = B->lab + 3 * (TS - 8) * (TS - 8);
    B->gmm = (mm_t *)(B->drv + 8 * (TS - 10) * (TS - 10));
    B->homo = (unsigned char *)(B->gmm + TS * (TS / 2));
    B->homosum = B->homo + 8 * TS * TS;
    B->homosummax = B->homosum + 8 * TS * TS;
}

/* hexagon tables and solitary-green position, :1887-1925, :1933-1943 */
static void xt_setup(xt_t *P) {
    static const short orth[12] = {1, 0, 0, 1, -1, 0, 0, -1, 1, 0, 0, 1};
    static const short patt[2][16] = {{0, 1, 0, -1, 2, 0, -1, 0, 1, 1, 1, -1, 0, 0, 0, 0},
                                      {0, 1, 0, -2, 1, 0, -2, 0, 1, 1, -2, -2, 1, -1, -1, 1}};
    P->sgrow = P->sgcol = 0;
    memset(P->allhex, 0, sizeof P->allhex);
    for (int row = 0; row < 3; row++)
        for (int col = 0; col < 3; col++) {
            int gint = isgreen(P, row, col), ng = 0;
            for (int d = 0; d < 10; d += 2) {
                if (isgreen(P, row + orth[d] + 6, col + orth[d + 2] + 6)) ng = 0;
                else ng++;
                if (ng == 4) { P->sgrow = row; P->sgcol = col; }
                if (ng == gint + 1)
                    for (int c = 0; c < 8; c++) {
                        int v = orth[d] * patt[gint][c * 2] + orth[d + 1] * patt[gint][c * 2 + 1];
                        int h = orth[d + 2] * patt[gint][c * 2] + orth[d + 3] * patt[gint][c * 2 + 1];
                        P->allhex[0][row][col][c ^ (gint * 2 & d)] = (short)(h + v * P->W);
                        P->allhex[1][row][col][c ^ (gint * 2 & d)] = (short)(h + v * TS);
                    }
            }
        }
    for (int row = 0; row < 3; row++) {
        int gc = 0;
        for (int col = 0; col < 3; col++) gc += isgreen(P, row, col);
        P->right_shift[row] = gc == 2;
    }
}

int oracle_xtrans3(const float *raw3, float *rgb3, int w, int h, const signed char *xtrans36,
                   float maximum_value, int tile_mode) {
    if (!raw3 || !rgb3 || !xtrans36 || w <= 0 || h <= 0) return -1;
    xt_t P;
    P.raw = raw3;
    P.out = rgb3;
    P.W = w;
    P.H = h;
    P.maxv = maximum_value;
    for (int i = 0; i < 36; i++) P.xt[i / 6][i % 6] = xtrans36[i];
    xt_setup(&P);
    const size_t bytes = xbuf_bytes();
    if (tile_mode == ORACLE_TILES_SEQUENTIAL) {
        char *mem = calloc(1, bytes);
        xbuf_t B;
        xbuf_init(&B, mem);
        for (int top = 3; top < h - 19; top += TS - 16)
            for (int left = 3; left < w - 19; left += TS - 16) xtrans_tile(&P, &B, top, left);
        free(mem);
    } else {
        int ntr = 0, ntc = 0;
        for (int top = 3; top < h - 19; top += TS - 16) ntr++;
        for (int left = 3; left < w - 19; left += TS - 16) ntc++;
#pragma omp parallel
        {
            char *mem = malloc(bytes);
            xbuf_t B;
            xbuf_init(&B, mem);
#pragma omp for schedule(dynamic)
            for (int t = 0; t < ntr * ntc; t++) {
                memset(mem, 0, bytes);
                xtrans_tile(&P, &B, 3 + (t / ntc) * (TS - 16), 3 + (t % ntc) * (TS - 16));
            }
            free(mem);
        }
    }
    border8(&P, 8);
    return 0;
}
