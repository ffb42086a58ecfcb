/*
 * TEST INFRASTRUCTURE ONLY — CPU restatement of the integer preprocessing in
 * front of the demosaics.  Never called by the product.
 *
 *  oracle_black_level   apply_black_level_correction, core/libraw_wrapper.cpp:401-491:
 *                       every one of the 4 channels of every pixel,
 *                       val = clamp(val - pos_black - cblack[c], 0, maximum) when a
 *                       positional table is present, else clamp(val - cblack[c]) when
 *                       any cblack[c] is set, else clamp(val - black)
 *  oracle_ingest_u16    the same on a one-sample-per-site CFA plane, followed by the
 *                       white balance of apply_wb_bayer / apply_wb_xtrans
 *                       (core/cpu_accelerator.cpp:2813-2869) into a float CFA plane,
 *                       and the data maximum of adjust_maximum
 *                       (core/libraw_wrapper.cpp:1010-1089), which reads each pixel's
 *                       channel through fcol_bayer even for X-Trans (:1051)
 *  oracle_adjust_maximum  the maximum rule of :1070-1077
 *  oracle_green_matching  apply_green_matching, :496-585
 *  oracle_remove_zeroes   remove_zeroes, :150-212
 *  oracle_auto_wb         the automatic multipliers of scale_colors, :621-663
 *  oracle_scale_aber      the legacy radial aberration resampling of scale_colors, :833-880
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

int oracle_black_level(uint16_t *img4, int w, int h, const int *cblack4, const int *pos_table,
                       int pos_rows, int pos_cols, int black, int maximum) {
    const int positional = pos_table && pos_rows && pos_cols;
    const int any = cblack4[0] || cblack4[1] || cblack4[2] || cblack4[3];
    for (size_t q = 0; q < (size_t)w * h; q++)
        for (int c = 0; c < 4; c++) {
            int v = img4[4 * q + c];
            if (positional) {
                const int row = (int)(q / (size_t)w), col = (int)(q % (size_t)w);
                v -= pos_table[(row % pos_rows) * pos_cols + (col % pos_cols)];
                v -= cblack4[c];
            } else if (any) {
                v -= cblack4[c];
            } else {
                v -= black;
            }
            img4[4 * q + c] = (uint16_t)clampi(v, 0, maximum);
        }
    return 0;
}

static inline unsigned bayer_native(uint32_t f, unsigned r, unsigned c) {
    return f >> ((((r << 1) & 14) | (c & 1)) << 1) & 3;
}

int oracle_ingest_u16(const uint16_t *raw, float *cfa, int w, int h, uint32_t filters,
                      const signed char *xtrans36, const int *cblack4, int black, int maximum,
                      const float *wb4, float *data_max) {
    const int any = cblack4[0] || cblack4[1] || cblack4[2] || cblack4[3];
    float dmax = 0.f;
    for (int row = 0; row < h; row++)
        for (int col = 0; col < w; col++) {
            const size_t i = (size_t)row * w + col;
            unsigned c4, c3;
            if (filters == 9) {
                c4 = c3 = (unsigned)xtrans36[((row + 6) % 6) * 6 + (col + 6) % 6];
            } else {
                c4 = bayer_native(filters, row, col);
                c3 = c4 == 3 ? 1 : c4;
            }
            int v = raw[i];
            v -= any ? cblack4[c4] : black;
            v = clampi(v, 0, maximum);
            const float f = (float)(uint16_t)v * wb4[c4];
            cfa[i] = f;
            /* adjust_maximum reads channel fcol_bayer(row, col, filters) of the
               white-balanced 3-channel buffer; only the native channel is non-zero */
            unsigned cq = bayer_native(filters, row, col);
            cq = cq == 3 ? 1 : cq;
            const float val = (cq == c3) ? f : 0.f;
            if (val > dmax) dmax = val;
        }
    if (data_max) *data_max = dmax;
    return 0;
}

float oracle_adjust_maximum(float data_max, float maximum, float threshold) {
    if (threshold < 0.00001f) return maximum;
    float thr = threshold > 0.99999f ? 0.75f : threshold;
    if (data_max > 0 && data_max < maximum && data_max > maximum * thr) return data_max;
    return maximum;
}

/* apply_green_matching, core/libraw_wrapper.cpp:496-585: for Bayer data, every
 * site (j, i), j = 3, 5, .. < h-3, i = 2, 4, .. < w-3, whose channel-3 sample is
 * below 0.95 maximum and whose four diagonal channel-1 and four axial channel-3
 * neighbours are flat (mean absolute pair difference below 0.01 maximum) gets
 * its channel-3 sample scaled by mean(ch1) / mean(ch3); neighbours are read
 * from the unmodified image. */
int oracle_green_matching(uint16_t *img4, int w, int h, uint32_t filters, int maximum) {
    if (filters == 9) return 0;
    const size_t n = (size_t)w * h * 4;
    uint16_t *img = (uint16_t *)malloc(n * sizeof(uint16_t));
    if (!img) return -1;
    memcpy(img, img4, n * sizeof(uint16_t));
    const float thr = 0.01f;
#define IMG(r, c, ch) img[((size_t)(r) * w + (c)) * 4 + (ch)]
    for (int j = 3; j < h - 3; j += 2)
        for (int i = 2; i < w - 3; i += 2) {
            const int o1_1 = IMG(j - 1, i - 1, 1), o1_2 = IMG(j - 1, i + 1, 1);
            const int o1_3 = IMG(j + 1, i - 1, 1), o1_4 = IMG(j + 1, i + 1, 1);
            const int o2_1 = IMG(j - 2, i, 3), o2_2 = IMG(j + 2, i, 3);
            const int o2_3 = IMG(j, i - 2, 3), o2_4 = IMG(j, i + 2, 3);
            const double m1 = (o1_1 + o1_2 + o1_3 + o1_4) / 4.0;
            const double m2 = (o2_1 + o2_2 + o2_3 + o2_4) / 4.0;
            const double c1 = (abs(o1_1 - o1_2) + abs(o1_1 - o1_3) + abs(o1_1 - o1_4) +
                               abs(o1_2 - o1_3) + abs(o1_3 - o1_4) + abs(o1_2 - o1_4)) / 6.0;
            const double c2 = (abs(o2_1 - o2_2) + abs(o2_1 - o2_3) + abs(o2_1 - o2_4) +
                               abs(o2_2 - o2_3) + abs(o2_3 - o2_4) + abs(o2_2 - o2_4)) / 6.0;
            if ((IMG(j, i, 3) < maximum * 0.95) && (c1 < maximum * thr) && (c2 < maximum * thr) &&
                (m2 > 0.1)) {
                const float correction = (float)(img4[((size_t)j * w + i) * 4 + 3] * m1 / m2);
                img4[((size_t)j * w + i) * 4 + 3] =
                    correction > maximum ? (uint16_t)maximum : (uint16_t)correction;
            }
        }
#undef IMG
    free(img);
    return 0;
}

/* remove_zeroes, core/libraw_wrapper.cpp:150-212: in raster order, a zero
 * sample at its native channel becomes the integer mean of the non-zero
 * samples of the same native channel in its 5x5 window (clipped at the frame);
 * samples replaced earlier count with their new value. */
int oracle_remove_zeroes(uint16_t *img4, int w, int h, uint32_t filters) {
    for (int row = 0; row < h; row++)
        for (int col = 0; col < w; col++) {
            const unsigned ch = bayer_native(filters, row, col);
            uint16_t *cur = &img4[((size_t)row * w + col) * 4 + ch];
            if (*cur != 0) continue;
            unsigned sum = 0, cnt = 0;
            for (int r = row - 2; r <= row + 2; r++)
                for (int c = col - 2; c <= col + 2; c++) {
                    if (r < 0 || r >= h || c < 0 || c >= w) continue;
                    const unsigned nch = bayer_native(filters, r, c);
                    const uint16_t v = img4[((size_t)r * w + c) * 4 + nch];
                    if (nch == ch && v != 0) {
                        sum += v;
                        cnt++;
                    }
                }
            if (cnt > 0) *cur = (uint16_t)(sum / cnt);
        }
    return 0;
}

/* Automatic white balance of scale_colors, core/libraw_wrapper.cpp:621-663
 * (p = 1): the mean of every native channel (3 folded into 1) over the whole
 * image; pre_mul = (mean G / mean R, 1, mean G / mean B, 1).  Sums of integers
 * in double are exact, so their order does not matter. */
int oracle_auto_wb(const uint16_t *img4, int w, int h, uint32_t filters, const signed char *xtrans36,
                   float *pre_mul4) {
    double norm[4] = {0, 0, 0, 0}, count[4] = {0, 0, 0, 0};
    for (int row = 0; row < h; row++)
        for (int col = 0; col < w; col++) {
            const unsigned cc = filters == 9 ? (unsigned)xtrans36[((row + 6) % 6) * 6 + (col + 6) % 6]
                                             : bayer_native(filters, row, col);
            norm[cc] += pow(img4[((size_t)row * w + col) * 4 + cc], 1.0);
            count[cc] += 1;
        }
    norm[1] += norm[3];
    count[1] += count[3];
    for (int c = 0; c < 3; c++) norm[c] = pow(norm[c] / count[c], 1.0 / 1.0);
    pre_mul4[0] = (float)(norm[1] / norm[0]);
    pre_mul4[1] = 1.0f;
    pre_mul4[2] = (float)(norm[1] / norm[2]);
    pre_mul4[3] = 1.0f;
    return 0;
}

/* float -> unsigned as the reference's x86-64 build converts it: the value is
 * truncated to a 64-bit integer and its low 32 bits are kept, so sources left
 * of / above the frame by a pixel or more become huge and are skipped. */
static unsigned f2u_x86(float f) { return (unsigned)(long long)f; }

/* Legacy radial chromatic aberration of scale_colors, :833-880: channels 0 and
 * 2 resampled bilinearly at (x - cx) * aber[c] + cx, (y - cy) * aber[c] + cy
 * from a copy of the channel; samples whose source cell leaves the frame keep
 * their value. */
int oracle_scale_aber(uint16_t *img4, int w, int h, const float *aber4) {
    if (w <= 1 || h <= 1) return 0;
    const size_t size = (size_t)w * h;
    uint16_t *plane = (uint16_t *)malloc(size * sizeof(uint16_t));
    if (!plane) return -1;
    const float cy = (float)h * 0.5f, cx = (float)w * 0.5f;
    for (int c = 0; c < 4; c += 2) {
        const float scale = aber4[c];
        if (fabsf(scale - 1.0f) < 1e-6f || scale <= 0.0f) continue;
        for (size_t i = 0; i < size; i++) plane[i] = img4[i * 4 + c];
        for (int row = 0; row < h; row++)
            for (int col = 0; col < w; col++) {
                const float sc = ((float)col - cx) * scale + cx;
                const float sr = ((float)row - cy) * scale + cy;
                const unsigned ur = f2u_x86(sr);
                if (ur >= (unsigned)h - 1) continue;
                const float fr = sr - (float)ur;
                const unsigned uc = f2u_x86(sc);
                if (uc >= (unsigned)w - 1) continue;
                const float fc = sc - (float)uc;
                const uint16_t *p = plane + (size_t)ur * w + uc;
                const float v = (p[0] * (1.0f - fc) + p[1] * fc) * (1.0f - fr) +
                                (p[w] * (1.0f - fc) + p[w + 1] * fc) * fr;
                const float lo = (0.0f < v) ? v : 0.0f;            /* std::max(0.0f, v) */
                const float cl = (lo < 65535.0f) ? lo : 65535.0f;   /* std::min(65535.0f, lo) */
                img4[((size_t)row * w + col) * 4 + c] = (uint16_t)cl;
            }
    }
    free(plane);
    return 0;
}
