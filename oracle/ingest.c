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
 */
#include <stdint.h>
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
