/*
 * TEST INFRASTRUCTURE ONLY — CPU restatement of the float highlight recovery.
 * Never called by the product.
 *
 *  oracle_recover_highlights  core/libraw_wrapper.cpp:1095-1383
 *                             (Impl::recover_highlights), called by
 *                             postprocess for highlight_mode > 2 (:1659-1683)
 *                             and by recover_highlights_numpy (:3063-3095)
 *
 * The reference function lives in the LibRaw-bound wrapper and is not built
 * here.  The restatement keeps its statement order: one raster pass for the
 * soft mask and the R/G, B/G ratio sums (serial float sums), the highlight
 * rewrite in raster order, the 8-neighbour average of blown pixels (all
 * averages read the state before any of them is written), then the LoCA
 * green lift over the highlight pixels in raster order, in place: a pixel
 * reads the already-updated values of the pixels before it.
 *
 * loca_order: ORACLE_LOCA_RASTER (0) is the reference; ORACLE_LOCA_SNAPSHOT
 * (1) reads every neighbour from the image as it was before the LoCA pass (a
 * parallel reading that the reference does not do), used by the tests to show
 * that their images separate the two; ORACLE_LOCA_OFF (2) stops before the
 * LoCA pass, so tests can compare with the state it starts from.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static inline float maxf_std(float a, float b) { return (a < b) ? b : a; } /* std::max */
static inline float minf_std(float a, float b) { return (b < a) ? b : a; } /* std::min */
static inline float clampf_std(float v, float lo, float hi) {               /* std::clamp */
    return v < lo ? lo : (hi < v ? hi : v);
}
static inline float max3(const float *p) { return maxf_std(p[0], maxf_std(p[1], p[2])); }

int oracle_recover_highlights(float *rgb3, int w, int h, float sat, float *mask_out,
                              const float *wb_rgb, float *max_val_out, long long *counts3,
                              int loca_order) {
    if (!rgb3 || w <= 0 || h <= 0) return -1;
    const size_t width = (size_t)w, height = (size_t)h, n = width * height;

    /* mask edges and WB undo (:1135-1147) */
    const float mask_t_hi = 1.00f, mask_t_lo = 0.95f;
    const float mask_inv = 1.f / maxf_std(1e-6f, mask_t_hi - mask_t_lo);
    const float chroma_t_lo = wb_rgb ? 0.30f : 0.50f, chroma_t_hi = wb_rgb ? 0.60f : 0.70f;
    const float chroma_inv = 1.f / maxf_std(1e-6f, chroma_t_hi - chroma_t_lo);
    const float inv_wb_r = wb_rgb ? 1.f / maxf_std(1e-6f, wb_rgb[0]) : 1.f;
    const float inv_wb_g = wb_rgb ? 1.f / maxf_std(1e-6f, wb_rgb[1]) : 1.f;
    const float inv_wb_b = wb_rgb ? 1.f / maxf_std(1e-6f, wb_rgb[2]) : 1.f;

    unsigned char *hl = calloc(n, 1); /* membership of the highlight list, raster order */
    size_t *white = malloc(n * sizeof(size_t));
    if (!hl || !white) {
        free(hl);
        free(white);
        return -2;
    }

    /* pass 1: mask and ratio sums (:1149-1190) */
    float grf = 0.f, gbf = 0.f, count = 0.f;
    size_t n_highlight = 0;
    for (size_t idx = 0; idx < n; ++idx) {
        const float *px = rgb3 + 3 * idx;
        if (mask_out) {
            float t = (px[1] - mask_t_lo) * mask_inv;
            if (t < 0.f) t = 0.f;
            else if (t > 1.f) t = 1.f;
            const float w_g = t * t * (3.f - 2.f * t);
            const float r_n = px[0] * inv_wb_r, g_n = px[1] * inv_wb_g, b_n = px[2] * inv_wb_b;
            const float mx = maxf_std(r_n, maxf_std(g_n, b_n));
            const float mn = minf_std(r_n, minf_std(g_n, b_n));
            const float chroma = (mx > 1e-6f) ? (mx - mn) / mx : 0.f;
            float ct = (chroma - chroma_t_lo) * chroma_inv;
            if (ct < 0.f) ct = 0.f;
            else if (ct > 1.f) ct = 1.f;
            const float neutrality = 1.f - ct * ct * (3.f - 2.f * ct);
            mask_out[idx] = w_g * neutrality;
        }
        if (px[0] >= sat && px[2] >= sat) {
            hl[idx] = 1;
            n_highlight++;
            if (px[0] < 0.95f && px[1] >= sat && px[1] < 0.95f && px[2] < 0.95f) {
                grf += px[0] / px[1];
                gbf += px[2] / px[1];
                count += 1.f;
            }
        }
    }
    if (count > 0.f) {
        grf /= count;
        gbf /= count;
    } else {
        grf = 1.f;
        gbf = 1.f;
    }

    /* pass 2: highlight rewrite (:1192-1242) */
    size_t n_white = 0;
    float max_val = 0.f;
    for (size_t idx = 0; idx < n; ++idx) {
        if (!hl[idx]) continue;
        float *px = rgb3 + 3 * idx;
        for (int i = 0; i < 3; ++i)
            if (px[i] > max_val) max_val = px[i];
        px[1] = (px[0] / grf + px[2] / gbf) * 0.5f;
        const float sp = (px[0] < px[2]) ? px[0] : px[2];
        const float sl = (minf_std(sp, 1.f) - sat) / (sat);
        px[0] = px[0] * (1.f - sl) + (px[1] * grf) * sl;
        px[2] = px[2] * (1.f - sl) + (px[1] * gbf) * sl;
        px[1] = (px[0] / grf + px[2] / gbf) * 0.5f;
        if (px[1] >= 0.95f) white[n_white++] = idx;
    }

    /* pass 3: blown pixels take the mean of their 8 neighbours, all means
     * computed before any is written (:1244-1285) */
    float *avg = malloc((n_white ? n_white : 1) * 3 * sizeof(float));
    size_t *avg_idx = malloc((n_white ? n_white : 1) * sizeof(size_t));
    if (!avg || !avg_idx) {
        free(hl);
        free(white);
        free(avg);
        free(avg_idx);
        return -2;
    }
    size_t n_avg = 0;
    for (size_t k = 0; k < n_white; ++k) {
        const size_t idx = white[k], y = idx / width, x = idx % width;
        if (x <= 0 || x >= width - 1 || y <= 0 || y >= height - 1) continue;
        const float *I = rgb3;
        for (int c = 0; c < 3; ++c)
            avg[3 * n_avg + c] =
                (I[3 * (idx - width - 1) + c] + I[3 * (idx - width) + c] + I[3 * (idx - width + 1) + c] +
                 I[3 * (idx - 1) + c] + I[3 * (idx + 1) + c] + I[3 * (idx + width - 1) + c] +
                 I[3 * (idx + width) + c] + I[3 * (idx + width + 1) + c]) *
                (1.f / 8.f);
        avg_idx[n_avg++] = idx;
    }
    for (size_t k = 0; k < n_avg; ++k)
        for (int c = 0; c < 3; ++c) rgb3[3 * avg_idx[k] + c] = avg[3 * k + c];
    free(avg);
    free(avg_idx);

    /* pass 4: LoCA green lift over the highlight list, in raster order, in
     * place (:1287-1374) */
    if (loca_order == ORACLE_LOCA_OFF) {
        free(hl);
        free(white);
        if (max_val_out) *max_val_out = max_val;
        if (counts3) {
            counts3[0] = (long long)n_highlight;
            counts3[1] = (long long)n_white;
            counts3[2] = 0;
        }
        return 0;
    }
    float *snap = NULL;
    if (loca_order == ORACLE_LOCA_SNAPSHOT) {
        snap = malloc(n * 3 * sizeof(float));
        if (!snap) {
            free(hl);
            free(white);
            return -2;
        }
        memcpy(snap, rgb3, n * 3 * sizeof(float));
    }
    const float *rd = snap ? snap : rgb3;
    long long n_loca = 0;
    for (size_t idx = 0; idx < n; ++idx) {
        if (!hl[idx]) continue;
        const size_t y = idx / width, x = idx % width;
        if (x <= 1 || x >= width - 2 || y <= 1 || y >= height - 2) continue;
        float *pw = rgb3 + 3 * idx;
        float p[3] = {rd[3 * idx], rd[3 * idx + 1], rd[3 * idx + 2]};
        const float l = max3(p);
        if (l < sat * 0.90f) continue;
        float local_max = 0.0f, local_min = FLT_MAX;
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const size_t nidx = (size_t)((int)y + dy) * width + (size_t)((int)x + dx);
                const float nl = max3(rd + 3 * nidx);
                local_max = maxf_std(local_max, nl);
                local_min = minf_std(local_min, nl);
            }
        if ((local_max - local_min) < 0.015f) continue;
        const float rb = 0.5f * (p[0] + p[2]);
        if (rb <= p[1] * 1.04f) continue;
        double ratio_sum = 0.0, ratio_w = 0.0;
        for (int dy = -2; dy <= 2; ++dy)
            for (int dx = -2; dx <= 2; ++dx) {
                const size_t nidx = (size_t)((int)y + dy) * width + (size_t)((int)x + dx);
                const float *np = rd + 3 * nidx;
                const float nrb = 0.5f * (np[0] + np[2]);
                const float nl = max3(np);
                if (nl < sat * 0.75f || nrb <= 1e-6f) continue;
                const float ratio = np[1] / nrb;
                if (ratio < 0.70f || ratio > 1.35f) continue;
                const double wt = 1.0 / (double)(1 + abs(dx) + abs(dy));
                ratio_sum += ratio * wt;
                ratio_w += wt;
            }
        const float local_ratio = (float)(ratio_w > 0.0 ? ratio_sum / ratio_w : 1.0);
        const float desired_g = rb * clampf_std(local_ratio, 0.95f, 1.08f);
        const float g_missing = desired_g - p[1];
        if (g_missing <= 0.0f) continue;
        const float sat_w =
            clampf_std((l - sat) / maxf_std(1e-6f, maxf_std(1.0f, max_val) - sat), 0.0f, 1.0f);
        const float edge_w = clampf_std((local_max - local_min) / 0.08f, 0.0f, 1.0f);
        const float alpha = clampf_std(0.30f + 0.55f * sat_w * edge_w, 0.0f, 0.90f);
        const float g_boost = alpha * minf_std(g_missing, (rb - p[1]) * 0.90f);
        p[1] += g_boost;
        const float rb_residual = maxf_std(0.0f, 0.5f * (p[0] + p[2]) - p[1]);
        if (rb_residual > 0.0f) {
            const float trim = minf_std(rb_residual * alpha * 0.45f, 0.5f * (p[0] + p[2]) * 0.12f);
            p[0] = maxf_std(0.0f, p[0] - trim);
            p[2] = maxf_std(0.0f, p[2] - trim);
        }
        p[1] = maxf_std(0.0f, p[1]);
        pw[0] = p[0];
        pw[1] = p[1];
        pw[2] = p[2];
        n_loca++;
    }
    free(snap);
    if (max_val_out) *max_val_out = max_val;
    if (counts3) {
        counts3[0] = (long long)n_highlight;
        counts3[1] = (long long)n_white;
        counts3[2] = n_loca;
    }
    free(hl);
    free(white);
    return 0;
}
