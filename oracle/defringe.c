/*
 * TEST INFRASTRUCTURE ONLY — CPU restatement of the reference defringe
 * (core/cpu_accelerator.cpp:3939-5064) with its smoothstep helper (:3322-3325).
 * Never called by the product.
 *
 * Structure: global max of green -> guide/log-ratio planes (:3961-3991);
 * Sobel magnitude and its global max (:3996-4027); separable Gaussian blurs
 * with clamp-to-edge, horizontal then vertical, taps summed k = -ks..ks
 * (:4033-4081); support seed (:4162-4248) and its blur; 8-direction line link
 * (:4296-4323); solo seeds and their blurs (:4325-4376); per-pixel main
 * correction and solo cleanup (:4377-4987).  In place (in == out) the solo
 * cleanup reads the pixel's already-corrected values, out of place the input
 * values, as in the reference.  Transcendentals are glibc's expf/logf/sqrtf,
 * as the reference's std::exp/std::log/std::sqrt on float.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static inline float maxf_(float a, float b) { return (a < b) ? b : a; }   /* std::max */
static inline float minf_(float a, float b) { return (b < a) ? b : a; }   /* std::min */
static inline float clampf_(float v, float lo, float hi) { return v < lo ? lo : (hi < v ? hi : v); }
static inline float sstep(float e0, float e1, float x) {
    float t = clampf_((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

typedef struct {
    int W, H, ks;
    float *kernel;
    float *tmp;
} blur_t;

/* :4055-4081 */
static void gblur(const blur_t *B, const float *in, float *out) {
    const int W = B->W, H = B->H, ks = B->ks;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            float acc = 0.f;
            for (int k = -ks; k <= ks; k++) {
                int xx = x + k < 0 ? 0 : (x + k > W - 1 ? W - 1 : x + k);
                acc += B->kernel[k + ks] * in[(size_t)y * W + xx];
            }
            B->tmp[(size_t)y * W + x] = acc;
        }
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            float acc = 0.f;
            for (int k = -ks; k <= ks; k++) {
                int yy = y + k < 0 ? 0 : (y + k > H - 1 ? H - 1 : y + k);
                acc += B->kernel[k + ks] * B->tmp[(size_t)yy * W + x];
            }
            out[(size_t)y * W + x] = acc;
        }
}

int oracle_defringe(const float *src, float *dst, int W, int H, float radius, float edge_threshold,
                    float chroma_threshold, float strength, int enable_green,
                    float green_strength_scale) {
    if (!src || !dst || W <= 0 || H <= 0) return -1;
    green_strength_scale = clampf_(green_strength_scale, 0.f, 1.f);
    const size_t N = (size_t)W * H;
    const int inplace = (src == dst);

    /* step 1 */
    float max_guide = 0.f;
    for (size_t i = 0; i < N; i++) {
        float g = maxf_(0.f, src[3 * i + 1]);
        if (g > max_guide) max_guide = g;
    }
    max_guide = maxf_(max_guide, 1e-6f);
    const float ratio_eps = maxf_(max_guide * 1e-4f, 1e-6f);
    float *guide_linear = malloc(N * sizeof(float)), *guide_signal = malloc(N * sizeof(float));
    float *log_rg = malloc(N * sizeof(float)), *log_bg = malloc(N * sizeof(float));
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; i++) {
        const float r = src[3 * i], g = src[3 * i + 1], b = src[3 * i + 2];
        const float gg = maxf_(g, 0.f);
        guide_linear[i] = gg;
        guide_signal[i] = sqrtf(clampf_(gg / max_guide, 0.f, 1.f));
        log_rg[i] = logf((maxf_(r, 0.f) + ratio_eps) / (gg + ratio_eps));
        log_bg[i] = logf((maxf_(b, 0.f) + ratio_eps) / (gg + ratio_eps));
        if (!inplace) {
            dst[3 * i] = r;
            dst[3 * i + 1] = g;
            dst[3 * i + 2] = b;
        }
    }

    /* step 2: Sobel */
    float *edge = calloc(N, sizeof(float)), *edge_raw = calloc(N, sizeof(float));
    {
        float max_edge = 0.f;
        const int ly = H > 1 ? H - 1 : 1, lx = W > 1 ? W - 1 : 1;
        const float *s = guide_signal;
        for (int y = 1; y < ly; y++)
            for (int x = 1; x < lx; x++) {
                size_t u = (size_t)(y - 1) * W, m = (size_t)y * W, d = (size_t)(y + 1) * W;
                float gx = -s[u + x - 1] + s[u + x + 1] - 2.f * s[m + x - 1] + 2.f * s[m + x + 1] -
                           s[d + x - 1] + s[d + x + 1];
                float gy = -s[u + x - 1] - 2.f * s[u + x] - s[u + x + 1] + s[d + x - 1] +
                           2.f * s[d + x] + s[d + x + 1];
                float mag = sqrtf(gx * gx + gy * gy);
                edge[m + x] = mag;
                edge_raw[m + x] = mag;
                if (mag > max_edge) max_edge = mag;
            }
        float nf = maxf_(max_edge, 1e-6f);
        if (nf > 0.f) {
            float inv = 1.f / nf;
            for (size_t i = 0; i < N; i++) edge[i] *= inv;
        }
    }

    /* step 3: Gaussian taps, :4033-4048 */
    const float sigma = maxf_(radius / 3.f, 0.5f);
    const int ks = (int)ceilf(3.f * sigma);
    blur_t B = {W, H, ks, malloc((2 * ks + 1) * sizeof(float)), calloc(N, sizeof(float))};
    {
        float ksum = 0.f;
        for (int k = -ks; k <= ks; k++) {
            float v = expf(-0.5f * (float)(k * k) / (sigma * sigma));
            B.kernel[k + ks] = v;
            ksum += v;
        }
        for (int k = 0; k < 2 * ks + 1; k++) B.kernel[k] /= ksum;
    }
    float *blur_guide = malloc(N * sizeof(float)), *blur_rg = malloc(N * sizeof(float)), *blur_bg = malloc(N * sizeof(float));
    float *blur_edge = malloc(N * sizeof(float)), *blur_edge_raw = malloc(N * sizeof(float));
    gblur(&B, guide_linear, blur_guide);
    gblur(&B, log_rg, blur_rg);
    gblur(&B, log_bg, blur_bg);
    gblur(&B, edge, blur_edge);
    gblur(&B, edge_raw, blur_edge_raw);

    const float sensitivity = clampf_(sqrtf(maxf_(strength, 1.0f)), 1.0f, 2.5f);
    const float isens = 1.0f / sensitivity;
    const float mask_cutoff = 0.02f * isens;
    const float scene_gate = sstep(0.55f, 0.75f, max_guide);

    /* support seed, :4162-4248 */
    float *fs = calloc(N, sizeof(float)), *bfs = calloc(N, sizeof(float));
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; i++) {
        const float rg = log_rg[i], bg = log_bg[i];
        const float dr = rg - blur_rg[i], db = bg - blur_bg[i];
        const float r0 = maxf_(src[3 * i], 0.f), g0 = maxf_(src[3 * i + 1], 0.f), b0 = maxf_(src[3 * i + 2], 0.f);
        const float purple_excess = minf_(maxf_(0.f, dr), maxf_(0.f, db));
        const float denom = maxf_(blur_guide[i], 0.05f);
        const float g_over_min = maxf_(0.f, g0 - minf_(r0, b0)) / denom;
        const float green_dom = sstep(0.010f, 0.060f, g_over_min);
        const float green_excess = maxf_(maxf_(0.f, -dr), maxf_(0.f, -db)) * green_dom;
        const int is_purple = purple_excess >= green_excess;
        const float excess = is_purple ? purple_excess : green_excess;
        const float min_dir = is_purple ? maxf_(0.f, minf_(rg, bg)) : maxf_(0.f, minf_(-rg, -bg));
        const float max_dir = is_purple ? maxf_(1e-4f, maxf_(rg, bg)) : maxf_(1e-4f, maxf_(-rg, -bg));
        const float balance = min_dir / max_dir;
        const float c_edge = sstep(edge_threshold * 0.35f * isens, edge_threshold * 1.50f * isens, blur_edge[i]) *
                             sstep(0.55f * isens, 1.35f * isens, blur_edge_raw[i]);
        const float c_bright = sstep(0.24f * isens, 0.52f * isens, blur_guide[i]);
        const float c_chroma = sstep(chroma_threshold * 0.45f * isens, chroma_threshold * 1.60f * isens, excess);
        const float red_strength = maxf_(0.f, r0 - maxf_(g0, b0)) / maxf_(r0, 1e-4f);
        const float rb_balance = b0 / maxf_(r0, 1e-4f);
        const float blue_lift = maxf_(0.f, b0 - g0) / maxf_(r0 - g0, 1e-4f);
        const float red_purple = maxf_(sstep(0.55f, 0.78f, rb_balance), sstep(0.16f, 0.34f, blue_lift));
        const float rp_signal = sstep(0.50f, 0.72f, rb_balance) * sstep(0.24f, 0.42f, blue_lift);
        const float balanced_hue = sstep(0.06f * isens, 0.22f * isens, min_dir) * sstep(0.35f, 0.75f, balance);
        const float rp_hue = is_purple ? minf_(1.f, rp_signal * 1.50f) *
                                             sstep(0.035f * isens, 0.16f * isens, min_dir) *
                                             sstep(0.18f, 0.58f, balance)
                                       : 0.f;
        const float c_hue = maxf_(balanced_hue, rp_hue);
        const float red_guard = 1.f - sstep(0.08f, 0.25f, red_strength) * (1.f - red_purple);
        const float green_enable = (enable_green || is_purple) ? 1.f : 0.f;
        fs[i] = c_edge * c_bright * c_hue * c_chroma * red_guard * green_enable;
    }
    gblur(&B, fs, bfs);

    /* line link, :4266-4323 */
    float *ls = malloc(N * sizeof(float)), *lfs = calloc(N, sizeof(float));
    for (size_t i = 0; i < N; i++) ls[i] = maxf_(fs[i], bfs[i] * 0.55f);
    static const int dirs[8][2] = {{1, 0}, {0, 1}, {1, 1}, {1, -1}, {2, 1}, {1, 2}, {2, -1}, {1, -2}};
#define SAMPLE(map, X, Y) \
    (map)[(size_t)((Y) < 0 ? 0 : ((Y) > H - 1 ? H - 1 : (Y))) * W + ((X) < 0 ? 0 : ((X) > W - 1 ? W - 1 : (X)))]
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            float best = 0.f;
            for (int d = 0; d < 8; d++) {
                const int tx = dirs[d][0], ty = dirs[d][1], nx = -ty, ny = tx;
                float along = 0.f;
                const int max_step = d < 4 ? 14 : 6;
                for (int s = 1; s <= max_step; s++) {
                    along = maxf_(along, SAMPLE(ls, x + tx * s, y + ty * s));
                    along = maxf_(along, SAMPLE(ls, x - tx * s, y - ty * s));
                }
                const float across = maxf_(SAMPLE(fs, x + nx * 2, y + ny * 2), SAMPLE(fs, x - nx * 2, y - ny * 2));
                const float thin = 1.f - sstep(along * 0.45f, along * 0.95f, across);
                best = maxf_(best, along * thin);
            }
            lfs[(size_t)y * W + x] = best;
        }
#undef SAMPLE

    /* solo seeds, :4325-4376 */
    float *sr = calloc(N, sizeof(float)), *sb = calloc(N, sizeof(float)), *srb = calloc(N, sizeof(float)), *sbb = calloc(N, sizeof(float));
    const float solo_enable = sstep(1.25f, 3.0f, strength);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; i++) {
        const float g = maxf_(src[3 * i + 1], 0.f);
        const float tr = g * expf(blur_rg[i]);
        const float tb = g * expf(blur_bg[i]);
        const float solo_r = maxf_(0.f, src[3 * i] - tr) / maxf_(blur_guide[i], 0.05f);
        const float solo_b = maxf_(0.f, src[3 * i + 2] - tb) / maxf_(blur_guide[i], 0.05f);
        const float linked = sstep(0.004f, 0.045f, maxf_(bfs[i], lfs[i] * 0.90f));
        const float c_edge = sstep(edge_threshold * 0.35f * isens, edge_threshold * 1.50f * isens, blur_edge[i]) *
                             sstep(0.55f * isens, 1.35f * isens, blur_edge_raw[i]);
        const float c_bright = sstep(0.24f * isens, 0.52f * isens, blur_guide[i]);
        const float r0 = maxf_(src[3 * i], 0.f), g0 = maxf_(src[3 * i + 1], 0.f), b0 = maxf_(src[3 * i + 2], 0.f);
        const float red_strength = maxf_(0.f, r0 - maxf_(g0, b0)) / maxf_(r0, 1e-4f);
        const float red_guard = 1.f - 0.90f * sstep(0.12f, 0.32f, red_strength);
        sr[i] = solo_enable * c_edge * c_bright * linked * red_guard *
                sstep(chroma_threshold * 0.22f * isens, chroma_threshold * 0.95f * isens, solo_r);
        sb[i] = solo_enable * c_edge * c_bright *
                sstep(chroma_threshold * 0.22f * isens, chroma_threshold * 0.95f * isens, solo_b);
    }
    gblur(&B, sr, srb);
    gblur(&B, sb, sbb);

    /* main correction and solo cleanup, :4377-4987 */
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; i++) {
        /* the input pixel, read before this pixel's output is written */
        const float s0 = src[3 * i], s1 = src[3 * i + 1], s2 = src[3 * i + 2];
        const float rg = log_rg[i], bg = log_bg[i];
        const float dr = rg - blur_rg[i], db = bg - blur_bg[i];
        const float r0 = maxf_(s0, 0.f), g0 = maxf_(s1, 0.f), b0 = maxf_(s2, 0.f);
        const float rb_avg_pre = 0.5f * (r0 + b0);
        const float purple_r = maxf_(0.f, dr), purple_b = maxf_(0.f, db);
        const float green_r = maxf_(0.f, -dr), green_b = maxf_(0.f, -db);
        const float purple_excess = minf_(purple_r, purple_b);
        const float denom = maxf_(blur_guide[i], 0.05f);
        const float g_over_min = maxf_(0.f, g0 - minf_(r0, b0)) / denom;
        const float green_dom = sstep(0.010f, 0.060f, g_over_min);
        const float green_excess = maxf_(green_r, green_b) * green_dom;
        const int is_purple = purple_excess >= green_excess;
        const float excess = is_purple ? purple_excess : green_excess;
        const float min_dir = is_purple ? maxf_(0.f, minf_(rg, bg)) : maxf_(0.f, minf_(-rg, -bg));
        const float max_dir = is_purple ? maxf_(1e-4f, maxf_(rg, bg)) : maxf_(1e-4f, maxf_(-rg, -bg));
        const float balance = min_dir / max_dir;
        const float c_edge_core = sstep(edge_threshold * 0.35f * isens, edge_threshold * 1.50f * isens, blur_edge[i]) *
                                  sstep(0.55f * isens, 1.35f * isens, blur_edge_raw[i]);
        const float c_edge_low = sstep(edge_threshold * 0.10f * isens, edge_threshold * 0.80f * isens, blur_edge[i]) *
                                 sstep(0.20f * isens, 0.90f * isens, blur_edge_raw[i]);
        const float linked = sstep(0.004f, 0.045f, maxf_(bfs[i], lfs[i] * 0.90f));
        const float c_bright = sstep(0.24f * isens, 0.52f * isens, blur_guide[i]);
        const float c_chroma = sstep(chroma_threshold * 0.45f * isens, chroma_threshold * 1.60f * isens, excess);
        const float rb_balance_hue = b0 / maxf_(r0, 1e-4f);
        const float blue_lift_hue = maxf_(0.f, b0 - g0) / maxf_(r0 - g0, 1e-4f);
        const float rp_signal = sstep(0.50f, 0.72f, rb_balance_hue) * sstep(0.24f, 0.42f, blue_lift_hue);
        const float balanced_hue = sstep(0.06f * isens, 0.22f * isens, min_dir) * sstep(0.35f, 0.75f, balance);
        const float rp_gate = is_purple ? minf_(1.f, rp_signal * 1.50f) * sstep(0.035f * isens, 0.16f * isens, min_dir) *
                                              sstep(0.18f, 0.58f, balance)
                                        : 0.f;
        const float green_unbal = (enable_green && !is_purple) ? sstep(0.035f * isens, 0.16f * isens, max_dir) : 0.f;
        const float c_hue = maxf_(balanced_hue, maxf_(rp_gate, green_unbal));
        const float rb_avg = rb_avg_pre;
        const float purple_abs = minf_(maxf_(0.f, r0 - g0), maxf_(0.f, b0 - g0)) / maxf_(blur_guide[i], 0.05f);
        const float green_abs_base = maxf_(0.f, g0 - rb_avg) / maxf_(blur_guide[i], 0.05f);
        const float green_abs_alt = maxf_(0.f, g0 - minf_(r0, b0)) / maxf_(blur_guide[i], 0.05f);
        const float green_abs = maxf_(green_abs_base, 0.55f * green_abs_alt);
        const int abs_is_purple = purple_abs >= green_abs;
        const float abs_excess = abs_is_purple ? purple_abs : green_abs;
        const float c_abs = sstep(chroma_threshold * 0.35f * isens, chroma_threshold * 1.20f * isens, abs_excess);
        const float low_contrast = sstep(chroma_threshold * 1.20f * isens, chroma_threshold * 2.40f * isens, abs_excess);
        const float green_soft = (enable_green && !is_purple) ? c_edge_low * sstep(0.05f * isens, 0.20f * isens, max_dir) : 0.f;
        const float c_edge = maxf_(c_edge_core, maxf_(c_edge_low * linked * 1.45f,
                                                      maxf_(c_edge_low * low_contrast * 0.35f, green_soft)));
        const float ratio_mask = c_edge * c_bright * c_hue * c_chroma;
        const float direct_mask = c_edge * c_bright * c_hue * c_abs;
        const int use_direct = direct_mask > ratio_mask;
        float mask = maxf_(ratio_mask, direct_mask) * scene_gate;
        const int correct_purple = use_direct ? abs_is_purple : is_purple;
        float local_gate;
        if (correct_purple) {
            local_gate = sstep(chroma_threshold * 0.10f * isens, chroma_threshold * 0.55f * isens, purple_abs);
            mask *= local_gate;
        } else {
            local_gate = sstep(chroma_threshold * 0.10f * isens, chroma_threshold * 0.55f * isens, green_abs);
            if (enable_green) {
                const float peak_chroma = sstep(chroma_threshold * 0.80f * isens, chroma_threshold * 2.50f * isens, green_abs);
                const float peak_nb = sstep(0.005f, 0.030f, bfs[i]);
                const float green_peak_mask = peak_chroma * peak_nb * c_bright * c_hue;
                mask = maxf_(mask, green_peak_mask * scene_gate);
            }
            const float hl_protect = 1.f - 0.85f * sstep(0.62f, 0.88f, blur_guide[i]);
            const float rb_to_green = rb_avg / maxf_(g0, 1e-4f);
            const float natural_hl = sstep(0.22f, 0.46f, rb_to_green) * sstep(0.30f, 0.55f, g0);
            const float nat_suppress = 1.f - 0.99f * natural_hl;
            const float nat_suppress_sq = nat_suppress * nat_suppress;
            const float halo = sstep(chroma_threshold * 0.55f * isens, chroma_threshold * 1.65f * isens, green_abs);
            const float gain_base = enable_green ? 0.60f : 0.35f;
            const float gain_slope = enable_green ? 1.10f : 0.95f;
            const float green_gain = gain_base + gain_slope * halo;
            const float min_ratio_ref = minf_(blur_rg[i], blur_bg[i]);
            const float green_nb = sstep(0.02f, 0.10f, maxf_(0.f, -min_ratio_ref));
            const float smax = maxf_(r0, maxf_(g0, b0));
            const float smin = minf_(r0, minf_(g0, b0));
            const float schroma = smax - smin;
            const float ssat = schroma / maxf_(smax, 1e-4f);
            const float whiteish = 1.f - sstep(0.05f, 0.13f, ssat);
            const float foliage_guard = green_nb * whiteish * sstep(0.35f, 0.75f, blur_guide[i]) * (1.f - halo);
            const float strength_release = enable_green ? sstep(2.0f, 6.0f, strength) : 0.f;
            const float line_evidence = sstep(0.004f, 0.030f, lfs[i]) * sstep(0.20f, 0.55f, halo) * sstep(0.05f, 0.15f, ssat);
            const float not_natural = 1.f - natural_hl;
            const float release = strength_release * line_evidence * not_natural;
            const float rel_hl = hl_protect + (1.f - hl_protect) * release;
            const float rel_nat = nat_suppress_sq + (1.f - nat_suppress_sq) * release;
            const float foliage_term = 1.f - 0.999f * foliage_guard;
            const float rel_foliage = foliage_term + (1.f - foliage_term) * release;
            mask *= local_gate * green_gain * rel_hl * rel_nat * rel_foliage;
            const float kill_evidence = sstep(0.04f, 0.15f, lfs[i]);
            const float natural_blob = sqrtf(maxf_(0.f, natural_hl)) * (1.f - kill_evidence) * green_nb;
            mask *= 1.f - 0.999f * natural_blob;
            if (!enable_green) mask = 0.f;
        }

        /* current output values of this pixel */
        float o0 = s0, o1 = s1, o2 = s2;
        int skip_solo = 0;
        if (mask > mask_cutoff) {
            const float g = maxf_(s1, 0.f);
            const float red_strength = maxf_(0.f, r0 - maxf_(g0, b0)) / maxf_(r0, 1e-4f);
            const float rb_balance = b0 / maxf_(r0, 1e-4f);
            const float blue_lift = maxf_(0.f, b0 - g0) / maxf_(r0 - g0, 1e-4f);
            const float rp_balance = sstep(0.55f, 0.78f, rb_balance);
            const float rp_lift = sstep(0.16f, 0.34f, blue_lift);
            const float red_purple = maxf_(rp_balance, rp_lift);
            const float red_strength_guard = sstep(0.08f, 0.25f, red_strength);
            float rp_context = 0.f, ctx_rp = 0.f, red_magenta = 0.f;
            if (red_purple > 0.50f) {
                rp_context = linked;
                ctx_rp = red_purple * rp_context;
                if (rp_context < 0.20f && red_strength_guard < 0.50f && red_purple > 0.50f)
                    red_magenta = sstep(0.02f, 0.12f, purple_abs) * (1.f - rp_context);
            }
            float neutral_gray_surface = 0.f;
            if (strength > 3.f && rp_context > 0.80f && lfs[i] > 0.012f && red_strength_guard < 0.98f &&
                blue_lift > 1.10f && rb_balance > 0.82f && purple_abs > 0.015f) {
                const float ng_edge = sstep(1.10f, 1.80f, blue_lift) * sstep(0.82f, 1.05f, rb_balance) *
                                      (1.f - red_strength_guard);
                neutral_gray_surface = ng_edge * sstep(0.015f, 0.08f, purple_abs) * sstep(3.f, 8.f, strength);
            }
            const float red_object = maxf_(red_strength_guard, red_magenta);
            const float nwg_base = 1.f - red_object * (1.f - ctx_rp);
            const float nwg_sq = nwg_base * nwg_base * nwg_base * nwg_base;
            const float nwg = nwg_sq * nwg_sq;
            const float line_relief = sstep(0.012f, 0.06f, lfs[i]) * sstep(0.10f, 0.24f, blue_lift);
            const float red_object_guard = sstep(0.55f, 0.90f, red_object);
            const float linked_override = linked * 1.05f * (1.f - 0.85f * red_object_guard);
            const float line_override = line_relief * (1.f - 0.85f * red_object_guard);
            const int force_warm = (red_object_guard > 0.75f) && (linked < 0.65f) && (line_relief < 0.45f);
            const float warm_protect = correct_purple ? (force_warm ? nwg : maxf_(nwg, maxf_(linked_override, line_override))) : 1.f;
            const float red_wo_support = sstep(0.22f, 0.45f, red_strength) * (1.f - sstep(0.02f, 0.12f, bfs[i]));
            const float protect = correct_purple ? red_wo_support : 0.f;
            const float smooth_mask = sstep(mask_cutoff, mask_cutoff * 5.0f, mask);
            const float threshold_soften = 1.f - 0.86f * sstep(0.22f, 0.40f, chroma_threshold);
            const float smax = maxf_(r0, maxf_(g0, b0));
            const float smin = minf_(r0, minf_(g0, b0));
            const float schroma = smax - smin;
            const float ssat = schroma / maxf_(smax, 1e-4f);
            const float ng_base = sstep(0.008f, 0.030f, ssat);
            const float neutral_guard = 1.f - (1.f - ng_base) * sstep(0.35f, 0.80f, red_object);
            float amount = clampf_(smooth_mask * strength * 1.90f * (1.f - protect) * warm_protect * neutral_guard *
                                       threshold_soften,
                                   0.f, 1.f);
            const float warm_edge = sstep(0.006f, 0.035f, purple_abs) * sstep(0.25f, 0.65f, red_purple) *
                                    sstep(0.15f, 0.65f, rp_context) * (1.f - 0.55f * red_strength_guard) *
                                    sstep(3.f, 8.f, strength);
            const float safe_surface = maxf_(neutral_gray_surface, warm_edge);
            if (correct_purple && safe_surface > 1e-4f) amount = minf_(amount, 1.f - 0.99f * safe_surface);
            if (amount <= 1e-4f) {
                skip_solo = 1; /* `continue` in the reference skips the solo cleanup */
            } else {
                const float target_r = minf_(g, g * expf(blur_rg[i]));
                const float target_b = minf_(g, g * expf(blur_bg[i]));
                if (correct_purple) {
                    float cr = clampf_(s0 * (1.f - amount) + target_r * amount, 0.f, s0);
                    float cg = s1;
                    float cb = clampf_(s2 * (1.f - amount) + target_b * amount, 0.f, s2);
                    if (safe_surface > 1e-4f) {
                        const float luma_before = 0.2126f * s0 + 0.7152f * s1 + 0.0722f * s2;
                        const float luma_after = 0.2126f * cr + 0.7152f * cg + 0.0722f * cb;
                        const float max_drop = luma_before * 0.02f * (1.f - neutral_gray_surface);
                        const float min_luma = maxf_(0.f, luma_before - max_drop);
                        if (luma_after < min_luma) {
                            const float lift = min_luma - luma_after;
                            const float ceil_ = maxf_(s0, maxf_(s1, s2));
                            cr = minf_(s0, cr + lift);
                            cg = minf_(ceil_, cg + lift);
                            cb = minf_(s2, cb + lift);
                        }
                        const float smx = maxf_(s0, maxf_(s1, s2));
                        const float smn = minf_(s0, minf_(s1, s2));
                        const float cmx = maxf_(cr, maxf_(cg, cb));
                        const float cmn = minf_(cr, minf_(cg, cb));
                        const float sch = smx - smn;
                        const float cch = cmx - cmn;
                        const float min_chroma = sch * (1.f - 0.12f * safe_surface);
                        if (sch > 1e-4f && cch < min_chroma) {
                            const float dn = maxf_(sch - cch, 1e-4f);
                            const float keep = clampf_((sch - min_chroma) / dn, 0.f, 1.f);
                            cr = s0 * (1.f - keep) + cr * keep;
                            cg = s1 * (1.f - keep) + cg * keep;
                            cb = s2 * (1.f - keep) + cb * keep;
                        }
                    }
                    o0 = cr; o1 = cg; o2 = cb;
                } else {
                    float target_g = s1;
                    const float rg_ref = expf(blur_rg[i]);
                    const float bg_ref = expf(blur_bg[i]);
                    if (rg_ref > 1e-4f && bg_ref > 1e-4f) {
                        const float g_from_r = maxf_(s0, 0.f) / rg_ref;
                        const float g_from_b = maxf_(s2, 0.f) / bg_ref;
                        target_g = 0.5f * (g_from_r + g_from_b);
                    }
                    const float rb_avg_target = 0.5f * (maxf_(s0, 0.f) + maxf_(s2, 0.f));
                    target_g = maxf_(target_g, rb_avg_target);
                    const float green_amount = clampf_(amount * green_strength_scale, 0.f, 1.f);
                    o0 = s0;
                    o1 = clampf_(s1 * (1.f - green_amount) + target_g * green_amount, 0.f, s1);
                    o2 = s2;
                }
            }
        }
        if (!skip_solo && solo_enable > 0.f) {
            /* in place the reference reads src == dst here, i.e. this pixel's
               corrected values; out of place it reads the input */
            const float q0 = inplace ? o0 : s0, q1 = inplace ? o1 : s1, q2 = inplace ? o2 : s2;
            const float g = maxf_(q1, 0.f);
            const float target_r = g * expf(blur_rg[i]);
            const float target_b = g * expf(blur_bg[i]);
            const float solo_r = maxf_(0.f, q0 - target_r) / maxf_(blur_guide[i], 0.05f);
            const float solo_b = maxf_(0.f, q2 - target_b) / maxf_(blur_guide[i], 0.05f);
            const float local_r = sstep(chroma_threshold * 0.12f * isens, chroma_threshold * 0.65f * isens, solo_r);
            const float local_b = sstep(chroma_threshold * 0.12f * isens, chroma_threshold * 0.65f * isens, solo_b);
            const float qr0 = maxf_(q0, 0.f), qg0 = g, qb0 = maxf_(q2, 0.f);
            const float gom = maxf_(0.f, qg0 - minf_(qr0, qb0)) / maxf_(blur_guide[i], 0.05f);
            const float gdom = sstep(0.010f, 0.060f, gom);
            const float dsolo = maxf_(blur_guide[i], 0.05f);
            const float purpleish = sstep(0.010f, 0.060f, maxf_(maxf_(0.f, qr0 - qg0), maxf_(0.f, qb0 - qg0)) / dsolo);
            const float allow = 1.f - gdom * (1.f - purpleish);
            const float mr = srb[i] * local_r * allow;
            const float mb = sbb[i] * local_b * allow;
            if (mr > mask_cutoff) {
                const float a = clampf_(mr * strength * 0.85f, 0.f, 0.85f);
                o0 = clampf_(o0 * (1.f - a) + target_r * a, 0.f, o0);
            }
            if (mb > mask_cutoff) {
                const float a = clampf_(mb * strength * 0.95f, 0.f, 0.90f);
                o2 = clampf_(o2 * (1.f - a) + target_b * a, 0.f, o2);
            }
        }
        dst[3 * i] = o0;
        dst[3 * i + 1] = o1;
        dst[3 * i + 2] = o2;
    }

    free(guide_linear); free(guide_signal); free(log_rg); free(log_bg);
    free(edge); free(edge_raw); free(B.kernel); free(B.tmp);
    free(blur_guide); free(blur_rg); free(blur_bg); free(blur_edge); free(blur_edge_raw);
    free(fs); free(bfs); free(ls); free(lfs); free(sr); free(sb); free(srb); free(sbb);
    return 0;
}
