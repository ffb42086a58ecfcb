// Edge-gated Gaussian defringe for gfx950 (core/cpu_accelerator.cpp:3939-5064).
//
// Pipeline (planes are float, H x W, in the context scratch):
//   max(G) ─► guide / log R/G / log B/G and the Sobel |∇| of guide_signal + max ─►
//   5 Gaussian blurs (the support seed in the last one's epilogue) ─► blur ─►
//   8-direction line link (the solo seeds in its epilogue) ─► 2 blurs ─►
//   per-pixel correction.
// The two global maxima stay on the device (atomicMax on the bit pattern of a
// non-negative float), so the whole stage is enqueued without a host sync.
//
// Blurs sum the taps in the reference's order (k = -ks..ks, product then add,
// no contraction), so blurred planes are bit-identical to the reference's.  The
// logarithms and exponentials are evaluated in double and rounded to float
// (correctly rounded); the reference's glibc logf/expf differ from that in the
// last bit on ~0.1 % of arguments, which moves the result by ~1e-7.
#include "lre_common.h"
#include "lre_internal.h"

namespace {

constexpr int BLOCK = 256;
constexpr int MAX_TAPS = 129;  // radius <= 64 px

struct Scal {
    int W, H, ks, inplace, enable_green;
    float edge_thr, chroma_thr, strength, green_scale;
    float isens, mask_cutoff, solo_enable, threshold_soften;
};

LRE_DEV float sstep(float e0, float e1, float x) {  // :3322-3325
    float t = lre::clamp_std((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

// Smoothstep edges known before the launch: the ramp starts at e0 and is d
// wide; y is 1/d rounded to nearest.
struct Edge {
    float e0, d, y;
};

// n / d correctly rounded, from y = RN(1/d): the two residual corrections of
// the hardware division expansion (q = n y; r = n - d q; q += r y; twice)
// without its operand scaling, which is needed only near the ends of the
// exponent range.  Callers keep 2^-20 <= d <= 2^100 and use the quotient for
// 0 < n < d only; in that range the sequence is the IEEE quotient (checked
// against division for every numerator of two binades and 20000 divisors,
// and for every numerator in [d 2^-76, 2d] of the divisors used here).
LRE_DEV float quot(float n, float d, float y) {
    float q = n * y;
    float r = __builtin_fmaf(-d, q, n);
    q = __builtin_fmaf(r, y, q);
    r = __builtin_fmaf(-d, q, n);
    return __builtin_fmaf(r, y, q);
}

// smoothstep over a precomputed Edge.  Outside (0, d) the clamp decides
// (n >= d -> 1, n <= 0 -> 0, NaN stays NaN), as clamp(n / d, 0, 1) does; a
// t below 2^-76 gives 0 either way, since t * t underflows.
template <bool FAST>
LRE_DEV float sstep(const Edge &e, float x) {
    const float n = x - e.e0;
    float t;
    if (FAST) {
        t = quot(n, e.d, e.y);
        t = n >= e.d ? 1.0f : t;
        t = n <= 0.0f ? 0.0f : t;
    } else {
        t = lre::clamp_std(n / e.d, 0.0f, 1.0f);
    }
    return t * t * (3.0f - 2.0f * t);
}

// smoothstep with literal edges: d and 1/d fold at compile time
LRE_DEV float sstepc(float e0, float e1, float x) {
    const Edge e{e0, e1 - e0, 1.0f / (e1 - e0)};
    return sstep<true>(e, x);
}

// Edges that depend on the call's parameters (isens, edge and chroma
// thresholds, mask cutoff), formed on the host with the reference's float
// products in its order.
enum Gate {
    G_EDGE,          // et 0.35 i .. et 1.50 i, on the blurred normalised edge
    G_EDGE_RATIO,    // 0.55 i .. 1.35 i, on the blurred raw edge
    G_EDGE_LOW,      // et 0.10 i .. et 0.80 i
    G_RATIO_LOW,     // 0.20 i .. 0.90 i
    G_BRIGHT,        // 0.24 i .. 0.52 i
    G_CHROMA,        // ct 0.45 i .. ct 1.60 i
    G_BAL_HUE,       // 0.06 i .. 0.22 i
    G_DIR,           // 0.035 i .. 0.16 i
    G_GREEN_SOFT,    // 0.05 i .. 0.20 i
    G_SOLO,          // ct 0.22 i .. ct 0.95 i
    G_SOLO_LOCAL,    // ct 0.12 i .. ct 0.65 i
    G_LOCAL,         // ct 0.10 i .. ct 0.55 i
    G_ABS,           // ct 0.35 i .. ct 1.20 i
    G_LOW_CONTRAST,  // ct 1.20 i .. ct 2.40 i
    G_PEAK,          // ct 0.80 i .. ct 2.50 i
    G_HALO,          // ct 0.55 i .. ct 1.65 i
    G_MASK,          // mask_cutoff .. 5 mask_cutoff
    G_COUNT
};
struct Gates {
    Edge g[G_COUNT];
};
LRE_DEV float mx(float a, float b) { return lre::max_std(a, b); }
LRE_DEV float mn(float a, float b) { return lre::min_std(a, b); }
LRE_DEV float expf_cr(float x) { return (float)exp((double)x); }

// Exact shortcuts.  Where a pixel's edge gate is 0 the reference multiplies
// it into a product of factors that are all finite numbers in [0, 1] as long
// as the pixel's inputs are of moderate size, so the product is +0 (support
// and solo seeds), or the pixel is left as it is (correction).  A wavefront
// skips the rest of the work only when every one of its pixels qualifies;
// otherwise it computes every pixel in full, so the results are the same bits.
// moderate: |x| <= 1e30 (NaN fails), so differences and quotients by the
// reference's floors (1e-4, 0.05) stay finite.
LRE_DEV bool moderate(float x) { return fabsf(x) <= 1e30f; }
LRE_DEV bool wave_all(bool p) { return __all(p); }
LRE_DEV float logf_cr(float x) { return (float)log((double)x); }

// block-wide max of non-negative floats folded into *dst
LRE_DEV void block_max_to(float v, unsigned *dst) {
    __shared__ float red[BLOCK / 64];
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        float m = red[0];
        for (int k = 1; k < BLOCK / 64; k++) m = fmaxf(m, red[k]);
        // the maximum only grows, so a block below a value already stored
        // (even a stale one) leaves it unchanged: skip the contended atomic
        const unsigned u = __float_as_uint(m);
        if (u > __hip_atomic_load(dst, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(dst, u);
    }
}

__global__ void init_stats(unsigned *stats) {
    if (threadIdx.x < 4) stats[threadIdx.x] = 0u;
}

// max over pixels of max(0, G)  (:3961-3969)
__global__ __launch_bounds__(BLOCK) void max_green(const float *__restrict__ src, size_t n,
                                                   unsigned *stats) {
    const float m = lre::green_max(src, n, 0.f, [](float v) { return mx(0.f, v); });
    block_max_to(m, &stats[0]);
}

LRE_DEV float max_guide_of(const unsigned *stats) { return mx(__uint_as_float(stats[0]), 1e-6f); }


struct Planes {
    const float *src;
    float *dst;
    float *lrg, *lbg, *bg, *brg, *bbg, *be, *ber, *fs, *bfs, *lfs, *sr, *sb, *srb, *sbb;
};

// fringe support seed of pixel i (:4162-4248); ber is the blurred raw edge
// map at i (the caller may hold it before it is stored)
template <bool FAST>
LRE_DEV float support_px(const Planes &P, const Scal &C, const Gates &G, size_t i, float ber) {
    const auto gate = [&](Gate k, float x) { return sstep<FAST>(G.g[k], x); };
    const float rg = P.lrg[i], bg = P.lbg[i], brg = P.brg[i], bbg = P.bbg[i], bgd = P.bg[i];
    const float s0 = P.src[3 * i], s1 = P.src[3 * i + 1], s2 = P.src[3 * i + 2];
    const float c_edge = gate(G_EDGE, P.be[i]) * gate(G_EDGE_RATIO, ber);
    if (wave_all(c_edge == 0.f && moderate(rg) && moderate(bg) && moderate(brg) && moderate(bbg) &&
                 moderate(bgd) && moderate(s0) && moderate(s1) && moderate(s2)))
        return 0.f;
    const float dr = rg - brg, db = bg - bbg;
    const float r0 = mx(s0, 0.f), g0 = mx(s1, 0.f), b0 = mx(s2, 0.f);
    const float purple_excess = mn(mx(0.f, dr), mx(0.f, db));
    const float denom = mx(bgd, 0.05f);
    const float g_over_min = mx(0.f, g0 - mn(r0, b0)) / denom;
    const float green_dom = sstepc(0.010f, 0.060f, g_over_min);
    const float green_excess = mx(mx(0.f, -dr), mx(0.f, -db)) * green_dom;
    const bool is_purple = purple_excess >= green_excess;
    const float excess = is_purple ? purple_excess : green_excess;
    const float min_dir = is_purple ? mx(0.f, mn(rg, bg)) : mx(0.f, mn(-rg, -bg));
    const float max_dir = is_purple ? mx(1e-4f, mx(rg, bg)) : mx(1e-4f, mx(-rg, -bg));
    const float balance = min_dir / max_dir;
    const float c_bright = gate(G_BRIGHT, bgd);
    const float c_chroma = gate(G_CHROMA, excess);
    const float red_strength = mx(0.f, r0 - mx(g0, b0)) / mx(r0, 1e-4f);
    const float rb_balance = b0 / mx(r0, 1e-4f);
    const float blue_lift = mx(0.f, b0 - g0) / mx(r0 - g0, 1e-4f);
    const float red_purple = mx(sstepc(0.55f, 0.78f, rb_balance), sstepc(0.16f, 0.34f, blue_lift));
    const float rp_signal = sstepc(0.50f, 0.72f, rb_balance) * sstepc(0.24f, 0.42f, blue_lift);
    const float balanced_hue = gate(G_BAL_HUE, min_dir) * sstepc(0.35f, 0.75f, balance);
    const float rp_hue = is_purple ? mn(1.f, rp_signal * 1.50f) * gate(G_DIR, min_dir) *
                                         sstepc(0.18f, 0.58f, balance)
                                   : 0.f;
    const float c_hue = mx(balanced_hue, rp_hue);
    const float red_guard = 1.f - sstepc(0.08f, 0.25f, red_strength) * (1.f - red_purple);
    const float green_enable = (C.enable_green || is_purple) ? 1.f : 0.f;
    return c_edge * c_bright * c_hue * c_chroma * red_guard * green_enable;
}

// ---- separable Gaussian, clamp-to-edge, taps in the reference's order ----
struct BlurSet {
    const float *in[5];
    float *tmp[5];
    float *out[5];
    int scale_by_edge[5];  // multiply the input by 1/max(edge) (the normalised edge map)
    int count;
};

// normalised taps, passed by value (no host-to-device copy per call)
struct Taps {
    float v[MAX_TAPS];
};

// Everything per image of a defringe launch: its planes, the plane sets of
// the three blur passes, its two global maxima and its size.  The kernels
// take it through lre::One<DfImg> (one image) or lre::Many<DfImg> (several);
// the parameters common to the batch (Scal without the size, Gates, Taps) are
// kernel arguments.
struct DfImg {
    Planes P;
    float *gl, *er;
    BlurSet sets[3];
    unsigned *stats;
    int W, H, inplace;
};
// the batch-wide scalars with one image's size and aliasing filled in
LRE_DEV Scal scal_of(const Scal &C, const DfImg &I) {
    Scal c = C;
    c.W = I.W;
    c.H = I.H;
    c.inplace = I.inplace;
    return c;
}
constexpr int DF_CHUNK = BLOCK * 8;  // pixels per workgroup of the pointwise passes

// guide planes, log chroma ratios (:3970-3991) and the 3x3 Sobel magnitude of
// guide_signal, 0 on the frame (:3996-4027), in one pass: a workgroup takes a
// 64 x 16 tile and puts guide_signal of the tile and a one-pixel ring in LDS;
// guide_signal itself is never stored.
constexpr int PS_TX = 64, PS_TY = 16, PS_W = PS_TX + 2, PS_H = PS_TY + 2;

template <class Src>
__global__ __launch_bounds__(BLOCK) void prep_sobel(Src S) {
    const lre::TileOf tl = S.tile();
    const DfImg &I = S.args(tl);
    const float *src = I.P.src;
    float *__restrict__ gl = I.gl, *__restrict__ lrg = I.P.lrg, *__restrict__ lbg = I.P.lbg;
    float *__restrict__ er = I.er;
    const int W = I.W, H = I.H;
    unsigned *stats = I.stats;
    __shared__ float sgs[PS_H * PS_W];
    const float max_guide = max_guide_of(stats);
    const float ratio_eps = mx(max_guide * 1e-4f, 1e-6f);
    const int x0 = tl.tx * PS_TX, y0 = tl.ty * PS_TY;
    // ring cells outside the image only feed frame pixels, whose magnitude is 0
    lre::stage2d<BLOCK>(
        PS_H, PS_W,
        [&](int r, int c) {
            const int y = min(max(y0 - 1 + r, 0), H - 1), x = min(max(x0 - 1 + c, 0), W - 1);
            return src[3 * ((size_t)y * W + x) + 1];
        },
        [&](int r, int c, float g) {
            sgs[r * PS_W + c] = sqrtf(lre::clamp_std(mx(g, 0.f) / max_guide, 0.f, 1.f));
        });
    __syncthreads();
    const int ly = H > 1 ? H - 1 : 1, lx = W > 1 ? W - 1 : 1;
    float m = 0.f;
    for (int t = threadIdx.x; t < PS_TX * PS_TY; t += BLOCK) {
        const int ty = t / PS_TX, tx = t % PS_TX, y = y0 + ty, x = x0 + tx;
        if (x >= W || y >= H) continue;
        const size_t i = (size_t)y * W + x;
        const float r = src[3 * i], g = src[3 * i + 1], b = src[3 * i + 2];
        const float gg = mx(g, 0.f);
        gl[i] = gg;
        lrg[i] = logf_cr((mx(r, 0.f) + ratio_eps) / (gg + ratio_eps));
        lbg[i] = logf_cr((mx(b, 0.f) + ratio_eps) / (gg + ratio_eps));
        float mag = 0.f;
        if (y >= 1 && y < ly && x >= 1 && x < lx) {
            const float *u = sgs + ty * PS_W + tx, *c = u + PS_W, *d = c + PS_W;  // rows y-1, y, y+1 at x-1
            float gx = -u[0] + u[2] - 2.f * c[0] + 2.f * c[2] - d[0] + d[2];
            float gy = -u[0] - 2.f * u[1] - u[2] + d[0] + 2.f * d[1] + d[2];
            mag = sqrtf(gx * gx + gy * gy);
            if (mag > m) m = mag;
        }
        er[i] = mag;
    }
    block_max_to(m, &stats[1]);
}

constexpr int HB_W = 256;  // outputs per block along x

__global__ __launch_bounds__(BLOCK) void blur_h(BlurSet S, Taps taps, int W, int H, int ks,
                                                const unsigned *stats) {
    __shared__ float seg[HB_W + 2 * (MAX_TAPS / 2)];
    __shared__ float kt[MAX_TAPS];
    const int y = blockIdx.y, x0 = blockIdx.x * HB_W;
    for (int t = threadIdx.x; t < 2 * ks + 1; t += BLOCK) kt[t] = taps.v[t];
    const float inv_edge = 1.f / mx(__uint_as_float(stats[1]), 1e-6f);
    for (int p = 0; p < S.count; p++) {
        const float *in = S.in[p] + (size_t)y * W;
        const float sc = S.scale_by_edge[p] ? inv_edge : 1.f;
        __syncthreads();
        for (int t = threadIdx.x; t < HB_W + 2 * ks; t += BLOCK) {
            int xx = x0 - ks + t;
            xx = xx < 0 ? 0 : (xx > W - 1 ? W - 1 : xx);
            seg[t] = S.scale_by_edge[p] ? in[xx] * sc : in[xx];
        }
        __syncthreads();
        const int x = x0 + threadIdx.x;
        if (x < W) {
            float acc = 0.f;
            for (int k = 0; k <= 2 * ks; k++) acc += kt[k] * seg[threadIdx.x + k];
            S.tmp[p][(size_t)y * W + x] = acc;
        }
    }
}

constexpr int VB_W = 64, VB_H = 64;

// vertical pass: each workgroup stages (VB_H + 2 ks) rows x 64 columns of the
// horizontal result in LDS (rows clamped at the frame) and sums the taps
__global__ __launch_bounds__(BLOCK) void blur_v(BlurSet S, Taps taps, int W, int H, int ks) {
    extern __shared__ float col_tile[];  // (VB_H + 2 ks) x VB_W
    __shared__ float kt[MAX_TAPS];
    for (int t = threadIdx.x; t < 2 * ks + 1; t += BLOCK) kt[t] = taps.v[t];
    const int x0 = blockIdx.x * VB_W, y0 = blockIdx.y * VB_H;
    const int rows = VB_H + 2 * ks;
    const int tx = threadIdx.x % VB_W, ty = threadIdx.x / VB_W;
    const int x = x0 + tx;
    for (int p = 0; p < S.count; p++) {
        const float *tm = S.tmp[p];
        __syncthreads();
        for (int t = threadIdx.x; t < rows * VB_W; t += BLOCK) {
            const int r = t / VB_W, c = t % VB_W;
            int yy = y0 - ks + r;
            yy = yy < 0 ? 0 : (yy > H - 1 ? H - 1 : yy);
            const int xx = x0 + c;
            col_tile[t] = xx < W ? tm[(size_t)yy * W + xx] : 0.f;
        }
        __syncthreads();
        if (x < W)
            for (int r = ty; r < VB_H; r += BLOCK / VB_W) {
                const int y = y0 + r;
                if (y >= H) break;
                float acc = 0.f;
                for (int k = 0; k <= 2 * ks; k++) acc += kt[k] * col_tile[(r + k) * VB_W + tx];
                S.out[p][(size_t)y * W + x] = acc;
            }
    }
}

// Both passes in one kernel for ks <= FB_KSMAX: a workgroup stages the
// (64 + 2 ks)^2 input window of a 64 x 64 output tile in LDS (coordinates
// clamped at the frame), runs the horizontal taps over every staged row (the
// rows of the separable form's intermediate image at the clamped row
// indices), then the vertical taps.  Each thread accumulates FB_RB outputs
// along the filter direction; every output sums its taps in the order
// k = 0 .. 2 ks, as the separable kernels do.
// Output tile of the fused blur: FB_TX columns by TY rows, TY one of two
// heights.  Every output sums its taps in the same order whatever the tile,
// so both heights give the same bits; the short tiles put more workgroups on a
// small image, the tall ones redo fewer rows of the horizontal pass on a large
// one (0.3 MP: 0.139 against 0.161 ms per defringe; 24 MP: 1.93 against
// 1.83 ms).
constexpr int FB_TX = 64, FB_TY_SHORT = 48, FB_TY_TALL = 96, FB_KSMAX = 16, FB_RB = 4;
constexpr int FB_BLOCK = 512;
static_assert(FB_KSMAX <= (MAX_TAPS - 1) / 2, "fused blur radius within the tap table");

// With SUP, the last plane of the set is the blurred raw edge map and each
// output of it also yields the fringe support seed of its pixel: the thread
// that writes it wrote the pixel's other blurred planes in the earlier plane
// iterations (barriers between them).
template <int KS, bool SUP, int FB_TY, class Src>
__global__ __launch_bounds__(FB_BLOCK) void blur2d(Src SR, int which, Taps taps, Scal Cb, Gates G) {
    extern __shared__ float fb[];
    const lre::TileOf tl = SR.tile();
    const DfImg &I = SR.args(tl);
    const BlurSet &S = I.sets[which];
    const Planes &P = I.P;
    const Scal C = scal_of(Cb, I);
    const int W = I.W, H = I.H;
    const unsigned *stats = I.stats;
    constexpr int ks = KS, nt = 2 * KS + 1;
    constexpr int IW = FB_TX + 2 * ks, IH = FB_TY + 2 * ks;
    // odd row pitches: lanes that walk down rows hit distinct LDS banks
    constexpr int PW = IW | 1;
    constexpr int PH = FB_TX + 1;
    float *win = fb;              // IH x IW input window (pitch PW)
    float *hor = fb + IH * PW;    // IH x FB_TX horizontal result (pitch PH)
    float kt[nt];                 // taps in registers (the loops below unroll fully)
#pragma unroll
    for (int t = 0; t < nt; t++) kt[t] = taps.v[t];
    const float inv_edge = 1.f / mx(__uint_as_float(stats[1]), 1e-6f);
    const int x0 = tl.tx * FB_TX, y0 = tl.ty * FB_TY;
    for (int p = 0; p < S.count; p++) {
        const float *in = S.in[p];
        const bool scaled = S.scale_by_edge[p] != 0;
        bool zero = true;
        __syncthreads();
        lre::stage2d<FB_BLOCK>(
            IH, IW,
            [&](int r, int c) {
                int yy = y0 - ks + r, xx = x0 - ks + c;
                yy = yy < 0 ? 0 : (yy > H - 1 ? H - 1 : yy);
                xx = xx < 0 ? 0 : (xx > W - 1 ? W - 1 : xx);
                return in[(size_t)yy * W + xx];
            },
            [&](int r, int c, float v) {
                const float x = scaled ? v * inv_edge : v;
                win[r * PW + c] = x;
                zero = zero && x == 0.f;
            });
        // A window of zeros blurs to +0 everywhere (taps times zeros, summed
        // from +0): the sparse seed planes have many such tiles.  The support
        // seeds of the last plane need the pass itself.
        if (__syncthreads_and(zero) && !(SUP && p == S.count - 1)) {
            float *out = S.out[p];
            for (int t = threadIdx.x; t < FB_TY * FB_TX; t += FB_BLOCK) {
                const int x = x0 + t % FB_TX, y = y0 + t / FB_TX;
                if (x < W && y < H) out[(size_t)y * W + x] = 0.f;
            }
            continue;
        }
        for (int t = threadIdx.x; t < IH * (FB_TX / FB_RB); t += FB_BLOCK) {
            const int r = t % IH, c = (t / IH) * FB_RB;
            const float *row = win + r * PW + c;
            float acc[FB_RB], w[FB_RB];
#pragma unroll
            for (int u = 0; u < FB_RB; u++) { acc[u] = 0.f; w[u] = row[u]; }
#pragma unroll
            for (int k = 0; k < nt; k++) {
                const float tk = kt[k];
#pragma unroll
                for (int u = 0; u < FB_RB; u++) acc[u] += tk * w[u];
                if (k + 1 == nt) break;
#pragma unroll
                for (int u = 0; u < FB_RB - 1; u++) w[u] = w[u + 1];
                w[FB_RB - 1] = row[k + FB_RB];
            }
#pragma unroll
            for (int u = 0; u < FB_RB; u++) hor[r * PH + c + u] = acc[u];
        }
        __syncthreads();
        float *out = S.out[p];
        for (int t = threadIdx.x; t < (FB_TY / FB_RB) * FB_TX; t += FB_BLOCK) {
            const int c = t % FB_TX, r = (t / FB_TX) * FB_RB;
            const float *col = hor + r * PH + c;
            float acc[FB_RB], w[FB_RB];
#pragma unroll
            for (int u = 0; u < FB_RB; u++) { acc[u] = 0.f; w[u] = col[u * PH]; }
#pragma unroll
            for (int k = 0; k < nt; k++) {
                const float tk = kt[k];
#pragma unroll
                for (int u = 0; u < FB_RB; u++) acc[u] += tk * w[u];
                if (k + 1 == nt) break;
#pragma unroll
                for (int u = 0; u < FB_RB - 1; u++) w[u] = w[u + 1];
                w[FB_RB - 1] = col[(k + FB_RB) * PH];
            }
            const int x = x0 + c;
            const bool seed = SUP && p == S.count - 1;
            if (x < W)
#pragma unroll
                for (int u = 0; u < FB_RB; u++) {
                    const int y = y0 + r + u;
                    if (y >= H) continue;
                    const size_t o = (size_t)y * W + x;
                    out[o] = acc[u];
                    if (seed) P.fs[o] = support_px<true>(P, C, G, o, acc[u]);
                }
        }
    }
}

// LDS of blur2d at radius ks and tile height ty: the input window and the
// horizontal result
constexpr size_t fb_lds(int ks, int ty) {
    return sizeof(float) * ((size_t)(ty + 2 * ks) * ((FB_TX + 2 * ks) | 1) + (size_t)(ty + 2 * ks) * (FB_TX + 1));
}

// the LDS window of blur2d can exceed 64 KiB: raise the limit once per instance
template <int KS, bool SUP, int TY, class Src>
hipError_t blur2d_lds_attr() {
    static const hipError_t rc = hipFuncSetAttribute((const void *)blur2d<KS, SUP, TY, Src>,
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)fb_lds(KS, TY));
    return rc;
}

// fringe support seed, as a pass of its own (the blur kernels with a
// radius above the fused limit, and parameters that need the dividing smoothsteps)
template <bool FAST>
__global__ __launch_bounds__(BLOCK) void support(Planes P, Scal C, Gates G) {
    const size_t n = (size_t)C.W * C.H;
    for (size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (size_t)gridDim.x * BLOCK)
        P.fs[i] = support_px<FAST>(P, C, G, i, P.ber[i]);
}

// solo seeds of pixel i (:4325-4376), given its line-linked support lfs
template <bool FAST>
LRE_DEV void solo_px(const Planes &P, const Scal &C, const Gates &G, size_t i, float lfs) {
    const auto gate = [&](Gate k, float x) { return sstep<FAST>(G.g[k], x); };
    const float s0 = P.src[3 * i], s1 = P.src[3 * i + 1], s2 = P.src[3 * i + 2];
    const float bgd = P.bg[i], bfs = P.bfs[i];
    const float c_edge = gate(G_EDGE, P.be[i]) * gate(G_EDGE_RATIO, P.ber[i]);
    // solo_r and solo_b are finite whatever the exponentials give (mx(0, x)
    // maps -inf and NaN to 0), so with the edge gate at 0 both seeds are +0
    if (wave_all(c_edge == 0.f && moderate(s0) && moderate(s1) && moderate(s2) && moderate(bgd) &&
                 moderate(bfs) && moderate(lfs))) {
        P.sr[i] = 0.f;
        P.sb[i] = 0.f;
        return;
    }
    const float g = mx(s1, 0.f);
    const float tr = g * expf_cr(P.brg[i]);
    const float tb = g * expf_cr(P.bbg[i]);
    const float solo_r = mx(0.f, s0 - tr) / mx(bgd, 0.05f);
    const float solo_b = mx(0.f, s2 - tb) / mx(bgd, 0.05f);
    const float linked = sstepc(0.004f, 0.045f, mx(bfs, lfs * 0.90f));
    const float c_bright = gate(G_BRIGHT, bgd);
    const float r0 = mx(s0, 0.f), g0 = mx(s1, 0.f), b0 = mx(s2, 0.f);
    const float red_strength = mx(0.f, r0 - mx(g0, b0)) / mx(r0, 1e-4f);
    const float red_guard = 1.f - 0.90f * sstepc(0.12f, 0.32f, red_strength);
    P.sr[i] = C.solo_enable * c_edge * c_bright * linked * red_guard *
              gate(G_SOLO, solo_r);
    P.sb[i] = C.solo_enable * c_edge * c_bright * gate(G_SOLO, solo_b);
}

// line-linked support (:4266-4323): ls = max(fs, 0.55 * blur(fs)) sampled with
// clamp-to-edge.  Each workgroup stages ls and fs for its 32x32 outputs plus a
// 14-px halo in LDS (coordinates clamped at load time, which equals clamping
// every sample).
constexpr int LL_T = 32, LL_HALO = 14, LL_R = LL_T + 2 * LL_HALO;

template <bool FAST, class Src>
__global__ __launch_bounds__(BLOCK) void line_link(Src SR, Scal Cb, Gates G) {
    const lre::TileOf tl = SR.tile();
    const DfImg &I = SR.args(tl);
    const Planes &P = I.P;
    const Scal C = scal_of(Cb, I);
    const int W = C.W, H = C.H;
    __shared__ float s_ls[LL_R * LL_R], s_fs[LL_R * LL_R];
    const int x0 = tl.tx * LL_T, y0 = tl.ty * LL_T;
    bool flat = true;
    lre::stage2d<BLOCK>(
        LL_R, LL_R,
        [&](int r, int c) {
            int y = y0 - LL_HALO + r, x = x0 - LL_HALO + c;
            x = x < 0 ? 0 : (x > W - 1 ? W - 1 : x);
            y = y < 0 ? 0 : (y > H - 1 ? H - 1 : y);
            const size_t i = (size_t)y * W + x;
            return make_float2(P.fs[i], P.bfs[i]);
        },
        [&](int r, int c, float2 v) {
            const float ls = mx(v.x, v.y * 0.55f);
            s_fs[r * LL_R + c] = v.x;
            s_ls[r * LL_R + c] = ls;
            flat = flat && ls == 0.f;
        });
    // With every staged ls at 0 each direction's `along` is +0, so along * thin
    // is +0 or NaN (0 / 0 in the smoothstep), which mx(best, .) ignores: every
    // output of the tile is +0 (:4296-4323)
    flat = __syncthreads_and(flat);
    // direction d: (tx, ty) = (1,0) (0,1) (1,1) (1,-1) (2,1) (1,2) (2,-1) (1,-2)
    for (int o = threadIdx.x; o < LL_T * LL_T; o += BLOCK) {
        const int ly = o / LL_T, lx = o % LL_T;
        const int x = x0 + lx, y = y0 + ly;
        if (x >= W || y >= H) continue;
        const int c = (ly + LL_HALO) * LL_R + (lx + LL_HALO);
        float best = 0.f;
        if (!flat)
#pragma unroll
        for (int d = 0; d < 8; d++) {
            const int tx = d == 1 ? 0 : ((d == 4 || d == 6) ? 2 : 1);
            const int ty = d == 0 ? 0 : (d == 1 || d == 2 || d == 4) ? 1 : (d == 3 || d == 6) ? -1 : (d == 5 ? 2 : -2);
            const int step = ty * LL_R + tx;
            const int nstep = tx * LL_R - ty;  // (nx, ny) = (-ty, tx)
            float along = 0.f;
            const int max_step = d < 4 ? 14 : 6;
            for (int s = 1; s <= max_step; s++) {
                along = mx(along, s_ls[c + step * s]);
                along = mx(along, s_ls[c - step * s]);
            }
            // along starts at +0 and takes only larger values, so it is +0 or
            // positive; at +0, along * thin is +0 or NaN (0 / 0 in the
            // smoothstep) and mx(best, .) keeps best either way: the
            // direction's division is skipped
            if (along > 0.f) {
                const float across = mx(s_fs[c + nstep * 2], s_fs[c - nstep * 2]);
                const float thin = 1.f - sstep(along * 0.45f, along * 0.95f, across);
                best = mx(best, along * thin);
            }
        }
        const size_t i = (size_t)y * W + x;
        P.lfs[i] = best;
        solo_px<FAST>(P, C, G, i, best);  // the solo seeds need this pixel's lfs only
    }
}

// main correction and solo cleanup (:4377-4987)
template <bool FAST, class Src>
__global__ __launch_bounds__(BLOCK) void correct(Src SR, Scal Cb, Gates G) {
    const lre::TileOf tl = SR.tile();
    const DfImg &I = SR.args(tl);
    const Planes &P = I.P;
    const Scal C = scal_of(Cb, I);
    const unsigned *stats = I.stats;
    const size_t n = (size_t)C.W * C.H;
    const size_t chunk_end = ((size_t)tl.tx + 1) * DF_CHUNK < n ? ((size_t)tl.tx + 1) * DF_CHUNK : n;
    const float strength = C.strength;
    const auto gate = [&](Gate k, float x) { return sstep<FAST>(G.g[k], x); };
    const bool eg = C.enable_green != 0;
    const float scene_gate = sstepc(0.55f, 0.75f, max_guide_of(stats));
    const float mask_cutoff = C.mask_cutoff;
    for (size_t i = (size_t)tl.tx * DF_CHUNK + threadIdx.x; i < chunk_end; i += BLOCK) {
        const float s0 = P.src[3 * i], s1 = P.src[3 * i + 1], s2 = P.src[3 * i + 2];
        const float bguide = P.bg[i], brg = P.brg[i], bbg = P.bbg[i], bfs = P.bfs[i], lfs = P.lfs[i];
        const float rg = P.lrg[i], bg = P.lbg[i], be = P.be[i], ber = P.ber[i];
        const float c_edge_core = gate(G_EDGE, be) * gate(G_EDGE_RATIO, ber);
        const float c_edge_low = gate(G_EDGE_LOW, be) * gate(G_RATIO_LOW, ber);
        // With both edge gates at 0 (and the green correction off) both masks
        // are +0, so the main correction does not run; the solo masks are
        // products of srb / sbb with two factors in [0, 1], so below the cutoff
        // when those are.  The pixel is then returned as it came in.
        if (wave_all(!eg && c_edge_core == 0.f && c_edge_low == 0.f &&
                     (!(C.solo_enable > 0.f) || (P.srb[i] <= mask_cutoff && P.sbb[i] <= mask_cutoff)) &&
                     moderate(s0) && moderate(s1) && moderate(s2) && moderate(bguide) && moderate(brg) &&
                     moderate(bbg) && moderate(bfs) && moderate(lfs) && moderate(rg) && moderate(bg))) {
            if (P.dst != P.src) {
                P.dst[3 * i] = s0;
                P.dst[3 * i + 1] = s1;
                P.dst[3 * i + 2] = s2;
            }
            continue;
        }
        const float dr = rg - brg, db = bg - bbg;
        const float r0 = mx(s0, 0.f), g0 = mx(s1, 0.f), b0 = mx(s2, 0.f);
        const float rb_avg = 0.5f * (r0 + b0);
        const float purple_excess = mn(mx(0.f, dr), mx(0.f, db));
        const float denom = mx(bguide, 0.05f);
        const float g_over_min = mx(0.f, g0 - mn(r0, b0)) / denom;
        const float green_dom = sstepc(0.010f, 0.060f, g_over_min);
        const float green_excess = mx(mx(0.f, -dr), mx(0.f, -db)) * green_dom;
        const bool is_purple = purple_excess >= green_excess;
        const float excess = is_purple ? purple_excess : green_excess;
        const float min_dir = is_purple ? mx(0.f, mn(rg, bg)) : mx(0.f, mn(-rg, -bg));
        const float max_dir = is_purple ? mx(1e-4f, mx(rg, bg)) : mx(1e-4f, mx(-rg, -bg));
        const float balance = min_dir / max_dir;
        const float linked = sstepc(0.004f, 0.045f, mx(bfs, lfs * 0.90f));
        const float c_bright = gate(G_BRIGHT, bguide);
        const float c_chroma = gate(G_CHROMA, excess);
        const float rb_balance_hue = b0 / mx(r0, 1e-4f);
        const float blue_lift_hue = mx(0.f, b0 - g0) / mx(r0 - g0, 1e-4f);
        const float rp_signal = sstepc(0.50f, 0.72f, rb_balance_hue) * sstepc(0.24f, 0.42f, blue_lift_hue);
        const float balanced_hue = gate(G_BAL_HUE, min_dir) * sstepc(0.35f, 0.75f, balance);
        const float rp_gate = is_purple ? mn(1.f, rp_signal * 1.50f) * gate(G_DIR, min_dir) *
                                              sstepc(0.18f, 0.58f, balance)
                                        : 0.f;
        const float green_unbal = (eg && !is_purple) ? gate(G_DIR, max_dir) : 0.f;
        const float c_hue = mx(balanced_hue, mx(rp_gate, green_unbal));
        const float purple_abs = mn(mx(0.f, r0 - g0), mx(0.f, b0 - g0)) / mx(bguide, 0.05f);
        const float green_abs_base = mx(0.f, g0 - rb_avg) / mx(bguide, 0.05f);
        const float green_abs_alt = mx(0.f, g0 - mn(r0, b0)) / mx(bguide, 0.05f);
        const float green_abs = mx(green_abs_base, 0.55f * green_abs_alt);
        const bool abs_is_purple = purple_abs >= green_abs;
        const float abs_excess = abs_is_purple ? purple_abs : green_abs;
        const float c_abs = gate(G_ABS, abs_excess);
        const float low_contrast = gate(G_LOW_CONTRAST, abs_excess);
        const float green_soft = (eg && !is_purple) ? c_edge_low * gate(G_GREEN_SOFT, max_dir) : 0.f;
        const float c_edge = mx(c_edge_core, mx(c_edge_low * linked * 1.45f, mx(c_edge_low * low_contrast * 0.35f, green_soft)));
        const float ratio_mask = c_edge * c_bright * c_hue * c_chroma;
        const float direct_mask = c_edge * c_bright * c_hue * c_abs;
        const bool use_direct = direct_mask > ratio_mask;
        float mask = mx(ratio_mask, direct_mask) * scene_gate;
        const bool correct_purple = use_direct ? abs_is_purple : is_purple;
        if (correct_purple) {
            mask *= gate(G_LOCAL, purple_abs);
        } else if (!eg) {
            // the reference computes the green terms below and then sets the
            // mask to 0 without the green correction (:4698-4700): nothing else
            // reads them
            mask = 0.f;
        } else {
            const float local_gate = gate(G_LOCAL, green_abs);
            const float peak_chroma = gate(G_PEAK, green_abs);
            const float peak_nb = sstepc(0.005f, 0.030f, bfs);
            mask = mx(mask, peak_chroma * peak_nb * c_bright * c_hue * scene_gate);
            const float hl_protect = 1.f - 0.85f * sstepc(0.62f, 0.88f, bguide);
            const float rb_to_green = rb_avg / mx(g0, 1e-4f);
            const float natural_hl = sstepc(0.22f, 0.46f, rb_to_green) * sstepc(0.30f, 0.55f, g0);
            const float nat_suppress = 1.f - 0.99f * natural_hl;
            const float nat_suppress_sq = nat_suppress * nat_suppress;
            const float halo = gate(G_HALO, green_abs);
            const float green_gain = 0.60f + 1.10f * halo;  // with the green correction
            const float green_nb = sstepc(0.02f, 0.10f, mx(0.f, -mn(brg, bbg)));
            const float smax = mx(r0, mx(g0, b0)), smin = mn(r0, mn(g0, b0));
            const float ssat = (smax - smin) / mx(smax, 1e-4f);
            const float whiteish = 1.f - sstepc(0.05f, 0.13f, ssat);
            const float foliage_guard = green_nb * whiteish * sstepc(0.35f, 0.75f, bguide) * (1.f - halo);
            const float strength_release = sstepc(2.0f, 6.0f, strength);
            const float line_evidence = sstepc(0.004f, 0.030f, lfs) * sstepc(0.20f, 0.55f, halo) * sstepc(0.05f, 0.15f, ssat);
            const float release = strength_release * line_evidence * (1.f - natural_hl);
            const float rel_hl = hl_protect + (1.f - hl_protect) * release;
            const float rel_nat = nat_suppress_sq + (1.f - nat_suppress_sq) * release;
            const float foliage_term = 1.f - 0.999f * foliage_guard;
            const float rel_foliage = foliage_term + (1.f - foliage_term) * release;
            mask *= local_gate * green_gain * rel_hl * rel_nat * rel_foliage;
            const float kill_evidence = sstepc(0.04f, 0.15f, lfs);
            const float natural_blob = sqrtf(mx(0.f, natural_hl)) * (1.f - kill_evidence) * green_nb;
            mask *= 1.f - 0.999f * natural_blob;
        }

        float o0 = s0, o1 = s1, o2 = s2;
        bool skip_solo = false;
        float e_rg = 0.f, e_bg = 0.f;
        bool have_exp = false;
        if (mask > mask_cutoff) {
            const float g = mx(s1, 0.f);
            const float red_strength = mx(0.f, r0 - mx(g0, b0)) / mx(r0, 1e-4f);
            const float rb_balance = b0 / mx(r0, 1e-4f);
            const float blue_lift = mx(0.f, b0 - g0) / mx(r0 - g0, 1e-4f);
            const float red_purple = mx(sstepc(0.55f, 0.78f, rb_balance), sstepc(0.16f, 0.34f, blue_lift));
            const float rsg = sstepc(0.08f, 0.25f, red_strength);
            float rp_context = 0.f, ctx_rp = 0.f, red_magenta = 0.f;
            if (red_purple > 0.50f) {
                rp_context = linked;
                ctx_rp = red_purple * rp_context;
                if (rp_context < 0.20f && rsg < 0.50f && red_purple > 0.50f)
                    red_magenta = sstepc(0.02f, 0.12f, purple_abs) * (1.f - rp_context);
            }
            float ngs = 0.f;
            if (strength > 3.f && rp_context > 0.80f && lfs > 0.012f && rsg < 0.98f && blue_lift > 1.10f &&
                rb_balance > 0.82f && purple_abs > 0.015f) {
                const float ng_edge = sstepc(1.10f, 1.80f, blue_lift) * sstepc(0.82f, 1.05f, rb_balance) * (1.f - rsg);
                ngs = ng_edge * sstepc(0.015f, 0.08f, purple_abs) * sstepc(3.f, 8.f, strength);
            }
            const float red_object = mx(rsg, red_magenta);
            const float nwg_base = 1.f - red_object * (1.f - ctx_rp);
            const float nwg_sq = nwg_base * nwg_base * nwg_base * nwg_base;
            const float nwg = nwg_sq * nwg_sq;
            const float line_relief = sstepc(0.012f, 0.06f, lfs) * sstepc(0.10f, 0.24f, blue_lift);
            const float rog = sstepc(0.55f, 0.90f, red_object);
            const float linked_override = linked * 1.05f * (1.f - 0.85f * rog);
            const float line_override = line_relief * (1.f - 0.85f * rog);
            const bool force_warm = (rog > 0.75f) && (linked < 0.65f) && (line_relief < 0.45f);
            const float warm_protect = correct_purple ? (force_warm ? nwg : mx(nwg, mx(linked_override, line_override))) : 1.f;
            const float red_wo_support = sstepc(0.22f, 0.45f, red_strength) * (1.f - sstepc(0.02f, 0.12f, bfs));
            const float protect = correct_purple ? red_wo_support : 0.f;
            const float smooth_mask = gate(G_MASK, mask);
            const float smax = mx(r0, mx(g0, b0)), smin = mn(r0, mn(g0, b0));
            const float ssat = (smax - smin) / mx(smax, 1e-4f);
            const float ng_base = sstepc(0.008f, 0.030f, ssat);
            const float neutral_guard = 1.f - (1.f - ng_base) * sstepc(0.35f, 0.80f, red_object);
            float amount = lre::clamp_std(smooth_mask * strength * 1.90f * (1.f - protect) * warm_protect *
                                              neutral_guard * C.threshold_soften,
                                          0.f, 1.f);
            const float warm_edge = sstepc(0.006f, 0.035f, purple_abs) * sstepc(0.25f, 0.65f, red_purple) *
                                    sstepc(0.15f, 0.65f, rp_context) * (1.f - 0.55f * rsg) * sstepc(3.f, 8.f, strength);
            const float safe_surface = mx(ngs, warm_edge);
            if (correct_purple && safe_surface > 1e-4f) amount = mn(amount, 1.f - 0.99f * safe_surface);
            if (amount <= 1e-4f) {
                skip_solo = true;
            } else {
                e_rg = expf_cr(brg);
                e_bg = expf_cr(bbg);
                have_exp = true;
                const float target_r = mn(g, g * e_rg);
                const float target_b = mn(g, g * e_bg);
                if (correct_purple) {
                    float cr = lre::clamp_std(s0 * (1.f - amount) + target_r * amount, 0.f, s0);
                    float cg = s1;
                    float cb = lre::clamp_std(s2 * (1.f - amount) + target_b * amount, 0.f, s2);
                    if (safe_surface > 1e-4f) {
                        const float luma_before = 0.2126f * s0 + 0.7152f * s1 + 0.0722f * s2;
                        const float luma_after = 0.2126f * cr + 0.7152f * cg + 0.0722f * cb;
                        const float max_drop = luma_before * 0.02f * (1.f - ngs);
                        const float min_luma = mx(0.f, luma_before - max_drop);
                        if (luma_after < min_luma) {
                            const float lift = min_luma - luma_after;
                            const float ceil_ = mx(s0, mx(s1, s2));
                            cr = mn(s0, cr + lift);
                            cg = mn(ceil_, cg + lift);
                            cb = mn(s2, cb + lift);
                        }
                        const float smx = mx(s0, mx(s1, s2)), smn = mn(s0, mn(s1, s2));
                        const float cmx = mx(cr, mx(cg, cb)), cmn = mn(cr, mn(cg, cb));
                        const float sch = smx - smn, cch = cmx - cmn;
                        const float min_chroma = sch * (1.f - 0.12f * safe_surface);
                        if (sch > 1e-4f && cch < min_chroma) {
                            const float dn = mx(sch - cch, 1e-4f);
                            const float keep = lre::clamp_std((sch - min_chroma) / dn, 0.f, 1.f);
                            cr = s0 * (1.f - keep) + cr * keep;
                            cg = s1 * (1.f - keep) + cg * keep;
                            cb = s2 * (1.f - keep) + cb * keep;
                        }
                    }
                    o0 = cr; o1 = cg; o2 = cb;
                } else {
                    float target_g = s1;
                    if (e_rg > 1e-4f && e_bg > 1e-4f) {
                        const float g_from_r = mx(s0, 0.f) / e_rg;
                        const float g_from_b = mx(s2, 0.f) / e_bg;
                        target_g = 0.5f * (g_from_r + g_from_b);
                    }
                    target_g = mx(target_g, 0.5f * (mx(s0, 0.f) + mx(s2, 0.f)));
                    const float green_amount = lre::clamp_std(amount * C.green_scale, 0.f, 1.f);
                    o1 = lre::clamp_std(s1 * (1.f - green_amount) + target_g * green_amount, 0.f, s1);
                }
            }
        }
        if (!skip_solo && C.solo_enable > 0.f) {
            const float q0 = C.inplace ? o0 : s0, q1 = C.inplace ? o1 : s1, q2 = C.inplace ? o2 : s2;
            if (!have_exp) { e_rg = expf_cr(brg); e_bg = expf_cr(bbg); }
            const float g = mx(q1, 0.f);
            const float target_r = g * e_rg;
            const float target_b = g * e_bg;
            const float solo_r = mx(0.f, q0 - target_r) / mx(bguide, 0.05f);
            const float solo_b = mx(0.f, q2 - target_b) / mx(bguide, 0.05f);
            const float local_r = gate(G_SOLO_LOCAL, solo_r);
            const float local_b = gate(G_SOLO_LOCAL, solo_b);
            const float qr0 = mx(q0, 0.f), qb0 = mx(q2, 0.f);
            const float gom = mx(0.f, g - mn(qr0, qb0)) / mx(bguide, 0.05f);
            const float gdom = sstepc(0.010f, 0.060f, gom);
            const float purpleish = sstepc(0.010f, 0.060f, mx(mx(0.f, qr0 - g), mx(0.f, qb0 - g)) / mx(bguide, 0.05f));
            const float allow = 1.f - gdom * (1.f - purpleish);
            const float mr = P.srb[i] * local_r * allow;
            const float mb = P.sbb[i] * local_b * allow;
            if (mr > mask_cutoff) {
                const float a = lre::clamp_std(mr * strength * 0.85f, 0.f, 0.85f);
                o0 = lre::clamp_std(o0 * (1.f - a) + target_r * a, 0.f, o0);
            }
            if (mb > mask_cutoff) {
                const float a = lre::clamp_std(mb * strength * 0.95f, 0.f, 0.90f);
                o2 = lre::clamp_std(o2 * (1.f - a) + target_b * a, 0.f, o2);
            }
        }
        P.dst[3 * i] = o0;
        P.dst[3 * i + 1] = o1;
        P.dst[3 * i + 2] = o2;
    }
}

int grid_for(size_t n) {
    size_t b = (n + BLOCK - 1) / BLOCK;
    if (b > 8192) b = 8192;
    return (int)(b > 0 ? b : 1);
}

}  // namespace

// planes: 20 full planes + taps + stats
size_t lre_defringe_scratch_bytes(int w, int h) {
    const size_t plane = (((size_t)w * h * 4) + 255) & ~(size_t)255;
    return plane * 20 + 4096;
}

int lre_check_defringe_params(float radius) {
    const float sigma = (radius / 3.f < 0.5f) ? 0.5f : radius / 3.f;
    if (!std::isfinite(radius) || 2.f * std::ceil(3.f * sigma) + 1.f > (float)MAX_TAPS) {
        lre_set_error("defringe radius %g not supported (finite, at most %d taps)", radius, MAX_TAPS);
        return LRE_ERR_ARG;
    }
    return LRE_OK;
}

namespace {

// the parameters common to every image of a call
struct DfParams {
    Taps taps;
    Scal C;  // W, H and inplace are taken per image
    Gates G;
    int ks;
    bool fast;  // every ramp width inside the quotient shortcut's range
};

DfParams df_params(float radius, float edge_threshold, float chroma_threshold, float strength, int defringe_green,
                   float green_strength_scale) {
    DfParams D{};
    // Gaussian taps (:4033-4048), computed on the host with the C library's expf
    green_strength_scale = green_strength_scale < 0.f ? 0.f : (1.f < green_strength_scale ? 1.f : green_strength_scale);
    const float sigma = (radius / 3.f < 0.5f) ? 0.5f : radius / 3.f;
    const int ks = (int)std::ceil(3.f * sigma);
    D.ks = ks;
    {
        float ksum = 0.f;
        for (int k = -ks; k <= ks; k++) {
            float v = std::exp(-0.5f * (float)(k * k) / (sigma * sigma));
            D.taps.v[k + ks] = v;
            ksum += v;
        }
        for (int k = 0; k < 2 * ks + 1; k++) D.taps.v[k] /= ksum;
    }
    Scal &C = D.C;
    C.W = C.H = 0;
    C.ks = ks;
    C.inplace = 0;
    C.enable_green = defringe_green ? 1 : 0;
    C.edge_thr = edge_threshold; C.chroma_thr = chroma_threshold; C.strength = strength;
    C.green_scale = green_strength_scale;
    {
        float sv = std::sqrt(strength < 1.0f ? 1.0f : strength);
        const float sens = sv < 1.0f ? 1.0f : (2.5f < sv ? 2.5f : sv);
        C.isens = 1.0f / sens;
        C.mask_cutoff = 0.02f * C.isens;
        float t = (strength - 1.25f) / (3.0f - 1.25f);
        t = t < 0.f ? 0.f : (1.f < t ? 1.f : t);
        C.solo_enable = t * t * (3.0f - 2.0f * t);
        float u = (chroma_threshold - 0.22f) / (0.40f - 0.22f);
        u = u < 0.f ? 0.f : (1.f < u ? 1.f : u);
        C.threshold_soften = 1.f - 0.86f * (u * u * (3.0f - 2.0f * u));
    }
    // parameter-dependent smoothstep edges; the quotient shortcut needs every
    // ramp width inside [2^-20, 2^100] (else the kernels divide)
    D.fast = true;
    {
        const float i = C.isens, et = edge_threshold, ct = chroma_threshold, mc = C.mask_cutoff;
        auto set = [&](Gate k, float e0, float e1) {
            const float d = e1 - e0;
            D.G.g[k] = Edge{e0, d, 1.0f / d};
            if (!(d >= 0x1p-20f && d <= 0x1p100f)) D.fast = false;
        };
        set(G_EDGE, et * 0.35f * i, et * 1.50f * i);
        set(G_EDGE_RATIO, 0.55f * i, 1.35f * i);
        set(G_EDGE_LOW, et * 0.10f * i, et * 0.80f * i);
        set(G_RATIO_LOW, 0.20f * i, 0.90f * i);
        set(G_BRIGHT, 0.24f * i, 0.52f * i);
        set(G_CHROMA, ct * 0.45f * i, ct * 1.60f * i);
        set(G_BAL_HUE, 0.06f * i, 0.22f * i);
        set(G_DIR, 0.035f * i, 0.16f * i);
        set(G_GREEN_SOFT, 0.05f * i, 0.20f * i);
        set(G_SOLO, ct * 0.22f * i, ct * 0.95f * i);
        set(G_SOLO_LOCAL, ct * 0.12f * i, ct * 0.65f * i);
        set(G_LOCAL, ct * 0.10f * i, ct * 0.55f * i);
        set(G_ABS, ct * 0.35f * i, ct * 1.20f * i);
        set(G_LOW_CONTRAST, ct * 1.20f * i, ct * 2.40f * i);
        set(G_PEAK, ct * 0.80f * i, ct * 2.50f * i);
        set(G_HALO, ct * 0.55f * i, ct * 1.65f * i);
        set(G_MASK, mc, mc * 5.0f);
    }
    return D;
}

// One image's planes, carved from its scratch: 20 planes and the statistics,
// or with `compact` (the fused blur only, which needs no separable
// temporaries) 15 planes: the temporaries t2..t4 then alias planes the fused
// blur never touches through them.
DfImg df_image(const float *d_in, float *d_out, int w, int h, void *d_scratch, bool compact = false) {
    const size_t n = (size_t)w * h;
    const size_t plane = ((n * 4) + 255) & ~(size_t)255;
    char *base = static_cast<char *>(d_scratch);
    auto take = [&](int k) { return reinterpret_cast<float *>(base + plane * k); };
    const int o = compact ? 3 : 0;  // planes 10.. move down by three
    float *gl = take(0), *gs = take(1), *lrg = take(2), *lbg = take(3), *er = take(4);
    float *t0 = take(5), *t1 = take(6);
    float *t2 = compact ? t0 : take(7), *t3 = compact ? t0 : take(8), *t4 = compact ? t0 : take(9);
    float *bg = take(10 - o), *brg = take(11 - o), *bbg = take(12 - o), *be = take(13 - o), *ber = take(14 - o);
    float *fs = take(15 - o), *bfs = take(16 - o), *lfs = take(17 - o);
    float *sr = gl, *sb = gs, *srb = t0, *sbb = t1;  // reuse: the guide planes are dead by then
    DfImg I{};
    I.P.src = d_in; I.P.dst = d_out;
    I.P.lrg = lrg; I.P.lbg = lbg; I.P.bg = bg; I.P.brg = brg; I.P.bbg = bbg; I.P.be = be; I.P.ber = ber;
    I.P.fs = fs; I.P.bfs = bfs; I.P.lfs = lfs; I.P.sr = sr; I.P.sb = sb; I.P.srb = srb; I.P.sbb = sbb;
    I.gl = gl;
    I.er = er;
    I.stats = reinterpret_cast<unsigned *>(base + plane * (compact ? 15 : 20));
    I.W = w; I.H = h; I.inplace = d_in == d_out;
    // the five blurred guide planes (the normalised edge map scaled on load),
    // the support seeds, the two solo seed planes
    BlurSet &S5 = I.sets[0];
    S5.count = 5;
    const float *ins[5] = {gl, lrg, lbg, er, er};
    float *tms[5] = {t0, t1, t2, t3, t4};
    float *outs[5] = {bg, brg, bbg, be, ber};
    for (int k = 0; k < 5; k++) { S5.in[k] = ins[k]; S5.tmp[k] = tms[k]; S5.out[k] = outs[k]; S5.scale_by_edge[k] = 0; }
    S5.scale_by_edge[3] = 1;
    BlurSet &S1 = I.sets[1];
    S1.count = 1; S1.in[0] = fs; S1.tmp[0] = t2; S1.out[0] = bfs;
    BlurSet &S2 = I.sets[2];
    S2.count = 2;
    S2.in[0] = sr; S2.tmp[0] = t3; S2.out[0] = srb;
    S2.in[1] = sb; S2.tmp[1] = t4; S2.out[1] = sbb;
    return I;
}

// the fused blur of set `which` over a source (seed: the support seeds in its
// last plane); fused radii only
template <int TY, class Src>
int blur_fused_rows(Src src, dim3 grid, int which, bool seed, const DfParams &D, hipStream_t s) {
    const int ks = D.ks;
    const size_t lds = fb_lds(ks, TY);
    switch (ks) {
#define LRE_BLUR2D(K)                                                                                          \
    case K:                                                                                                    \
        if (seed) {                                                                                            \
            LRE_HIP_CHECK((blur2d_lds_attr<K, true, TY, Src>()));                                              \
            hipLaunchKernelGGL((blur2d<K, true, TY, Src>), grid, dim3(FB_BLOCK), lds, s, src, which, D.taps, \
                               D.C, D.G);                                                                      \
        } else {                                                                                               \
            LRE_HIP_CHECK((blur2d_lds_attr<K, false, TY, Src>()));                                             \
            hipLaunchKernelGGL((blur2d<K, false, TY, Src>), grid, dim3(FB_BLOCK), lds, s, src, which, D.taps,  \
                               D.C, D.G);                                                                      \
        }                                                                                                      \
        break;
        LRE_BLUR2D(1) LRE_BLUR2D(2) LRE_BLUR2D(3) LRE_BLUR2D(4) LRE_BLUR2D(5) LRE_BLUR2D(6)
        LRE_BLUR2D(7) LRE_BLUR2D(8) LRE_BLUR2D(9) LRE_BLUR2D(10) LRE_BLUR2D(11)
        LRE_BLUR2D(12) LRE_BLUR2D(13) LRE_BLUR2D(14) LRE_BLUR2D(15) LRE_BLUR2D(16)
#undef LRE_BLUR2D
    default:
        lre_set_error("defringe: radius %d has no fused blur", ks);
        return LRE_ERR_ARG;
    }
    LRE_HIP_CHECK(hipGetLastError());
    return LRE_OK;
}

// the grid of the blur tiles must be grid_blur(.., rows) of the same rows
template <class Src>
int blur_fused(Src src, dim3 grid, int which, bool seed, const DfParams &D, int rows, hipStream_t s) {
    return rows == FB_TY_SHORT ? blur_fused_rows<FB_TY_SHORT>(src, grid, which, seed, D, s)
                               : blur_fused_rows<FB_TY_TALL>(src, grid, which, seed, D, s);
}

// Grids of the tiled passes of one w x h image
dim3 grid_prep(int w, int h) { return dim3((w + PS_TX - 1) / PS_TX, (h + PS_TY - 1) / PS_TY); }
dim3 grid_blur(int w, int h, int rows) { return dim3((w + FB_TX - 1) / FB_TX, (h + rows - 1) / rows); }
dim3 grid_link(int w, int h) { return dim3((w + LL_T - 1) / LL_T, (h + LL_T - 1) / LL_T); }
int chunks_of(int w, int h) { return (int)(((size_t)w * h + DF_CHUNK - 1) / DF_CHUNK); }

// several images: statistics reset (one workgroup per image) and max(G)
// folded over DF_CHUNK-pixel chunks
__global__ void init_stats_many(lre::Many<DfImg> S) {
    const lre::TileOf tl = S.tile();
    if (threadIdx.x < 4) S.args(tl).stats[threadIdx.x] = 0u;
}
__global__ __launch_bounds__(BLOCK) void max_green_many(lre::Many<DfImg> S) {
    const lre::TileOf tl = S.tile();
    const DfImg &I = S.args(tl);
    const size_t n = (size_t)I.W * I.H, b = (size_t)tl.tx * DF_CHUNK, e = b + DF_CHUNK < n ? b + DF_CHUNK : n;
    const float m = lre::green_max_range(I.P.src, b, e, 0.f, [](float v) { return mx(0.f, v); });
    block_max_to(m, &I.stats[0]);
}

}  // namespace

int lre_defringe_blur_rows(int blur_rows, size_t tall_tiles) {
    if (blur_rows == FB_TY_SHORT || blur_rows == FB_TY_TALL) return blur_rows;
    // by size: short tiles until the tall ones give each CU four
    return tall_tiles < (size_t)4 * lre_current_cus() ? FB_TY_SHORT : FB_TY_TALL;
}

size_t lre_defringe_tall_tiles(int w, int h) {
    const dim3 g = grid_blur(w, h, FB_TY_TALL);
    return (size_t)g.x * g.y;
}

int lre_launch_defringe(const float *d_in, float *d_out, int w, int h, float radius,
                        float edge_threshold, float chroma_threshold, float strength,
                        int defringe_green, float green_strength_scale, void *d_scratch,
                        hipStream_t s, int blur_rows) {
    if (const int rc = lre_check_defringe_params(radius)) return rc;
    const size_t n = (size_t)w * h;
    const int rows = lre_defringe_blur_rows(blur_rows, lre_defringe_tall_tiles(w, h));
    const DfParams D = df_params(radius, edge_threshold, chroma_threshold, strength, defringe_green,
                                 green_strength_scale);
    const DfImg I = df_image(d_in, d_out, w, h, d_scratch);
    using One = lre::One<DfImg>;
    const One one{I};
    const int ks = D.ks;
    const Scal C = [&] {
        Scal c = D.C;
        c.W = w; c.H = h; c.inplace = I.inplace;
        return c;
    }();
    const int g = grid_for(n);
    hipLaunchKernelGGL(init_stats, dim3(1), dim3(64), 0, s, I.stats);
    hipLaunchKernelGGL(max_green, dim3(g), dim3(BLOCK), 0, s, d_in, n, I.stats);
    hipLaunchKernelGGL(prep_sobel<One>, grid_prep(w, h), dim3(BLOCK), 0, s, one);
    LRE_HIP_CHECK(hipGetLastError());
    // blur of a plane set; with seed, also the support seeds (the fused form
    // exists for the fused radii and the quotient smoothsteps only)
    auto blur = [&](int which, bool seed) -> int {
        if (ks <= FB_KSMAX) return blur_fused(one, grid_blur(w, h, rows), which, seed, D, rows, s);
        const BlurSet &S = I.sets[which];
        dim3 gh((w + HB_W - 1) / HB_W, h);
        hipLaunchKernelGGL(blur_h, gh, dim3(BLOCK), 0, s, S, D.taps, w, h, ks, I.stats);
        dim3 gv((w + VB_W - 1) / VB_W, (h + VB_H - 1) / VB_H);
        hipLaunchKernelGGL(blur_v, gv, dim3(BLOCK), sizeof(float) * (VB_H + 2 * ks) * VB_W, s, S, D.taps, w, h, ks);
        LRE_HIP_CHECK(hipGetLastError());
        return LRE_OK;
    };
    const bool seed_in_blur = D.fast && ks <= FB_KSMAX;
    int rc = blur(0, seed_in_blur);
    if (rc) return rc;
    if (!seed_in_blur) {
        if (D.fast) hipLaunchKernelGGL(support<true>, dim3(g), dim3(BLOCK), 0, s, I.P, C, D.G);
        else hipLaunchKernelGGL(support<false>, dim3(g), dim3(BLOCK), 0, s, I.P, C, D.G);
    }
    if ((rc = blur(1, false))) return rc;
    if (D.fast) hipLaunchKernelGGL((line_link<true, One>), grid_link(w, h), dim3(BLOCK), 0, s, one, D.C, D.G);
    else hipLaunchKernelGGL((line_link<false, One>), grid_link(w, h), dim3(BLOCK), 0, s, one, D.C, D.G);
    if ((rc = blur(2, false))) return rc;
    const dim3 gc(chunks_of(w, h));
    if (D.fast) hipLaunchKernelGGL((correct<true, One>), gc, dim3(BLOCK), 0, s, one, D.C, D.G);
    else hipLaunchKernelGGL((correct<false, One>), gc, dim3(BLOCK), 0, s, one, D.C, D.G);
    LRE_HIP_CHECK(hipGetLastError());
    return LRE_OK;
}

// ---- many images: one launch per pass over all of them ----------------------
// Six tables over the same per-image arguments, one per workgroup geometry:
// the statistics reset (one per image), max(G) chunks, the prep / Sobel tiles,
// the blur tiles, the line-link tiles and the correction chunks.
namespace {
using DfTable = lre::ManyTable<DfImg>;
size_t df_part(int n) { return (DfTable::bytes(n) + 255) & ~(size_t)255; }
}  // namespace

size_t lre_defringe_multi_table_bytes(int n) { return 6 * df_part(n); }

// the compact plane layout of the multi-image pass
size_t lre_defringe_multi_scratch_bytes(int w, int h) {
    return ((((size_t)w * h * 4) + 255) & ~(size_t)255) * 15 + 256;
}

int lre_defringe_multi_supported(float radius, float edge_threshold, float chroma_threshold, float strength,
                                 int defringe_green, float green_strength_scale) {
    if (lre_check_defringe_params(radius) != LRE_OK) return 0;
    const DfParams D = df_params(radius, edge_threshold, chroma_threshold, strength, defringe_green,
                                 green_strength_scale);
    return D.fast && D.ks <= FB_KSMAX;
}

int lre_defringe_multi_fill(const lre_post_job *jobs, int n, void *h_table, int grids[6], int rows) {
    auto args = [&](int i) {
        return df_image(jobs[i].d_in, jobs[i].d_out, jobs[i].w, jobs[i].h, jobs[i].d_scratch, true);
    };
    char *h = static_cast<char *>(h_table);
    auto as2 = [](dim3 g) { return make_int2((int)g.x, (int)g.y); };
    grids[0] = DfTable::fill(h, n, [](int) { return make_int2(1, 1); }, args);
    grids[1] = DfTable::fill(h + df_part(n), n, [&](int i) { return make_int2(chunks_of(jobs[i].w, jobs[i].h), 1); },
                             args);
    grids[2] = DfTable::fill(h + 2 * df_part(n), n, [&](int i) { return as2(grid_prep(jobs[i].w, jobs[i].h)); }, args);
    grids[3] = DfTable::fill(h + 3 * df_part(n), n, [&](int i) { return as2(grid_blur(jobs[i].w, jobs[i].h, rows)); }, args);
    grids[4] = DfTable::fill(h + 4 * df_part(n), n, [&](int i) { return as2(grid_link(jobs[i].w, jobs[i].h)); }, args);
    grids[5] = DfTable::fill(h + 5 * df_part(n), n, [&](int i) { return make_int2(chunks_of(jobs[i].w, jobs[i].h), 1); },
                             args);
    return LRE_OK;
}

int lre_launch_defringe_multi(const void *d_table, int n, const int grids[6], float radius, float edge_threshold,
                              float chroma_threshold, float strength, int defringe_green,
                              float green_strength_scale, int rows, hipStream_t s) {
    if (n == 0) return LRE_OK;
    if (!lre_defringe_multi_supported(radius, edge_threshold, chroma_threshold, strength, defringe_green,
                                      green_strength_scale)) {
        lre_set_error("defringe: these parameters have no multi-image pass");
        return LRE_ERR_ARG;
    }
    const DfParams D = df_params(radius, edge_threshold, chroma_threshold, strength, defringe_green,
                                 green_strength_scale);
    const char *d = static_cast<const char *>(d_table);
    auto view = [&](int k) { return DfTable::view(d + k * df_part(n), n); };
    using Many = lre::Many<DfImg>;
    hipLaunchKernelGGL(init_stats_many, dim3(grids[0]), dim3(64), 0, s, view(0));
    hipLaunchKernelGGL(max_green_many, dim3(grids[1]), dim3(BLOCK), 0, s, view(1));
    hipLaunchKernelGGL(prep_sobel<Many>, dim3(grids[2]), dim3(BLOCK), 0, s, view(2));
    LRE_HIP_CHECK(hipGetLastError());
    int rc;
    if ((rc = blur_fused(view(3), dim3(grids[3]), 0, true, D, rows, s))) return rc;
    if ((rc = blur_fused(view(3), dim3(grids[3]), 1, false, D, rows, s))) return rc;
    hipLaunchKernelGGL((line_link<true, Many>), dim3(grids[4]), dim3(BLOCK), 0, s, view(4), D.C, D.G);
    if ((rc = blur_fused(view(3), dim3(grids[3]), 2, false, D, rows, s))) return rc;
    hipLaunchKernelGGL((correct<true, Many>), dim3(grids[5]), dim3(BLOCK), 0, s, view(5), D.C, D.G);
    LRE_HIP_CHECK(hipGetLastError());
    return LRE_OK;
}
