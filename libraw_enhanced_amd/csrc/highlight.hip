// Float highlight recovery for gfx950, core/libraw_wrapper.cpp:1095-1383
// (Impl::recover_highlights).  Results equal the reference's serial order
// bit for bit:
//
//  1. hr_scan     the soft highlight mask and, per chunk of HR_CHUNK pixels,
//                 the number of pixels that enter the R/G, B/G ratio sums;
//                 the maximum channel of the highlight pixels (:1149-1215).
//  2. hr_sums     the ratio sums are serial float sums in raster order
//                 (:1180-1190): one wave walks the non-empty chunks in order
//                 and adds the quotients one by one.
//  3. hr_rewrite  per 64x16 tile with a one-pixel halo: the highlight
//                 rewrite (:1216-1241) of every staged pixel, then the
//                 8-neighbour mean of blown interior pixels (:1258-1285),
//                 which reads the rewritten neighbours and never another
//                 mean.  Flags the pixels the LoCA pass may change.
//  4. hr_loca     the LoCA green lift (:1287-1374) runs over the highlight
//                 pixels in raster order, in place, so a pixel reads the
//                 updated values of the pixels before it.  In the skewed
//                 coordinate t = x + 3y every pixel that (y, x) reads and
//                 that comes earlier in raster order has a smaller t, and
//                 every later one a larger t.  So all rows can advance
//                 together, one t per step.  The image is cut into tiles of
//                 64 rows x HR_TW values of t (parallelograms), and tile
//                 (band b, column k) needs only tiles (b, k-1), (b-1, k) and
//                 (b-1, k-1).  Each anti-diagonal b + k = d is one launch,
//                 and a tile with no flagged pixel returns at once.
#include "lre_common.h"
#include "lre_internal.h"

#include <algorithm>
#include <cfloat>

namespace {

constexpr int HR_CHUNK = 4096;  // pixels per chunk of the ratio scan
constexpr int HR_BLOCK = 256;
constexpr int RW_TX = 64, RW_TY = 16;  // rewrite tile (RW_TY divides the LoCA band height)
constexpr int RW_SX = RW_TX + 2, RW_SY = RW_TY + 2;
constexpr int HR_TR = 64;            // LoCA tile rows (one per lane of wave 0)
constexpr int HR_TW = 104;           // LoCA tile width in t
constexpr int HR_SR = HR_TR + 4;     // staged rows (two above, two below)
constexpr int HR_LW = HR_TW + 17;    // staged columns: t - 8 .. t + W + 8, odd stride
constexpr int HR_WORDS = (HR_TW + 31) / 32;
constexpr int HR_OW = HR_TW + 1;     // row stride of the tile's own pixels (odd: no LDS bank conflicts)
constexpr size_t HR_LDS = sizeof(float) * (2 * HR_SR * HR_LW + 3 * HR_TR * HR_OW) + HR_TR * HR_TW +
                          sizeof(unsigned) * HR_WORDS * (1 + HR_TR);

struct HrState {
    float grf, gbf;     // ratio means after the division by the count
    unsigned max_bits;  // max_val as float bits (never negative)
    unsigned total;     // pixels that enter the ratio sums
};

struct MaskConst {
    float lo, inv, chroma_lo, chroma_inv, inv_wb_r, inv_wb_g, inv_wb_b;
};

LRE_DEV float max3(float r, float g, float b) { return lre::max_std(r, lre::max_std(g, b)); }

// Green smoothstep times neutrality (:1152-1174)
LRE_DEV float mask_value(float r, float g, float b, const MaskConst &M) {
    float t = (g - M.lo) * M.inv;
    if (t < 0.f) t = 0.f;
    else if (t > 1.f) t = 1.f;
    const float w_g = t * t * (3.f - 2.f * t);
    const float r_n = r * M.inv_wb_r, g_n = g * M.inv_wb_g, b_n = b * M.inv_wb_b;
    const float mx = lre::max_std(r_n, lre::max_std(g_n, b_n));
    const float mn = lre::min_std(r_n, lre::min_std(g_n, b_n));
    const float chroma = (mx > 1e-6f) ? (mx - mn) / mx : 0.f;
    float ct = (chroma - M.chroma_lo) * M.chroma_inv;
    if (ct < 0.f) ct = 0.f;
    else if (ct > 1.f) ct = 1.f;
    const float neutrality = 1.f - ct * ct * (3.f - 2.f * ct);
    return w_g * neutrality;
}

LRE_DEV bool is_ratio_px(float r, float g, float b, float sat) {
    return r >= sat && b >= sat && r < 0.95f && g >= sat && g < 0.95f && b < 0.95f;
}

__global__ void hr_init(HrState *st, unsigned *tile_count, int ntiles) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        st->grf = st->gbf = 1.f;
        st->max_bits = 0u;
        st->total = 0u;
    }
    for (int k = i; k < ntiles; k += gridDim.x * blockDim.x) tile_count[k] = 0u;
}

// Mask, ratio quotients and max_val over one chunk.  The quotients r/g, b/g of
// the chunk's ratio pixels are written in raster order to the chunk's own list
// (capacity HR_CHUNK), so the serial sum reads only them.
__global__ __launch_bounds__(HR_BLOCK) void hr_scan(const float *__restrict__ in, float *__restrict__ mask,
                                                     size_t n, float sat, MaskConst M,
                                                     float2 *__restrict__ ratios,
                                                     unsigned *__restrict__ chunk_count, HrState *st) {
    __shared__ unsigned s_cnt[HR_BLOCK / 64];
    __shared__ float s_max[HR_BLOCK / 64];
    const size_t base = (size_t)blockIdx.x * HR_CHUNK;
    const int wv = threadIdx.x >> 6;
    unsigned running = 0;
    float mx = 0.f;
    for (int k = 0; k < HR_CHUNK; k += HR_BLOCK) {
        const size_t i = base + k + threadIdx.x;
        bool q = false;
        float qr = 0.f, qb = 0.f;
        if (i < n) {
            const float r = in[3 * i], g = in[3 * i + 1], b = in[3 * i + 2];
            if (mask) mask[i] = mask_value(r, g, b, M);
            if (r >= sat && b >= sat) {
                // max_val over the channels of the highlight pixels (:1199-1203)
                if (r > mx) mx = r;
                if (g > mx) mx = g;
                if (b > mx) mx = b;
                q = is_ratio_px(r, g, b, sat);
                if (q) {
                    qr = r / g;
                    qb = b / g;
                }
            }
        }
        const unsigned long long bal = __ballot(q);
        if ((threadIdx.x & 63) == 0) s_cnt[wv] = (unsigned)__popcll(bal);
        __syncthreads();
        unsigned before = running, total = 0;
#pragma unroll
        for (int v = 0; v < HR_BLOCK / 64; v++) {
            before += v < wv ? s_cnt[v] : 0u;
            total += s_cnt[v];
        }
        if (q) {
            // ratio pixels of this wave in lanes below this one
            const unsigned rank =
                __builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0u));
            ratios[base + before + rank] = make_float2(qr, qb);
        }
        running += total;
        __syncthreads();  // s_cnt is rewritten by the next round
    }
    for (int o = 32; o; o >>= 1) {
        const float other = __shfl_xor(mx, o);
        mx = other > mx ? other : mx;
    }
    if ((threadIdx.x & 63) == 0) s_max[wv] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        float m = 0.f;
        for (int v = 0; v < HR_BLOCK / 64; v++) m = s_max[v] > m ? s_max[v] : m;
        chunk_count[blockIdx.x] = running;
        if (running) atomicAdd(&st->total, running);
        if (m > 0.f) atomicMax(&st->max_bits, __float_as_uint(m));
    }
}

constexpr int HR_PRE = 16;  // quotients per chunk fetched ahead

// One wave: the serial float sums of r/g and b/g over the ratio pixels in
// raster order (:1180-1190), then the means (:1184-1190).  Each lane fetches
// the first HR_PRE quotients of one of 64 chunks; every lane then holds the
// same running sums, and readlane hands them the quotients one at a time.
__global__ __launch_bounds__(64) void hr_sums(const float2 *__restrict__ ratios,
                                              const unsigned *__restrict__ chunk_count, int nchunks,
                                              HrState *st) {
    const int lane = threadIdx.x;
    float grf = 0.f, gbf = 0.f, count = 0.f;
    const int walk = st->total ? nchunks : 0;
    auto add = [&](float2 v, int src) {
        grf += __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.x), src));
        gbf += __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.y), src));
        count += 1.f;
    };
    for (int c0 = 0; c0 < walk; c0 += 64) {
        const int cj = c0 + lane;
        const unsigned cnt = cj < nchunks ? chunk_count[cj] : 0u;
        unsigned long long live = __ballot(cnt != 0u);
        if (!live) continue;
        const float2 *lst = ratios + (size_t)cj * HR_CHUNK;
        float2 v[HR_PRE];
#pragma unroll
        for (int k = 0; k < HR_PRE; k++) v[k] = (unsigned)k < cnt ? lst[k] : make_float2(0.f, 0.f);
        while (live) {
            const int j = __ffsll(live) - 1;
            live &= live - 1;
            const unsigned m = (unsigned)__builtin_amdgcn_readlane((int)cnt, j);
#pragma unroll
            for (int k = 0; k < HR_PRE; k++)
                if ((unsigned)k < m) add(v[k], j);
            const float2 *lj = ratios + (size_t)(c0 + j) * HR_CHUNK;
            for (unsigned k = HR_PRE; k < m; k += 64) {  // the rest of a long list, 64 at a time
                const float2 u = k + lane < m ? lj[k + lane] : make_float2(0.f, 0.f);
                const unsigned mm = m - k < 64u ? m - k : 64u;
                for (unsigned i = 0; i < mm; i++) add(u, (int)i);
            }
        }
    }
    if (lane == 0) {
        if (count > 0.f) {
            grf /= count;
            gbf /= count;
        } else {
            grf = 1.f;
            gbf = 1.f;
        }
        st->grf = grf;
        st->gbf = gbf;
    }
}

// Highlight rewrite of one pixel (:1216-1241); returns whether it is blown.
LRE_DEV bool rewrite_px(float &r, float &g, float &b, float sat, float grf, float gbf) {
    g = (r / grf + b / gbf) * 0.5f;
    const float sp = (r < b) ? r : b;
    const float sl = (lre::min_std(sp, 1.f) - sat) / (sat);
    r = r * (1.f - sl) + (g * grf) * sl;
    b = b * (1.f - sl) + (g * gbf) * sl;
    g = (r / grf + b / gbf) * 0.5f;
    return g >= 0.95f;
}

// Stages the rewritten 66 x 18 window, takes the 8-neighbour mean for blown
// interior pixels, writes the tile and flags the LoCA candidates: highlight
// pixels at least two from the frame whose own values pass the LoCA tests that
// read only the pixel itself (the other tests read neighbours).
__global__ __launch_bounds__(HR_BLOCK) void hr_rewrite(const float *__restrict__ in, float *__restrict__ out,
                                                       int w, int h, float sat, const HrState *st,
                                                       unsigned char *__restrict__ flags,
                                                       unsigned *__restrict__ tile_count, int tcols) {
    __shared__ float sR[RW_SY][RW_SX], sG[RW_SY][RW_SX], sB[RW_SY][RW_SX];
    __shared__ unsigned char sW[RW_SY][RW_SX];  // bit 0 highlight, bit 1 blown
    // LoCA candidates per LoCA tile: the 16 rows lie in one band of 64, and
    // x + 3y spans fewer than 2 * HR_TW + 1 values, so at most 3 tiles
    __shared__ unsigned s_tc[3];
    const int x0 = blockIdx.x * RW_TX, y0 = blockIdx.y * RW_TY;
    const int k0 = (x0 + 3 * y0) / HR_TW;
    if (threadIdx.x < 3) s_tc[threadIdx.x] = 0u;
    const float grf = st->grf, gbf = st->gbf;
    for (int e = threadIdx.x; e < RW_SY * RW_SX; e += HR_BLOCK) {
        const int sy = e / RW_SX, sx = e - sy * RW_SX;
        const int gy = y0 - 1 + sy, gx = x0 - 1 + sx;
        float r = 0.f, g = 0.f, b = 0.f;
        unsigned char f = 0;
        if (gy >= 0 && gy < h && gx >= 0 && gx < w) {
            const size_t i = (size_t)gy * w + gx;
            r = in[3 * i];
            g = in[3 * i + 1];
            b = in[3 * i + 2];
            if (r >= sat && b >= sat) f = rewrite_px(r, g, b, sat, grf, gbf) ? 3 : 1;
        }
        sR[sy][sx] = r;
        sG[sy][sx] = g;
        sB[sy][sx] = b;
        sW[sy][sx] = f;
    }
    __syncthreads();
    const float sat09 = sat * 0.90f;
    for (int e = threadIdx.x; e < RW_TX * RW_TY; e += HR_BLOCK) {
        const int ty = e / RW_TX, tx = e - ty * RW_TX;
        const int gy = y0 + ty, gx = x0 + tx;
        if (gy >= h || gx >= w) continue;
        const int sy = ty + 1, sx = tx + 1;
        float v[3] = {sR[sy][sx], sG[sy][sx], sB[sy][sx]};
        const unsigned f = sW[sy][sx];
        if ((f & 2) && gx >= 1 && gx <= w - 2 && gy >= 1 && gy <= h - 2) {
            // neighbours summed in the reference's order, then times 1/8
            auto mean8 = [&](const float(&S)[RW_SY][RW_SX]) {
                return (S[sy - 1][sx - 1] + S[sy - 1][sx] + S[sy - 1][sx + 1] + S[sy][sx - 1] + S[sy][sx + 1] +
                        S[sy + 1][sx - 1] + S[sy + 1][sx] + S[sy + 1][sx + 1]) *
                       (1.f / 8.f);
            };
            v[0] = mean8(sR);
            v[1] = mean8(sG);
            v[2] = mean8(sB);
        }
        const size_t i = (size_t)gy * w + gx;
        out[3 * i] = v[0];
        out[3 * i + 1] = v[1];
        out[3 * i + 2] = v[2];
        const float l = max3(v[0], v[1], v[2]);
        const float rb = 0.5f * (v[0] + v[2]);
        const bool cand = (f & 1) && gx >= 2 && gx <= w - 3 && gy >= 2 && gy <= h - 3 && !(l < sat09) &&
                          !(rb <= v[1] * 1.04f);
        flags[i] = cand;
        if (cand) atomicAdd(&s_tc[(gx + 3 * gy) / HR_TW - k0], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 3 && s_tc[threadIdx.x])
        atomicAdd(&tile_count[(y0 / HR_TR) * tcols + k0 + (int)threadIdx.x], s_tc[threadIdx.x]);
}

// What a pixel contributes to a LoCA window (:1311-1345): NL = max(R, G, B)
// for the 3x3 contrast, RT = G / ((R + B) / 2) when the 5x5 ratio mean takes
// it, else 0 (no taken ratio is 0: those lie in [0.70, 1.35] or are NaN).
// Both depend only on the pixel's own values, so the window reads them
// instead of redoing the division for every tap, and a skipped tap adds
// 0 * w = +0 to the ratio sum as it stands.

LRE_DEV void tap_info(float r, float g, float b, float sat075, float &nl, float &rt) {
    nl = max3(r, g, b);
    const float nrb = 0.5f * (r + b);
    rt = 0.0f;
    if (nl < sat075 || nrb <= 1e-6f) return;
    const float ratio = g / nrb;
    if (ratio < 0.70f || ratio > 1.35f) return;
    rt = ratio;
}

struct LocaTile {
    float *NL, *RT;        // [HR_SR][HR_LW], skewed: pixel (r, s) of the tile at (r + 2) * HR_LW + s + 8
    float *OR, *OG, *OB;   // the tile's own pixels [HR_TR][HR_OW]
};

// LoCA green lift of the tile pixel (r, s) (:1299-1372).  Neighbour (dy, dx)
// sits at c + dy * (HR_LW + 3) + dx of the skewed planes.  Every LDS read of
// the step is issued first, so the step waits for the LDS once.  The pixel is
// a candidate: its own values, which no earlier step changes, already pass
// l >= 0.9 sat and rb > 1.04 G (:1308-1310, :1327-1330).
// A changed pixel is stored to the image at once (index gi): no later step of
// this tile reads it from there, and later tiles stage it after this launch.
LRE_DEV void loca_px(const LocaTile &L, int r, int s, float sat, float sat075, float max_val, float *img,
                     size_t gi) {
    const int o = r * HR_OW + s, c = (r + 2) * HR_LW + s + 8;
    constexpr int DY = HR_LW + 3;
    const float p0 = L.OR[o], p1 = L.OG[o], p2 = L.OB[o];
    float nl[9], rt[25];
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
        for (int dx = -1; dx <= 1; ++dx) nl[(dy + 1) * 3 + dx + 1] = L.NL[c + dy * DY + dx];
#pragma unroll
    for (int dy = -2; dy <= 2; ++dy)
#pragma unroll
        for (int dx = -2; dx <= 2; ++dx) rt[(dy + 2) * 5 + dx + 2] = L.RT[c + dy * DY + dx];
    const float l = max3(p0, p1, p2);
    // The reference's running max from 0 and min from FLT_MAX never take a
    // NaN (a comparison with NaN is false), so they are the maximum and
    // minimum over the numbers and either order gives them: a tree of the
    // hardware max/min, which skip NaN.  (Only the sign of a zero extreme can
    // differ, and a zero spread returns below either way.)
    float mx4[4], mn4[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        mx4[k] = fmaxf(nl[2 * k], nl[2 * k + 1]);
        mn4[k] = fminf(nl[2 * k], nl[2 * k + 1]);
    }
    const float local_max = fmaxf(fmaxf(fmaxf(mx4[0], mx4[1]), fmaxf(mx4[2], mx4[3])), fmaxf(nl[8], 0.0f));
    const float local_min = fminf(fminf(fminf(mn4[0], mn4[1]), fminf(mn4[2], mn4[3])), fminf(nl[8], FLT_MAX));
    if ((local_max - local_min) < 0.015f) return;
    const float rb = 0.5f * (p0 + p2);
    // The reference skips a tap; adding +0.0 instead leaves both sums unchanged
    // (no taken term is -0.0).  The sums run in the reference's order.
    double ratio_sum = 0.0, ratio_w = 0.0;
#pragma unroll
    for (int dy = -2; dy <= 2; ++dy)
#pragma unroll
        for (int dx = -2; dx <= 2; ++dx) {
            const float t = rt[(dy + 2) * 5 + dx + 2];
            const double wt = 1.0 / (double)(1 + (dx < 0 ? -dx : dx) + (dy < 0 ? -dy : dy));
            ratio_sum += t * wt;
            ratio_w += t != 0.0f ? wt : 0.0;
        }
    const float local_ratio = (float)(ratio_w > 0.0 ? ratio_sum / ratio_w : 1.0);
    const float desired_g = rb * lre::clamp_std(local_ratio, 0.95f, 1.08f);
    const float g_missing = desired_g - p1;
    if (g_missing <= 0.0f) return;
    const float sat_w =
        lre::clamp_std((l - sat) / lre::max_std(1e-6f, lre::max_std(1.0f, max_val) - sat), 0.0f, 1.0f);
    const float edge_w = lre::clamp_std((local_max - local_min) / 0.08f, 0.0f, 1.0f);
    const float alpha = lre::clamp_std(0.30f + 0.55f * sat_w * edge_w, 0.0f, 0.90f);
    const float g_boost = alpha * lre::min_std(g_missing, (rb - p1) * 0.90f);
    float q0 = p0, q1 = p1 + g_boost, q2 = p2;
    const float rb_residual = lre::max_std(0.0f, 0.5f * (q0 + q2) - q1);
    if (rb_residual > 0.0f) {
        const float trim = lre::min_std(rb_residual * alpha * 0.45f, 0.5f * (q0 + q2) * 0.12f);
        q0 = lre::max_std(0.0f, q0 - trim);
        q2 = lre::max_std(0.0f, q2 - trim);
    }
    q1 = lre::max_std(0.0f, q1);
    L.OR[o] = q0;
    L.OG[o] = q1;
    L.OB[o] = q2;
    tap_info(q0, q1, q2, sat075, L.NL[c], L.RT[c]);
    img[3 * gi] = q0;
    img[3 * gi + 1] = q1;
    img[3 * gi + 2] = q2;
}

// LDS writes of this wavefront's lanes are seen by its later LDS reads (the
// wave runs in order; this keeps the compiler from moving a read of another
// lane's cell above that lane's write)
LRE_DEV void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Tile (b_lo + blockIdx.x, d - b) of the anti-diagonal d.  Staged row sr holds
// image row Y - 2 + sr, column j holds x = T - 3y - 8 + j, so the pixel of tile
// row r at step s is at (r + 2, s + 8) for every row.  Lane r of wave 0 takes
// tile row r; the other waves stage, then keep step with the barriers.
__global__ __launch_bounds__(HR_BLOCK) void hr_loca(float *img, int w, int h, float sat, const HrState *st,
                                                    const unsigned char *__restrict__ flags,
                                                    const unsigned *__restrict__ tile_count, int tcols,
                                                    int d, int b_lo) {
    extern __shared__ __align__(16) float lds[];
    LocaTile L;
    L.NL = lds;
    L.RT = L.NL + HR_SR * HR_LW;
    L.OR = L.RT + HR_SR * HR_LW;
    L.OG = L.OR + HR_TR * HR_OW;
    L.OB = L.OG + HR_TR * HR_OW;
    unsigned char *F = reinterpret_cast<unsigned char *>(L.OB + HR_TR * HR_OW);
    unsigned *step_any = reinterpret_cast<unsigned *>(F + HR_TR * HR_TW);
    const int b = b_lo + (int)blockIdx.x, k = d - b;
    if (tile_count[b * tcols + k] == 0u) return;
    const int Y = b * HR_TR, T = k * HR_TW;
    const float sat075 = sat * 0.75f;
    if (threadIdx.x < HR_WORDS) step_any[threadIdx.x] = 0u;
    // 8 global loads in flight per thread (lre::stage2d)
    lre::stage2d<HR_BLOCK>(
        HR_SR, HR_LW,
        [&](int sr, int j) {
            const int y = Y - 2 + sr, x = T - 3 * y - 8 + j;
            float3 v = make_float3(0.f, 0.f, 0.f);
            if (y >= 0 && y < h && x >= 0 && x < w) {
                const size_t i = (size_t)y * w + x;
                v = make_float3(img[3 * i], img[3 * i + 1], img[3 * i + 2]);
            }
            return v;
        },
        [&](int sr, int j, float3 v) {
            const int e = sr * HR_LW + j;
            tap_info(v.x, v.y, v.z, sat075, L.NL[e], L.RT[e]);
            const int tr = sr - 2, ts = j - 8;
            if (tr >= 0 && tr < HR_TR && ts >= 0 && ts < HR_TW) {
                L.OR[tr * HR_OW + ts] = v.x;
                L.OG[tr * HR_OW + ts] = v.y;
                L.OB[tr * HR_OW + ts] = v.z;
            }
        });
    lre::stage2d<HR_BLOCK>(
        HR_TR, HR_TW,
        [&](int r, int s) {
            const int y = Y + r, x = T + s - 3 * y;
            return (y < h && x >= 0 && x < w) ? flags[(size_t)y * w + x] : (unsigned char)0;
        },
        [&](int r, int s, unsigned char f) { F[r * HR_TW + s] = f; });
    __syncthreads();
    // each row's flags as bits, one word per thread (HR_TR x HR_WORDS threads),
    // and their union over the rows per step
    unsigned *msk = step_any + HR_WORDS;
    static_assert(HR_TR * HR_WORDS <= HR_BLOCK, "one mask word per thread");
    if (threadIdx.x < HR_TR * HR_WORDS) {
        const int rr = threadIdx.x / HR_WORDS, wd = threadIdx.x % HR_WORDS;
        unsigned m = 0u;
        for (int q = 0; q < 32 && wd * 32 + q < HR_TW; q++) m |= (unsigned)(F[rr * HR_TW + wd * 32 + q] != 0) << q;
        msk[threadIdx.x] = m;
        if (m) atomicOr(&step_any[wd], m);
    }
    __syncthreads();
    // One wavefront runs the steps: lane r takes tile row r, and only the LDS
    // orders one step's writes before the next step's reads.
    if (threadIdx.x >= 64) return;
    const float max_val = __uint_as_float(st->max_bits);
    const int r = threadIdx.x;
    const int y = Y + r;
    unsigned fb[HR_WORDS];
#pragma unroll
    for (int wd = 0; wd < HR_WORDS; wd++) fb[wd] = msk[r * HR_WORDS + wd];
#pragma unroll
    for (int wd = 0; wd < HR_WORDS; wd++) {
        unsigned bits = step_any[wd];
        while (bits) {
            const int q = __ffs(bits) - 1;
            bits &= bits - 1;
            const int sidx = wd * 32 + q;
            if ((fb[wd] >> q) & 1u)
                loca_px(L, r, sidx, sat, sat075, max_val, img, (size_t)y * w + (T + sidx - 3 * y));
            wave_lds_sync();
        }
    }
}

__global__ void hr_emit(const HrState *st, float *max_val) { *max_val = __uint_as_float(st->max_bits); }

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

static_assert(HR_LDS <= 160 * 1024, "LoCA tile exceeds the LDS of a workgroup");
static_assert(HR_TR % RW_TY == 0 && RW_TX - 1 + 3 * (RW_TY - 1) < 2 * HR_TW,
              "a rewrite tile must fall in one LoCA band and at most 3 LoCA tile columns");

struct HrLayout {
    int nchunks, bands, tcols;
    size_t o_counts, o_tiles, o_flags, o_ratios, bytes;
};

HrLayout hr_layout(int w, int h) {
    HrLayout L;
    const size_t n = (size_t)w * h;
    L.nchunks = (int)((n + HR_CHUNK - 1) / HR_CHUNK);
    L.bands = (h + HR_TR - 1) / HR_TR;
    const long long tspan = (long long)w + 3LL * (h - 1);  // t = x + 3y < w + 3(h - 1)
    L.tcols = (int)((tspan + HR_TW - 1) / HR_TW);
    L.o_counts = 256;
    L.o_tiles = L.o_counts + align256((size_t)L.nchunks * 4);
    L.o_flags = L.o_tiles + align256((size_t)L.bands * L.tcols * 4);
    L.o_ratios = L.o_flags + align256(n);
    L.bytes = L.o_ratios + (size_t)L.nchunks * HR_CHUNK * sizeof(float2);
    return L;
}

hipError_t hr_lds_attr() {
    static const hipError_t rc =
        hipFuncSetAttribute((const void *)hr_loca, hipFuncAttributeMaxDynamicSharedMemorySize, (int)HR_LDS);
    return rc;
}

}  // namespace

size_t lre_highlight_scratch_bytes(int w, int h) { return hr_layout(w, h).bytes; }

int lre_launch_recover_highlights(const float *d_in, float *d_out, int w, int h, float sat, float *d_mask,
                                  const float *wb_rgb, float *d_max_val, void *d_scratch, hipStream_t s) {
    if (d_in == d_out) {
        lre_set_error("recover_highlights launcher needs distinct input and output");
        return LRE_ERR_ARG;
    }
    const HrLayout L = hr_layout(w, h);
    const size_t n = (size_t)w * h;
    char *base = static_cast<char *>(d_scratch);
    HrState *st = reinterpret_cast<HrState *>(base);
    unsigned *chunk_count = reinterpret_cast<unsigned *>(base + L.o_counts);
    unsigned *tile_count = reinterpret_cast<unsigned *>(base + L.o_tiles);
    unsigned char *flags = reinterpret_cast<unsigned char *>(base + L.o_flags);
    float2 *ratios = reinterpret_cast<float2 *>(base + L.o_ratios);
    // mask edges and the WB undo, in float as the reference computes them (:1135-1147)
    const float mask_t_hi = 1.00f, mask_t_lo = 0.95f;
    MaskConst M;
    M.lo = mask_t_lo;
    M.inv = 1.f / std::max(1e-6f, mask_t_hi - mask_t_lo);
    M.chroma_lo = wb_rgb ? 0.30f : 0.50f;
    const float chroma_hi = wb_rgb ? 0.60f : 0.70f;
    M.chroma_inv = 1.f / std::max(1e-6f, chroma_hi - M.chroma_lo);
    M.inv_wb_r = wb_rgb ? 1.f / std::max(1e-6f, wb_rgb[0]) : 1.f;
    M.inv_wb_g = wb_rgb ? 1.f / std::max(1e-6f, wb_rgb[1]) : 1.f;
    M.inv_wb_b = wb_rgb ? 1.f / std::max(1e-6f, wb_rgb[2]) : 1.f;
    const int ntiles = L.bands * L.tcols;
    LRE_HIP_CHECK(hr_lds_attr());
    hipLaunchKernelGGL(hr_init, dim3(std::min(64, (ntiles + 255) / 256 + 1)), dim3(256), 0, s, st, tile_count,
                       ntiles);
    hipLaunchKernelGGL(hr_scan, dim3(L.nchunks), dim3(HR_BLOCK), 0, s, d_in, d_mask, n, sat, M, ratios, chunk_count,
                       st);
    hipLaunchKernelGGL(hr_sums, dim3(1), dim3(64), 0, s, ratios, chunk_count, L.nchunks, st);
    hipLaunchKernelGGL(hr_rewrite, dim3((w + RW_TX - 1) / RW_TX, (h + RW_TY - 1) / RW_TY), dim3(HR_BLOCK), 0, s,
                       d_in, d_out, w, h, sat, st, flags, tile_count, L.tcols);
    LRE_HIP_CHECK(hipGetLastError());
    for (int d = 0; d <= L.bands + L.tcols - 2; d++) {
        const int b_lo = std::max(0, d - L.tcols + 1), b_hi = std::min(L.bands - 1, d);
        hipLaunchKernelGGL(hr_loca, dim3(b_hi - b_lo + 1), dim3(HR_BLOCK), HR_LDS, s, d_out, w, h, sat, st, flags,
                           tile_count, L.tcols, d, b_lo);
    }
    LRE_HIP_CHECK(hipGetLastError());
    if (d_max_val) hipLaunchKernelGGL(hr_emit, dim3(1), dim3(1), 0, s, st, d_max_val);
    LRE_HIP_CHECK(hipGetLastError());
    return LRE_OK;
}
