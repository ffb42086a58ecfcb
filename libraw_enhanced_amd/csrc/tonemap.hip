// Pointwise and box-filter stages for gfx950: ACES tone map, the
// detail-preserving tone map, white balance and CFA-plane extraction.
#include "lre_common.h"
#include "lre_epilogue.h"
#include "lre_internal.h"

namespace {

// ---- ACES x after_scale, core/cpu_accelerator.cpp:3034-3068 ----------------
__global__ void aces_kernel(const float4 *__restrict__ in, float4 *__restrict__ out, size_t n4,
                            const float *__restrict__ in_tail, float *__restrict__ out_tail,
                            int ntail, float after_scale) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n4; i += stride) {
        float4 v = in[i];
        v.x = lre::aces(v.x) * after_scale;
        v.y = lre::aces(v.y) * after_scale;
        v.z = lre::aces(v.z) * after_scale;
        v.w = lre::aces(v.w) * after_scale;
        out[i] = v;
    }
    if (blockIdx.x == 0 && (int)threadIdx.x < ntail)
        out_tail[threadIdx.x] = lre::aces(in_tail[threadIdx.x]) * after_scale;
}

// ---- detail-preserving tone map, core/libraw_wrapper.cpp:1398-1471 ---------
// Box means of radius 12 with the window clipped at the frame and divided by
// the true tap count, rows first then columns (core/libraw_wrapper.cpp:45-116),
// evaluated as the reference does: a running sum that drops the leaving tap
// and adds the entering one.  Each thread restarts the sum for its own run of
// outputs, so the rounding differs from the reference's whole-row runs by
// ~1e-7 relative.
//
// Two kernels.  dtm_guide writes the guide plane max(0, R, G, B).  dtm_tile
// takes a 64 x 64 output tile and stages the guide over the tile and a 24-px
// ring (the reach of two chained radius-12 boxes) in LDS; the means of the
// guide and its square, a and b, their means and the per-pixel gain all stay
// on chip.  Around the tile only the guide plane is read, so the image itself
// can be rewritten in place.  A tile whose staged guide never exceeds 0.5499
// is left as it is: there base <= max(guide) < 0.55 (b is a mean times 1 - a,
// so a g + b is a convex combination of values no larger than that maximum),
// the smoothstep weight is 0 and the reference leaves every pixel unchanged.
constexpr int DR = 12;
constexpr float DTM_DARK = 0.5499f;

// one image of a detail tone map launch
struct DtmImg {
    const float *guide;
    const float *in;
    float *out;
    int W, H;
};

// Geometry of an output tile of side T: the guide window (side T + 4 DR), the
// mean / a / b window (side T + 2 DR), odd LDS pitches, and LDS in floats: a
// and b first (the guide window shares their space; it is dead when they are
// written), then the row means of g and g*g, whose space later holds the row
// means of a and b.  Each thread sums runs of consecutive outputs: the row
// passes a line of M = T + 2 DR outputs in 8 runs and a line of T outputs in 8
// runs, the column passes runs of RUNC and RUND outputs.
template <int T>
struct DtmGeo {
    static constexpr int G = T + 4 * DR, M = T + 2 * DR;
    static constexpr int GP = G + 1, MP = M + 1, TP = T + 1;
    static constexpr int RUN1 = M / 8, RUN2 = T / 8;
    // the column passes: pass 2 sums NC runs of RUNC rows per column, pass 4 ND
    // runs of RUND (shorter runs for more threads measured no faster)
    static constexpr int RUNC = RUN1, NC = 8, RUND = RUN2, ND = 8;
    static_assert(NC * RUNC == M && ND * RUND == T, "column runs tile the windows");
    static constexpr int L_AB = 2 * M * MP > G * GP ? 2 * M * MP : G * GP, L_H = 2 * G * MP;
    static constexpr size_t LDS = sizeof(float) * (L_AB + L_H);
    static_assert(RUN1 * 8 == M && RUN2 * 8 == T, "runs tile the windows");
    static_assert(G * GP <= L_AB && 2 * M * TP <= L_H, "LDS regions");
};

// One thread's run of N box means along a line: tap k of the line is val(k)
// (two quantities at once), output j covers taps j .. j + 2 DR and sits at
// image coordinate p0 + j of an extent len (taps outside the image are 0).
// The mean is the sum times 1 / (tap count): for a run whose windows all lie
// inside the image the count is 2 DR + 1 for every output (a constant), else
// it is clipped per output and inverted by v_rcp_f32.  The reference divides;
// the two differ by an ulp or so, well inside the running-sum rounding.
template <int N, class V, class F>
LRE_DEV void box_run2(V val, int p0, int len, F emit) {
    float s0 = 0.f, s1 = 0.f;
#pragma unroll
    for (int k = 0; k <= 2 * DR; k++) {
        const float2 v = val(k);
        s0 += v.x;
        s1 += v.y;
    }
    const auto slide = [&](int j) {
        const float2 o = val(j), e = val(j + 2 * DR + 1);
        s0 -= o.x;
        s1 -= o.y;
        s0 += e.x;
        s1 += e.y;
    };
    if (p0 >= DR && p0 + N - 1 + DR <= len - 1) {
        constexpr float ic = 1.0f / (2 * DR + 1);
#pragma unroll
        for (int j = 0; j < N; j++) {
            emit(j, s0 * ic, s1 * ic);
            if (j + 1 < N) slide(j);
        }
    } else {
#pragma unroll
        for (int j = 0; j < N; j++) {
            const int p = p0 + j;
            const int cnt = min(len - 1, p + DR) - max(0, p - DR) + 1;
            // cnt < 1 only outside the image, where the mean is not used
            const float ic = __builtin_amdgcn_rcpf((float)(cnt > 0 ? cnt : 1));
            emit(j, s0 * ic, s1 * ic);
            if (j + 1 < N) slide(j);
        }
    }
}

// the guide planes of several images, one chunk of DTM_CHUNK pixels per workgroup
constexpr int DTM_CHUNK = 256 * 16;
__global__ __launch_bounds__(256) void dtm_guide_many(lre::Many<DtmImg> src) {
    const lre::TileOf tl = src.tile();
    const DtmImg &I = src.args(tl);
    const size_t n = (size_t)I.W * I.H, b = (size_t)tl.tx * DTM_CHUNK;
    const size_t e = b + DTM_CHUNK < n ? b + DTM_CHUNK : n;
    for (size_t t = b + threadIdx.x; t < e; t += 256) {
        const float *p = I.in + 3 * t;
        const_cast<float *>(I.guide)[t] = lre::max_std(0.0f, lre::max_std(p[0], lre::max_std(p[1], p[2])));
    }
}

// n4: pixel quads read as three float4 (0 when the image is not 16-byte aligned)
__global__ void dtm_guide(const float *__restrict__ rgb, float *__restrict__ guide, size_t n, size_t n4) {
    const float4 *v = reinterpret_cast<const float4 *>(rgb);
    float4 *g4 = reinterpret_cast<float4 *>(guide);
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    auto gd = [](float r, float g, float b) { return lre::max_std(0.0f, lre::max_std(r, lre::max_std(g, b))); };
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) {
        const float4 a = v[3 * i], b = v[3 * i + 1], c = v[3 * i + 2];
        g4[i] = make_float4(gd(a.x, a.y, a.z), gd(a.w, b.x, b.y), gd(b.z, b.w, c.x), gd(c.y, c.z, c.w));
    }
    for (size_t t = n4 * 4 + (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride)
        guide[t] = gd(rgb[3 * t], rgb[3 * t + 1], rgb[3 * t + 2]);
}

// One output tile whose guide window is in LDS (dark: no window value exceeds
// the dark limit, the same for every thread).  Returns with every LDS read of
// the window's space done before a barrier that all threads passed, and the
// last pass reading only ha/hb.
template <int T, int NT, bool EPI>
LRE_DEV void dtm_tile_one(const DtmImg &I, int x0, int y0, bool dark, float *dsm, const lre::Epilogue &E) {
    using D = DtmGeo<T>;
    constexpr int G = D::G, M = D::M, GP = D::GP, MP = D::MP, TP = D::TP;
    float *gw = dsm;                               // G x G guide window (pitch GP)
    float *pa = dsm, *pb = dsm + M * MP;           // M x M a and b (pitch MP)
    float *hg = dsm + D::L_AB, *hq = hg + G * MP;  // G x M row means of g and g*g
    float *ha = dsm + D::L_AB, *hb = ha + M * TP;  // M x T row means of a and b
    const float *in = I.in;
    float *out = I.out;
    const int W = I.W, H = I.H;
    const int tid = threadIdx.x;
    // the reference leaves every pixel of a dark tile unchanged
    if (dark) {
        if (!EPI && in == out) return;
        // each thread reads and then writes its own pixels (safe in place)
        constexpr int PT = (T * T + NT - 1) / NT;
        float v[PT][3];
#pragma unroll
        for (int k = 0; k < PT; k++) {
            const int t = tid + k * NT, x = x0 + (t % T), y = y0 + (t / T);
            if (t < T * T && x < W && y < H) {
                const size_t i = ((size_t)y * W + x) * 3;
                v[k][0] = in[i];
                v[k][1] = in[i + 1];
                v[k][2] = in[i + 2];
            }
        }
#pragma unroll
        for (int k = 0; k < PT; k++) {
            const int t = tid + k * NT, x = x0 + (t % T), y = y0 + (t / T);
            if (t < T * T && x < W && y < H) {
                const size_t i = ((size_t)y * W + x) * 3;
                float r = v[k][0], g = v[k][1], b = v[k][2];
                if (EPI) lre::epilogue_rgb(r, g, b, E);
                out[i] = r;
                out[i + 1] = g;
                out[i + 2] = b;
            }
        }
        return;
    }
    // This thread's pixels of the last pass (a run of RUND rows of one column)
    // are loaded now, so that their latency passes behind the box means: loaded
    // in the last pass, each would wait for its own round trip between the
    // stores, which may alias the input (in place).
    static_assert(T * D::ND <= NT, "one run of the last pass per thread");
    const int c4 = tid % T, run4 = tid / T, x4 = x0 + c4;
    float px[D::RUND][3];
    if (tid < T * D::ND) {
#pragma unroll
        for (int j = 0; j < D::RUND; j++) {
            const int y = y0 + run4 * D::RUND + j;
            if (x4 < W && y < H) {
                const size_t i = ((size_t)y * W + x4) * 3;
                px[j][0] = in[i];
                px[j][1] = in[i + 1];
                px[j][2] = in[i + 2];
            }
        }
    }
    // row means of g and g*g: rows [0, G), output columns [DR, DR + M)
    for (int t = tid; t < G * 8; t += NT) {
        const int r = t % G, run = t / G;
        const float *v = gw + r * GP + run * D::RUN1;
        box_run2<D::RUN1>([&](int k) { const float g = v[k]; return make_float2(g, g * g); },
                          x0 - DR + run * D::RUN1, W, [&](int j, float m0, float m1) {
                              hg[r * MP + run * D::RUN1 + j] = m0;
                              hq[r * MP + run * D::RUN1 + j] = m1;
                          });
    }
    __syncthreads();
    // column means -> mean, mean of squares -> a, b over the M x M window
    for (int t = tid; t < M * D::NC; t += NT) {
        const int c = t % M, run = t / M;
        const int x = x0 - DR + c;
        const float *v0 = hg + run * D::RUNC * MP + c, *v1 = hq + run * D::RUNC * MP + c;
        box_run2<D::RUNC>([&](int k) { return make_float2(v0[k * MP], v1[k * MP]); }, y0 - DR + run * D::RUNC, H,
                          [&](int j, float mean, float msq) {
                              const int rr = run * D::RUNC + j, y = y0 - DR + rr;
                              float a = 0.f, b = 0.f;  // outside the image: no taps
                              if (x >= 0 && x < W && y >= 0 && y < H) {
                                  const float var = lre::max_std(0.0f, msq - mean * mean);
                                  a = var / (var + 0.0025f);
                                  b = mean * (1.0f - a);
                              }
                              pa[rr * MP + c] = a;
                              pb[rr * MP + c] = b;
                          });
    }
    __syncthreads();
    // row means of a and b: rows [0, M), output columns [DR, DR + T)
    for (int t = tid; t < M * 8; t += NT) {
        const int r = t % M, run = t / M;
        const float *v0 = pa + r * MP + run * D::RUN2, *v1 = pb + r * MP + run * D::RUN2;
        box_run2<D::RUN2>([&](int k) { return make_float2(v0[k], v1[k]); }, x0 + run * D::RUN2, W,
                          [&](int j, float m0, float m1) {
                              ha[r * TP + run * D::RUN2 + j] = m0;
                              hb[r * TP + run * D::RUN2 + j] = m1;
                          });
    }
    __syncthreads();
    // column means -> A, B; base = max(1e-6, A guide + B); RGB *= 1 + (gain - 1) w
    if (tid < T * D::ND) {
        const int c = c4, run = run4;
        const int x = x4;
        const float *v0 = ha + run * D::RUND * TP + c, *v1 = hb + run * D::RUND * TP + c;
        box_run2<D::RUND>([&](int k) { return make_float2(v0[k * TP], v1[k * TP]); }, y0 + run * D::RUND, H,
                          [&](int j, float ma, float mb) {
                              const int y = y0 + run * D::RUND + j;
                              if (x >= W || y >= H) return;
                              const size_t i = ((size_t)y * W + x) * 3;
                              float r = px[j][0], g = px[j][1], b = px[j][2];
                              const float gd = lre::max_std(0.0f, lre::max_std(r, lre::max_std(g, b)));
                              const float base = lre::max_std(1e-6f, ma * gd + mb);
                              float gain = lre::clamp_std(lre::aces(base) / base, 0.0f, 1.0f);
                              float w = lre::clamp_std((base - 0.55f) / (0.90f - 0.55f), 0.0f, 1.0f);
                              w = w * w * (3.0f - 2.0f * w);
                              if (w > 0.0f) {
                                  gain = 1.0f + (gain - 1.0f) * w;
                                  r *= gain;
                                  g *= gain;
                                  b *= gain;
                              } else if (!EPI && in == out) {
                                  return;  // unchanged in place: nothing to store
                              }
                              if (EPI) lre::epilogue_rgb(r, g, b, E);
                              out[i] = r;
                              out[i + 1] = g;
                              out[i + 2] = b;
                          });
    }
}

// The guide window of a tile, as NT threads hold it in registers: thread
// (c, r0) = (tid % G, tid / G) holds column c of rows r0, r0 + PER, ...
template <int T, int NT>
struct DtmWin {
    static constexpr int G = DtmGeo<T>::G, PER = NT / G, K = (G + PER - 1) / PER;
    static_assert(PER >= 1, "a window row fits the workgroup");
    float v[K];
    LRE_DEV void load(const DtmImg &I, int x0, int y0) {
        const int tid = threadIdx.x, c = tid % G, r0 = tid / G;
        const int x = x0 - 2 * DR + c;
        const bool xin = x >= 0 && x < I.W && r0 < PER;
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int r = r0 + k * PER, y = y0 - 2 * DR + r;
            v[k] = (xin && r < G && y >= 0 && y < I.H) ? I.guide[(size_t)y * I.W + x] : 0.f;
        }
    }
    // into the LDS window (pitch G + 1); false if any value exceeds the dark limit
    LRE_DEV bool store(float *gw) const {
        const int tid = threadIdx.x, c = tid % G, r0 = tid / G;
        bool dark = true;
        if (r0 < PER) {
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int r = r0 + k * PER;
                if (r < G) {
                    gw[r * (G + 1) + c] = v[k];
                    dark = dark && v[k] <= DTM_DARK;  // NaN is not dark
                }
            }
        }
        return dark;
    }
};

// Tile t of a launch: a flat index over the tiles of one image (row-major,
// tiles_x from the width) or of several (the table's binary search).
template <int T>
LRE_DEV lre::TileOf dtm_tile_at(const lre::One<DtmImg> &s, int t) {
    const int tx = (s.a.W + T - 1) / T;
    return {0, t % tx, t / tx};
}
template <int T>
LRE_DEV lre::TileOf dtm_tile_at(const lre::Many<DtmImg> &s, int t) { return s.tile_at(t); }

// The tiles of workgroup w, XCD by XCD: workgroups are dealt round-robin to the
// 8 XCDs, so the workgroups of one XCD (w % 8) take one contiguous range of
// tiles in row-major order between them, and neighbouring tiles, which share
// most of their guide windows, meet in that XCD's L2.
struct DtmOrder {
    int t, step, end;
    LRE_DEV DtmOrder(int ntiles) {
        constexpr int XCDS = 8;
        const int g = (int)gridDim.x, w = (int)blockIdx.x, x = w % XCDS;
        int before = 0;
        for (int k = 0; k < x; k++) before += (g - k + XCDS - 1) / XCDS;
        const int mine = (g - x + XCDS - 1) / XCDS;
        const long long b0 = (long long)ntiles * before / g, b1 = (long long)ntiles * (before + mine) / g;
        t = (int)b0 + w / XCDS;
        step = mine;
        end = (int)b1;
    }
};

// EPI: the colour matrix and gamma stages that follow in a stage list run on
// the result before it is stored (the same float operations as their own pass).
// Src: lre::One<DtmImg> (one image) or lre::Many<DtmImg> (the tiles of several
// images).  The grid is persistent: each workgroup walks its tiles (DtmOrder)
// and loads the next tile's guide window into registers while it works on the
// current one, so the window's latency passes behind the box means.
template <int T, int NT, bool EPI, class Src>
__global__ __launch_bounds__(NT) void dtm_tile(Src src, int ntiles, lre::Epilogue E) {
    using D = DtmGeo<T>;
    constexpr int G = D::G, M = D::M, MP = D::MP, TP = D::TP;
    static_assert(D::GP == G + 1, "window pitch");
    extern __shared__ float dsm[];
    float *gw = dsm;                               // G x G guide window (pitch GP)
    float *pa = dsm, *pb = dsm + M * MP;           // M x M a and b (pitch MP)
    float *hg = dsm + D::L_AB, *hq = hg + G * MP;  // G x M row means of g and g*g
    float *ha = dsm + D::L_AB, *hb = ha + M * TP;  // M x T row means of a and b
    const int tid = threadIdx.x;
    DtmOrder ord(ntiles);
    DtmWin<T, NT> win;
    if (ord.t < ord.end) {
        const lre::TileOf tl = dtm_tile_at<T>(src, ord.t);
        win.load(src.args(tl), tl.tx * T, tl.ty * T);
    }
    for (; ord.t < ord.end; ord.t += ord.step) {
        const lre::TileOf tl = dtm_tile_at<T>(src, ord.t);
        const DtmImg &I = src.args(tl);
        // The previous tile's last pass reads only ha/hb, which the window does
        // not overlap; its passes that read the window's space ended at a barrier.
        const bool dark = win.store(gw);
        if (ord.t + ord.step < ord.end) {
            const lre::TileOf nt = dtm_tile_at<T>(src, ord.t + ord.step);
            win.load(src.args(nt), nt.tx * T, nt.ty * T);
        }
        dtm_tile_one<T, NT, EPI>(I, tl.tx * T, tl.ty * T, __syncthreads_and(dark), dsm, E);
    }
}

// ---- white balance, core/cpu_accelerator.cpp:2813-2869 ---------------------
struct XTransTable { signed char v[36]; };

__global__ void wb_kernel(const ushort4 *__restrict__ img4, float *__restrict__ rgb, int W, int H,
                          float4 wb, uint32_t filters, XTransTable xt) {
    const size_t n = (size_t)W * H;
    for (lre::GridWalk gw((size_t)blockIdx.x * blockDim.x + threadIdx.x, n, (size_t)gridDim.x * blockDim.x, W);
         gw.more(); gw.next()) {
        const size_t i = gw.i;
        const unsigned row = (unsigned)gw.row, col = (unsigned)gw.col;
        unsigned c = filters == 9 ? (unsigned)xt.v[(row % 6) * 6 + col % 6]
                                  : lre::bayer_native(filters, row, col);
        const ushort4 s = img4[i];
        const unsigned short u = c == 0 ? s.x : c == 1 ? s.y : c == 2 ? s.z : s.w;
        const float m = c == 0 ? wb.x : c == 1 ? wb.y : c == 2 ? wb.z : wb.w;
        const float v = (float)u * m;
        if (c == 3) c = 1;
        rgb[i * 3 + c] = v;
    }
}

__global__ void rgb3_to_cfa_kernel(const float *__restrict__ rgb3, float *__restrict__ cfa, int W,
                                   int H, uint32_t filters, XTransTable xt) {
    const size_t n = (size_t)W * H;
    for (lre::GridWalk gw((size_t)blockIdx.x * blockDim.x + threadIdx.x, n, (size_t)gridDim.x * blockDim.x, W);
         gw.more(); gw.next()) {
        const size_t i = gw.i;
        const unsigned row = (unsigned)gw.row, col = (unsigned)gw.col;
        const unsigned c = filters == 9 ? (unsigned)xt.v[(row % 6) * 6 + col % 6]
                                        : lre::bayer3(filters, row, col);
        cfa[i] = rgb3[i * 3 + c];
    }
}

// The tile kernel at side T with NT threads over ntiles tiles, one persistent
// workgroup per CU.  64 x 64 tiles of 1024 threads (one workgroup per CU,
// 142 KiB of LDS) measured faster on 24 MP than 32 x 32 tiles of 512 (two per
// CU): 0.44 against 0.54 ms, the smaller tiles redoing half again as many box
// means in their rings.
template <int T, int NT, class Src>
int dtm_launch(Src src, int ntiles, hipStream_t s, const lre::Epilogue *epi) {
    if (ntiles < 1) return LRE_OK;
    const int cus = lre_current_cus();
    const dim3 grid(ntiles < cus ? ntiles : cus);
    static const hipError_t attr = [] {
        hipError_t e = hipFuncSetAttribute((const void *)dtm_tile<T, NT, false, Src>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)DtmGeo<T>::LDS);
        if (e == hipSuccess)
            e = hipFuncSetAttribute((const void *)dtm_tile<T, NT, true, Src>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)DtmGeo<T>::LDS);
        return e;
    }();
    LRE_HIP_CHECK(attr);
    if (epi && epi->active())
        hipLaunchKernelGGL((dtm_tile<T, NT, true, Src>), grid, dim3(NT), DtmGeo<T>::LDS, s, src, ntiles, *epi);
    else
        hipLaunchKernelGGL((dtm_tile<T, NT, false, Src>), grid, dim3(NT), DtmGeo<T>::LDS, s, src, ntiles,
                           lre::Epilogue{});
    LRE_HIP_CHECK(hipGetLastError());
    return LRE_OK;
}

int grid_for(size_t n, int block) {
    size_t b = (n + block - 1) / block;
    if (b > 65536) b = 65536;
    return (int)(b > 0 ? b : 1);
}


// ---- micro-contrast enhancement, core/cpu_accelerator.cpp:3135-3319 -------
// Local Gaussian mean and deviation (9x9 taps of sigma 5/3.7, coordinates
// clamped at the frame, :3077-3128), then a gain on the detail that grows as
// the local contrast (deviation over its image maximum) falls below the
// target.  Every output sums its 81 taps in the reference's (ky, kx) order, so
// the means are bit-identical to the reference's.
constexpr int MC_K = 9, MC_R = 4, MC_T = 32, MC_W = MC_T + 2 * MC_R, MC_RB = 4;
constexpr int MC_BLOCK = MC_T * (MC_T / MC_RB);  // one 4-row strip of one column per thread

struct McTaps {
    float k[MC_K * MC_K];
};

__global__ void mc_init(unsigned *stats) {
    if (threadIdx.x == 0) stats[0] = 0u;
}

// pass 1: means of I and I*I per channel, the deviation sqrt(max(var, 0)),
// and the image maximum of the deviations (atomicMax on the bits of a
// non-negative float)
__global__ __launch_bounds__(MC_BLOCK) void mc_stats_kernel(const float *__restrict__ in, float *__restrict__ mean3,
                                                            float *__restrict__ sd3, int W, int H, McTaps T,
                                                            unsigned *stats) {
    __shared__ float win[3][MC_W * MC_W];
    const int x0 = blockIdx.x * MC_T, y0 = blockIdx.y * MC_T;
    lre::stage2d<MC_BLOCK>(
        MC_W, MC_W,
        [&](int r, int c) {
            const int y = min(max(y0 - MC_R + r, 0), H - 1), x = min(max(x0 - MC_R + c, 0), W - 1);
            const float *p = in + ((size_t)y * W + x) * 3;
            return make_float3(p[0], p[1], p[2]);
        },
        [&](int r, int c, float3 v) {
            const int t = r * MC_W + c;
            win[0][t] = v.x;
            win[1][t] = v.y;
            win[2][t] = v.z;
        });
    __syncthreads();
    const int lx = threadIdx.x % MC_T, ly0 = (threadIdx.x / MC_T) * MC_RB;
    float a[MC_RB][3], q[MC_RB][3];
#pragma unroll
    for (int u = 0; u < MC_RB; u++)
#pragma unroll
        for (int c = 0; c < 3; c++) a[u][c] = q[u][c] = 0.f;
    // window row ky feeds the outputs u with tap row ky - u in [0, 9): each
    // output still takes its taps row by row, left to right
#pragma unroll
    for (int ky = 0; ky < MC_K + MC_RB - 1; ky++)
#pragma unroll
        for (int kx = 0; kx < MC_K; kx++) {
            float v[3];
#pragma unroll
            for (int c = 0; c < 3; c++) v[c] = win[c][(ly0 + ky) * MC_W + lx + kx];
#pragma unroll
            for (int u = 0; u < MC_RB; u++) {
                const int ty = ky - u;
                if (ty < 0 || ty >= MC_K) continue;
                const float wgt = T.k[ty * MC_K + kx];
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    a[u][c] += v[c] * wgt;
                    q[u][c] += (v[c] * v[c]) * wgt;
                }
            }
        }
    float m = 0.f;
    const int x = x0 + lx;
#pragma unroll
    for (int u = 0; u < MC_RB; u++) {
        const int y = y0 + ly0 + u;
        if (x >= W || y >= H) continue;
        const size_t i = ((size_t)y * W + x) * 3;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float sd = sqrtf(lre::max_std(q[u][c] - a[u][c] * a[u][c], 0.f));
            mean3[i + c] = a[u][c];
            sd3[i + c] = sd;
            m = lre::max_std(m, sd);
        }
    }
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    __shared__ float red[MC_BLOCK / 64];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float v = red[0];
        for (int k = 1; k < MC_BLOCK / 64; k++) v = fmaxf(v, red[k]);
        const unsigned u = __float_as_uint(v);
        if (u > __hip_atomic_load(stats, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(stats, u);
    }
}

// pass 2, pointwise (:3212-3318); mask == nullptr selects the gate on green
__global__ void mc_apply_kernel(const float *in, float *out, const float *__restrict__ mean3,
                                const float *__restrict__ sd3, const float *__restrict__ mask, size_t n,
                                float threshold, float strength, float target, const unsigned *stats) {
    const float max_sd = __uint_as_float(stats[0]);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float s0 = in[3 * i], s1 = in[3 * i + 1], s2 = in[3 * i + 2];
        if (!mask && !(s1 >= threshold)) {
            out[3 * i] = s0;
            out[3 * i + 1] = s1;
            out[3 * i + 2] = s2;
            continue;
        }
        const float s[3] = {s0, s1, s2};
        float e[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float cm = sd3[3 * i + c] / max_sd;
            const float f = (cm < target) ? strength * (target - cm) / target : 0.f;
            const float hf = s[c] - mean3[3 * i + c];
            e[c] = mean3[3 * i + c] + hf * (1.f + f);
        }
        float r = e[0], g = e[1], b = e[2];
        const float maxc = lre::max_std(r, lre::max_std(g, b));
        if (maxc > 1.f) {
            const float scale = 1.f / maxc;
            r *= scale;
            g *= scale;
            b *= scale;
        }
        if (mask) {
            const float wt = lre::clamp_std(mask[i], 0.f, 1.f);
            r = s0 + (r - s0) * wt;
            g = s1 + (g - s1) * wt;
            b = s2 + (b - s2) * wt;
        }
        out[3 * i] = r;
        out[3 * i + 1] = g;
        out[3 * i + 2] = b;
    }
}

}  // namespace

int lre_launch_tone_mapping(const float *d_in, float *d_out, size_t n, float after_scale,
                            hipStream_t s) {
    const size_t n4 = n / 4;
    const int tail = (int)(n - n4 * 4);
    hipLaunchKernelGGL(aces_kernel, dim3(grid_for(n4, 256)), dim3(256), 0, s,
                       reinterpret_cast<const float4 *>(d_in), reinterpret_cast<float4 *>(d_out),
                       n4, d_in + n4 * 4, d_out + n4 * 4, tail, after_scale);
    LRE_HIP_CHECK(hipGetLastError());
    return LRE_OK;
}

// the guide plane
size_t lre_detail_tm_scratch_bytes(int w, int h) { return (size_t)w * h * sizeof(float); }

int lre_launch_detail_tonemap(const float *d_in, float *d_out, int w, int h, void *d_scratch,
                              hipStream_t s, const lre::Epilogue *epi) {
    float *guide = static_cast<float *>(d_scratch);
    const size_t n = (size_t)w * h;
    const size_t n4 = (reinterpret_cast<uintptr_t>(d_in) % 16 == 0) ? n / 4 : 0;
    hipLaunchKernelGGL(dtm_guide, dim3(grid_for(n4 > 0 ? n4 : n, 256)), dim3(256), 0, s, d_in, guide, n, n4);
    constexpr int T = 64;
    const lre::One<DtmImg> one{DtmImg{guide, d_in, d_out, w, h}};
    return dtm_launch<T, 1024>(one, ((w + T - 1) / T) * ((h + T - 1) / T), s, epi);
}

// Many images: two tables (guide chunks, then tiles) over the same per-image
// arguments, each guide plane at the start of the image's scratch.
using DtmTable = lre::ManyTable<DtmImg>;
size_t lre_detail_tm_multi_table_bytes(int n) { return 2 * ((DtmTable::bytes(n) + 255) & ~(size_t)255); }

int lre_detail_tm_multi_fill(const lre_post_job *jobs, int n, void *h_table, int grids[2]) {
    const size_t half = (DtmTable::bytes(n) + 255) & ~(size_t)255;
    auto args = [&](int i) {
        return DtmImg{static_cast<const float *>(jobs[i].d_scratch), jobs[i].d_in, jobs[i].d_out, jobs[i].w, jobs[i].h};
    };
    grids[0] = DtmTable::fill(h_table, n, [&](int i) {
        return make_int2((int)(((size_t)jobs[i].w * jobs[i].h + DTM_CHUNK - 1) / DTM_CHUNK), 1);
    }, args);
    grids[1] = DtmTable::fill(static_cast<char *>(h_table) + half, n, [&](int i) {
        return make_int2((jobs[i].w + 63) / 64, (jobs[i].h + 63) / 64);
    }, args);
    return LRE_OK;
}

int lre_launch_detail_tm_multi(const void *d_table, int n, const int grids[2], hipStream_t s,
                               const lre::Epilogue *epi) {
    if (n == 0 || grids[1] == 0) return LRE_OK;
    const size_t half = (DtmTable::bytes(n) + 255) & ~(size_t)255;
    hipLaunchKernelGGL(dtm_guide_many, dim3(grids[0]), dim3(256), 0, s, DtmTable::view(d_table, n));
    LRE_HIP_CHECK(hipGetLastError());
    return dtm_launch<64, 1024>(DtmTable::view(static_cast<const char *>(d_table) + half, n), grids[1], s, epi);
}

int lre_launch_white_balance(const uint16_t *d_img4, float *d_rgb, int w, int h,
                             const float wb[4], uint32_t filters, const signed char *xtrans,
                             hipStream_t s) {
    XTransTable xt{};
    if (xtrans)
        for (int i = 0; i < 36; i++) xt.v[i] = xtrans[i];
    hipLaunchKernelGGL(wb_kernel, dim3(grid_for((size_t)w * h, 256)), dim3(256), 0, s,
                       reinterpret_cast<const ushort4 *>(d_img4), d_rgb, w, h,
                       make_float4(wb[0], wb[1], wb[2], wb[3]), filters, xt);
    LRE_HIP_CHECK(hipGetLastError());
    return LRE_OK;
}

int lre_launch_rgb3_to_cfa(const float *d_rgb3, float *d_cfa, int w, int h, uint32_t filters,
                           const signed char *xtrans, hipStream_t s) {
    XTransTable xt{};
    if (xtrans)
        for (int i = 0; i < 36; i++) xt.v[i] = xtrans[i];
    hipLaunchKernelGGL(rgb3_to_cfa_kernel, dim3(grid_for((size_t)w * h, 256)), dim3(256), 0, s,
                       d_rgb3, d_cfa, w, h, filters, xt);
    LRE_HIP_CHECK(hipGetLastError());
    return LRE_OK;
}

// the deviation and mean planes (3 floats per pixel each) and the maximum
size_t lre_micro_contrast_scratch_bytes(int w, int h) {
    return (((size_t)w * h * 12 + 255) & ~(size_t)255) * 2 + 256;
}

int lre_launch_micro_contrast(const float *d_in, float *d_out, int w, int h, float threshold,
                              float strength, float target_contrast, const float *d_mask,
                              void *d_scratch, hipStream_t s) {
    const size_t n = (size_t)w * h, plane = ((n * 12 + 255) & ~(size_t)255);
    float *mean3 = static_cast<float *>(d_scratch);
    float *sd3 = reinterpret_cast<float *>(static_cast<char *>(d_scratch) + plane);
    unsigned *stats = reinterpret_cast<unsigned *>(static_cast<char *>(d_scratch) + 2 * plane);
    // create_gaussian_kernel(9, 5/3.7) (:3077-3099), in float with the C
    // library's expf, weights summed in row order and divided by the sum
    McTaps T;
    {
        const float sigma = 5.f / 3.7f;
        float sum = 0.f;
        for (int i = 0; i < MC_K; i++)
            for (int j = 0; j < MC_K; j++) {
                const float x = (float)(i - MC_R), y = (float)(j - MC_R);
                T.k[i * MC_K + j] = expf(-(x * x + y * y) / (2.0f * sigma * sigma));
                sum += T.k[i * MC_K + j];
            }
        for (int i = 0; i < MC_K * MC_K; i++) T.k[i] /= sum;
    }
    hipLaunchKernelGGL(mc_init, dim3(1), dim3(64), 0, s, stats);
    hipLaunchKernelGGL(mc_stats_kernel, dim3((w + MC_T - 1) / MC_T, (h + MC_T - 1) / MC_T), dim3(MC_BLOCK), 0, s,
                       d_in, mean3, sd3, w, h, T, stats);
    hipLaunchKernelGGL(mc_apply_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, d_in, d_out, mean3, sd3, d_mask,
                       n, threshold, strength, target_contrast, stats);
    LRE_HIP_CHECK(hipGetLastError());
    return LRE_OK;
}
