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
// the true tap count, rows first then columns (core/libraw_wrapper.cpp:45-116).
// Each workgroup stages a (64+24) x (64+24) input window in LDS, sums 25 taps
// horizontally into LDS, then 25 taps vertically, finishing each output in
// the thread that sums it.  The reference's running sums are evaluated in a
// different order; the results agree to ~1e-6 relative.
constexpr int DR = 12;                  // radius
constexpr int DTX = 64, DTY = 64;       // output tile
constexpr int DIX = DTX + 2 * DR, DIY = DTY + 2 * DR;
constexpr int DIXP = DIX + 1, DTXP = DTX + 1;  // odd LDS row pitches
constexpr int DBLOCK = 512;

LRE_DEV int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Box means over LDS windows.  hpass: horizontal window sums of the DIY x DIX
// window `in` (pitch DIXP, zero outside the image) divided by the clipped tap
// count, for the DTX output columns, into hs0 (and, with SQ, the sums of the
// squared values into hs1); pitch DTXP.  vpass: the vertical sums of hs0/hs1,
// handed to f(row in tile, column in tile, mean0, mean1).  Each thread sums RB
// neighbouring outputs along
// the filter direction with a sliding window; every output adds its 25 taps in
// order.  Lanes walk down rows; the odd pitches put them in distinct LDS banks.
constexpr int RB = 4, NT = 2 * DR + 1;

template <bool SQ>
LRE_DEV void hpass(const float *in, float *hs0, float *hs1, int x0, int W) {
    for (int t = threadIdx.x; t < DIY * (DTX / RB); t += DBLOCK) {
        const int ry = t % DIY, cx = (t / DIY) * RB;
        const float *r0 = in + ry * DIXP + cx;
        float s0[RB], s1[RB], w[RB];
#pragma unroll
        for (int u = 0; u < RB; u++) { s0[u] = s1[u] = 0.f; w[u] = r0[u]; }
#pragma unroll
        for (int k = 0; k < NT; k++) {
#pragma unroll
            for (int u = 0; u < RB; u++) {
                s0[u] += w[u];
                if (SQ) s1[u] += w[u] * w[u];
            }
            if (k + 1 < NT) {
#pragma unroll
                for (int u = 0; u < RB - 1; u++) w[u] = w[u + 1];
                w[RB - 1] = r0[k + RB];
            }
        }
#pragma unroll
        for (int u = 0; u < RB; u++) {
            const int x = x0 + cx + u;
            const float cnt = (float)(min(W - 1, x + DR) - max(0, x - DR) + 1);
            hs0[ry * DTXP + cx + u] = s0[u] / cnt;
            if (SQ) hs1[ry * DTXP + cx + u] = s1[u] / cnt;
        }
    }
}

template <class F>
LRE_DEV void vpass(const float *hs0, const float *hs1, int y0, int H, F f) {
    for (int t = threadIdx.x; t < (DTY / RB) * DTX; t += DBLOCK) {
        const int cx = t % DTX, oy = (t / DTX) * RB;
        const float *c0 = hs0 + oy * DTXP + cx, *c1 = hs1 + oy * DTXP + cx;
        float s0[RB], s1[RB], w0[RB], w1[RB];
#pragma unroll
        for (int u = 0; u < RB; u++) { s0[u] = s1[u] = 0.f; w0[u] = c0[u * DTXP]; w1[u] = c1[u * DTXP]; }
#pragma unroll
        for (int k = 0; k < NT; k++) {
#pragma unroll
            for (int u = 0; u < RB; u++) { s0[u] += w0[u]; s1[u] += w1[u]; }
            if (k + 1 < NT) {
#pragma unroll
                for (int u = 0; u < RB - 1; u++) { w0[u] = w0[u + 1]; w1[u] = w1[u + 1]; }
                w0[RB - 1] = c0[(k + RB) * DTXP];
                w1[RB - 1] = c1[(k + RB) * DTXP];
            }
        }
#pragma unroll
        for (int u = 0; u < RB; u++) {
            const int y = y0 + oy + u;
            const float cnt = (float)(min(H - 1, y + DR) - max(0, y - DR) + 1);
            f(oy + u, cx, s0[u] / cnt, s1[u] / cnt);
        }
    }
}


// pass 1: guide = max(0,R,G,B); a = var/(var+eps), b = mean(1-a) -> planes
__global__ __launch_bounds__(DBLOCK) void dtm_ab_kernel(const float *__restrict__ rgb,
                                                        float *__restrict__ pa,
                                                        float *__restrict__ pb, int W, int H) {
    extern __shared__ float dsm[];
    float *win = dsm, *hs0 = win + DIY * DIXP, *hs1 = hs0 + DIY * DTXP;
    const int x0 = blockIdx.x * DTX, y0 = blockIdx.y * DTY;
    lre::stage2d<DBLOCK>(
        DIY, DIX,
        [&](int r, int c) {
            const int y = y0 - DR + r, x = x0 - DR + c;
            float g = 0.f;
            if (x >= 0 && x < W && y >= 0 && y < H) {
                const float *p = rgb + ((size_t)y * W + x) * 3;
                g = lre::max_std(0.0f, lre::max_std(p[0], lre::max_std(p[1], p[2])));
            }
            return g;
        },
        [&](int r, int c, float g) { win[r * DIXP + c] = g; });
    __syncthreads();
    hpass<true>(win, hs0, hs1, x0, W);  // sums of g and of g * g
    __syncthreads();
    vpass(hs0, hs1, y0, H, [&](int ty, int tx, float mean, float msq) {
        const int x = x0 + tx, y = y0 + ty;
        if (x >= W || y >= H) return;
        const float var = lre::max_std(0.0f, msq - mean * mean);
        const float a = var / (var + 0.0025f);
        pa[(size_t)y * W + x] = a;
        pb[(size_t)y * W + x] = mean * (1.0f - a);
    });
}

// pass 2: box(a), box(b); base = max(1e-6, A*guide + B); RGB *= 1 + (gain-1) w.
// EPI: the colour matrix and gamma stages that follow in a stage list run on
// the result before it is stored (the same float operations as their own pass).
template <bool EPI>
__global__ __launch_bounds__(DBLOCK) void dtm_apply_kernel(const float *__restrict__ pa,
                                                           const float *__restrict__ pb,
                                                           const float *in, float *out, int W,
                                                           int H, lre::Epilogue E) {
    extern __shared__ float dsm[];
    float *win = dsm, *hs0 = win + DIY * DIXP, *hs1 = hs0 + DIY * DTXP;
    const int x0 = blockIdx.x * DTX, y0 = blockIdx.y * DTY;
    // the a and b windows pass through the same LDS buffer in turn
    for (int pl = 0; pl < 2; pl++) {
        const float *src = pl ? pb : pa;
        lre::stage2d<DBLOCK>(
            DIY, DIX,
            [&](int r, int c) {
                const int y = y0 - DR + r, x = x0 - DR + c;
                return (x >= 0 && x < W && y >= 0 && y < H) ? src[(size_t)y * W + x] : 0.f;
            },
            [&](int r, int c, float v) { win[r * DIXP + c] = v; });
        __syncthreads();
        hpass<false>(win, pl ? hs1 : hs0, nullptr, x0, W);
        __syncthreads();
    }
    vpass(hs0, hs1, y0, H, [&](int ty, int tx, float ma, float mb) {
        const int x = x0 + tx, y = y0 + ty;
        if (x >= W || y >= H) return;
        const size_t i = ((size_t)y * W + x) * 3;
        float r = in[i], g = in[i + 1], b = in[i + 2];
        const float guide = lre::max_std(0.0f, lre::max_std(r, lre::max_std(g, b)));
        const float base = lre::max_std(1e-6f, ma * guide + mb);
        float gain = lre::clamp_std(lre::aces(base) / base, 0.0f, 1.0f);
        float w = lre::clamp_std((base - 0.55f) / (0.90f - 0.55f), 0.0f, 1.0f);
        w = w * w * (3.0f - 2.0f * w);
        if (w > 0.0f) {
            gain = 1.0f + (gain - 1.0f) * w;
            r *= gain;
            g *= gain;
            b *= gain;
        }
        if (EPI) lre::epilogue_rgb(r, g, b, E);
        out[i] = r;
        out[i + 1] = g;
        out[i + 2] = b;
    });
}

// ---- white balance, core/cpu_accelerator.cpp:2813-2869 ---------------------
struct XTransTable { signed char v[36]; };

__global__ void wb_kernel(const ushort4 *__restrict__ img4, float *__restrict__ rgb, int W, int H,
                          float4 wb, uint32_t filters, XTransTable xt) {
    const size_t n = (size_t)W * H;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const unsigned row = (unsigned)(i / W), col = (unsigned)(i % W);
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
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const unsigned row = (unsigned)(i / W), col = (unsigned)(i % W);
        const unsigned c = filters == 9 ? (unsigned)xt.v[(row % 6) * 6 + col % 6]
                                        : lre::bayer3(filters, row, col);
        cfa[i] = rgb3[i * 3 + c];
    }
}

constexpr size_t DTM_LDS = sizeof(float) * (DIY * DIXP + 2 * DIY * DTXP);

// raise the dynamic LDS limit of both box-filter kernels once per process
hipError_t dtm_lds_attr() {
    static const hipError_t rc = [] {
        hipError_t e = hipFuncSetAttribute((const void *)dtm_ab_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)DTM_LDS);
        if (e == hipSuccess)
            e = hipFuncSetAttribute((const void *)dtm_apply_kernel<false>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)DTM_LDS);
        if (e == hipSuccess)
            e = hipFuncSetAttribute((const void *)dtm_apply_kernel<true>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)DTM_LDS);
        return e;
    }();
    return rc;
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

size_t lre_detail_tm_scratch_bytes(int w, int h) { return (size_t)w * h * sizeof(float) * 2; }

int lre_launch_detail_tonemap(const float *d_in, float *d_out, int w, int h, void *d_scratch,
                              hipStream_t s, const lre::Epilogue *epi) {
    float *pa = static_cast<float *>(d_scratch);
    float *pb = pa + (size_t)w * h;
    dim3 grid((w + DTX - 1) / DTX, (h + DTY - 1) / DTY);
    LRE_HIP_CHECK(dtm_lds_attr());
    hipLaunchKernelGGL(dtm_ab_kernel, grid, dim3(DBLOCK), DTM_LDS, s, d_in, pa, pb, w, h);
    LRE_HIP_CHECK(hipGetLastError());
    if (epi && epi->active())
        hipLaunchKernelGGL(dtm_apply_kernel<true>, grid, dim3(DBLOCK), DTM_LDS, s, pa, pb, d_in, d_out, w, h, *epi);
    else
        hipLaunchKernelGGL(dtm_apply_kernel<false>, grid, dim3(DBLOCK), DTM_LDS, s, pa, pb, d_in, d_out, w, h,
                           lre::Epilogue{});
    LRE_HIP_CHECK(hipGetLastError());
    return LRE_OK;
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
