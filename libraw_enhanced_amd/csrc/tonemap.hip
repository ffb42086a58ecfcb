// Pointwise and box-filter stages for gfx950: ACES tone map, the
// detail-preserving tone map, white balance and CFA-plane extraction.
#include "lre_common.h"
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
// Each workgroup stages a (64+24) x (32+24) input window in LDS, sums 25 taps
// horizontally into LDS, then 25 taps vertically.  The reference's running
// sums are evaluated in a different order; the results agree to ~1e-6
// relative.
constexpr int DR = 12;                  // radius
constexpr int DTX = 64, DTY = 32;       // output tile
constexpr int DIX = DTX + 2 * DR, DIY = DTY + 2 * DR;
constexpr int DBLOCK = 256;

LRE_DEV int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Box-filter two planes held in LDS (in0/in1: DIY x DIX, zero outside the
// image) into out0/out1 (DTY x DTX, in LDS; may alias in0/in1, which are dead
// after the horizontal pass).  hs0/hs1: DIY x DTX scratch.
LRE_DEV void box2(const float *in0, const float *in1, float *hs0, float *hs1, float *out0,
                  float *out1, int x0, int y0, int W, int H) {
    for (int t = threadIdx.x; t < DIY * DTX; t += DBLOCK) {
        const int ry = t / DTX, cx = t % DTX;
        const int x = x0 + cx;
        const float cnt = (float)(min(W - 1, x + DR) - max(0, x - DR) + 1);
        float s0 = 0.f, s1 = 0.f;
        const float *r0 = in0 + ry * DIX + cx, *r1 = in1 + ry * DIX + cx;
#pragma unroll
        for (int k = 0; k <= 2 * DR; k++) { s0 += r0[k]; s1 += r1[k]; }
        hs0[t] = s0 / cnt;
        hs1[t] = s1 / cnt;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < DTY * DTX; t += DBLOCK) {
        const int oy = t / DTX, cx = t % DTX;
        const int y = y0 + oy;
        const float cnt = (float)(min(H - 1, y + DR) - max(0, y - DR) + 1);
        float s0 = 0.f, s1 = 0.f;
#pragma unroll
        for (int k = 0; k <= 2 * DR; k++) { s0 += hs0[(oy + k) * DTX + cx]; s1 += hs1[(oy + k) * DTX + cx]; }
        out0[t] = s0 / cnt;
        out1[t] = s1 / cnt;
    }
    __syncthreads();
}

// pass 1: guide = max(0,R,G,B); a = var/(var+eps), b = mean(1-a) -> planes
__global__ __launch_bounds__(DBLOCK) void dtm_ab_kernel(const float *__restrict__ rgb,
                                                        float *__restrict__ pa,
                                                        float *__restrict__ pb, int W, int H) {
    __shared__ float in0[DIY * DIX], in1[DIY * DIX];
    __shared__ float hs0[DIY * DTX], hs1[DIY * DTX];
    float *o0 = in0, *o1 = in1;
    const int x0 = blockIdx.x * DTX, y0 = blockIdx.y * DTY;
    for (int t = threadIdx.x; t < DIY * DIX; t += DBLOCK) {
        const int y = y0 - DR + t / DIX, x = x0 - DR + t % DIX;
        float g = 0.f, g2 = 0.f;
        if (x >= 0 && x < W && y >= 0 && y < H) {
            const float *p = rgb + ((size_t)y * W + x) * 3;
            g = lre::max_std(0.0f, lre::max_std(p[0], lre::max_std(p[1], p[2])));
            g2 = g * g;
        }
        in0[t] = g;
        in1[t] = g2;
    }
    __syncthreads();
    box2(in0, in1, hs0, hs1, o0, o1, x0, y0, W, H);
    for (int t = threadIdx.x; t < DTY * DTX; t += DBLOCK) {
        const int x = x0 + t % DTX, y = y0 + t / DTX;
        if (x >= W || y >= H) continue;
        const float mean = o0[t];
        const float var = lre::max_std(0.0f, o1[t] - mean * mean);
        const float a = var / (var + 0.0025f);
        pa[(size_t)y * W + x] = a;
        pb[(size_t)y * W + x] = mean * (1.0f - a);
    }
}

// pass 2: box(a), box(b); base = max(1e-6, A*guide + B); RGB *= 1 + (gain-1) w
__global__ __launch_bounds__(DBLOCK) void dtm_apply_kernel(const float *__restrict__ pa,
                                                           const float *__restrict__ pb,
                                                           const float *in, float *out, int W,
                                                           int H) {
    __shared__ float in0[DIY * DIX], in1[DIY * DIX];
    __shared__ float hs0[DIY * DTX], hs1[DIY * DTX];
    float *o0 = in0, *o1 = in1;
    const int x0 = blockIdx.x * DTX, y0 = blockIdx.y * DTY;
    for (int t = threadIdx.x; t < DIY * DIX; t += DBLOCK) {
        const int y = y0 - DR + t / DIX, x = x0 - DR + t % DIX;
        const bool ok = x >= 0 && x < W && y >= 0 && y < H;
        in0[t] = ok ? pa[(size_t)y * W + x] : 0.f;
        in1[t] = ok ? pb[(size_t)y * W + x] : 0.f;
    }
    __syncthreads();
    box2(in0, in1, hs0, hs1, o0, o1, x0, y0, W, H);
    for (int t = threadIdx.x; t < DTY * DTX; t += DBLOCK) {
        const int x = x0 + t % DTX, y = y0 + t / DTX;
        if (x >= W || y >= H) continue;
        const size_t i = ((size_t)y * W + x) * 3;
        float r = in[i], g = in[i + 1], b = in[i + 2];
        const float guide = lre::max_std(0.0f, lre::max_std(r, lre::max_std(g, b)));
        const float base = lre::max_std(1e-6f, o0[t] * guide + o1[t]);
        float gain = lre::clamp_std(lre::aces(base) / base, 0.0f, 1.0f);
        float w = lre::clamp_std((base - 0.55f) / (0.90f - 0.55f), 0.0f, 1.0f);
        w = w * w * (3.0f - 2.0f * w);
        if (w > 0.0f) {
            gain = 1.0f + (gain - 1.0f) * w;
            r *= gain;
            g *= gain;
            b *= gain;
        }
        out[i] = r;
        out[i + 1] = g;
        out[i + 2] = b;
    }
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

int grid_for(size_t n, int block) {
    size_t b = (n + block - 1) / block;
    if (b > 65536) b = 65536;
    return (int)(b > 0 ? b : 1);
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
                              hipStream_t s) {
    float *pa = static_cast<float *>(d_scratch);
    float *pb = pa + (size_t)w * h;
    dim3 grid((w + DTX - 1) / DTX, (h + DTY - 1) / DTY);
    hipLaunchKernelGGL(dtm_ab_kernel, grid, dim3(DBLOCK), 0, s, d_in, pa, pb, w, h);
    LRE_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(dtm_apply_kernel, grid, dim3(DBLOCK), 0, s, pa, pb, d_in, d_out, w, h);
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
