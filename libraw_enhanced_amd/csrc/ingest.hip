// Integer preprocessing in front of the demosaics (gfx950):
//  * black subtraction and clip on the reference's uint16 x 4 image
//    (core/libraw_wrapper.cpp:401-491);
//  * fused ingest of a uint16 CFA plane: black subtraction and clip, white
//    balance into a float CFA plane (core/cpu_accelerator.cpp:2813-2869), and
//    the data maximum that adjust_maximum compares against the white level
//    (core/libraw_wrapper.cpp:1037-1077), reduced on the device.
//  * green matching (core/libraw_wrapper.cpp:496-585) and zero removal
//    (:150-212) on the uint16 x 4 image.
// Integer arithmetic, one float multiply per sample, and the reference's FP64
// green-matching ratios: bit-exact.
#include "lre_common.h"
#include "lre_internal.h"

namespace {

constexpr int BLOCK = 256;

struct Black {
    int c[4];
    int any, black, maximum;
    int pos_rows, pos_cols;
};

LRE_DEV int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__global__ __launch_bounds__(BLOCK) void black_level_kernel(ushort4 *img, int W, int H, Black B,
                                                           const int *__restrict__ pos) {
    const size_t n = (size_t)W * H;
    for (size_t q = (size_t)blockIdx.x * BLOCK + threadIdx.x; q < n; q += (size_t)gridDim.x * BLOCK) {
        ushort4 s = img[q];
        int v[4] = {s.x, s.y, s.z, s.w};
        int pb = 0;
        if (pos) pb = pos[((int)(q / W) % B.pos_rows) * B.pos_cols + (int)(q % W) % B.pos_cols];
        for (int c = 0; c < 4; c++) {
            int x = v[c];
            if (pos) { x -= pb; x -= B.c[c]; }
            else if (B.any) x -= B.c[c];
            else x -= B.black;
            v[c] = clampi(x, 0, B.maximum);
        }
        img[q] = make_ushort4((unsigned short)v[0], (unsigned short)v[1], (unsigned short)v[2],
                              (unsigned short)v[3]);
    }
}

struct XT { signed char v[36]; };

__global__ __launch_bounds__(BLOCK) void ingest_kernel(const unsigned short *__restrict__ raw,
                                                      float *__restrict__ cfa, int W, int H,
                                                      uint32_t filters, XT xt, Black B, float4 wb,
                                                      unsigned *dmax) {
    const size_t n = (size_t)W * H;
    float m = 0.f;
    for (lre::GridWalk gw((size_t)blockIdx.x * BLOCK + threadIdx.x, n, (size_t)gridDim.x * BLOCK, W); gw.more();
         gw.next()) {
        const size_t i = gw.i;
        const unsigned row = (unsigned)gw.row, col = (unsigned)gw.col;
        unsigned c4, c3;
        if (filters == 9) {
            c4 = c3 = (unsigned)xt.v[(row % 6) * 6 + col % 6];
        } else {
            c4 = lre::bayer_native(filters, row, col);
            c3 = c4 == 3 ? 1 : c4;
        }
        int v = raw[i];
        v -= B.any ? B.c[c4] : B.black;
        v = clampi(v, 0, B.maximum);
        const float mul = c4 == 0 ? wb.x : c4 == 1 ? wb.y : c4 == 2 ? wb.z : wb.w;
        const float f = (float)(unsigned short)v * mul;
        cfa[i] = f;
        if (lre::bayer3(filters, row, col) == c3 && f > m) m = f;
    }
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    __shared__ float red[BLOCK / 64];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0 && dmax) {
        float v = red[0];
        for (int k = 1; k < BLOCK / 64; k++) v = fmaxf(v, red[k]);
        // the maximum only grows: skip the contended atomic below a stored value
        const unsigned u = __float_as_uint(v);
        if (u > __hip_atomic_load(dmax, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(dmax, u);
    }
}

__global__ void zero_u32(unsigned *p) { *p = 0u; }

int grid_for(size_t n) {
    size_t b = (n + BLOCK - 1) / BLOCK;
    if (b > 8192) b = 8192;
    return (int)(b > 0 ? b : 1);
}

Black make_black(const int cblack[4], int black, int maximum) {
    Black B;
    for (int c = 0; c < 4; c++) B.c[c] = cblack ? cblack[c] : 0;
    B.any = B.c[0] || B.c[1] || B.c[2] || B.c[3];
    B.black = black;
    B.maximum = maximum;
    B.pos_rows = B.pos_cols = 0;
    return B;
}

// adjust_maximum (core/libraw_wrapper.cpp:1070-1077) on the device-side data
// maximum, so the demosaic can follow the ingest without a host round trip
__global__ void adjust_maximum_kernel(const float *d_data_max, float maximum, float threshold,
                                      float *d_maxv) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float m = maximum;
    if (!(threshold < 0.00001f)) {
        const float thr = threshold > 0.99999f ? 0.75f : threshold;
        const float d = *d_data_max;
        if (d > 0 && d < maximum && d > maximum * thr) m = d;
    }
    *d_maxv = m;
}

// The uint16 raw image of the preprocessing stages, as the reference holds it
// (uint16 x 4 per pixel, only the channel of the pixel's Bayer colour set) or
// as one sample per site.  get/set address channel ch of pixel (r, c); in the
// plane, a channel other than the site's colour reads 0 and is not stored,
// which is what the four-channel image shows to every later stage (they read
// a pixel's own channel only).
struct Img4View {
    unsigned short *p;
    int W;
    LRE_DEV unsigned short get(int r, int c, unsigned ch) const { return p[((size_t)r * W + c) * 4 + ch]; }
    LRE_DEV void set(int r, int c, unsigned ch, unsigned short v) const { p[((size_t)r * W + c) * 4 + ch] = v; }
};
struct PlaneView {
    unsigned short *p;
    int W;
    uint32_t filters;
    LRE_DEV bool own(int r, int c, unsigned ch) const {
        return lre::bayer_native(filters, (unsigned)r, (unsigned)c) == ch;
    }
    LRE_DEV unsigned short get(int r, int c, unsigned ch) const { return own(r, c, ch) ? p[(size_t)r * W + c] : 0; }
    LRE_DEV void set(int r, int c, unsigned ch, unsigned short v) const {
        if (own(r, c, ch)) p[(size_t)r * W + c] = v;
    }
};

// ---- green matching (:496-585): reads src, writes the channel-3 samples it
// corrects into dst (a copy of src) ----
template <class V>
__global__ __launch_bounds__(BLOCK) void green_match_kernel(V src, V dst, int W, int H, int maximum) {
    const int nj = (H - 3 - 3 + 1) / 2, ni = (W - 3 - 2 + 1) / 2;  // j = 3,5,.. < H-3; i = 2,4,.. < W-3
    if (nj <= 0 || ni <= 0) return;
    const size_t n = (size_t)nj * ni;
    const float thr = 0.01f;
    for (size_t t = (size_t)blockIdx.x * BLOCK + threadIdx.x; t < n; t += (size_t)gridDim.x * BLOCK) {
        const int j = 3 + 2 * (int)(t / ni), i = 2 + 2 * (int)(t % ni);
        const int o1_1 = src.get(j - 1, i - 1, 1), o1_2 = src.get(j - 1, i + 1, 1);
        const int o1_3 = src.get(j + 1, i - 1, 1), o1_4 = src.get(j + 1, i + 1, 1);
        const int o2_1 = src.get(j - 2, i, 3), o2_2 = src.get(j + 2, i, 3);
        const int o2_3 = src.get(j, i - 2, 3), o2_4 = src.get(j, i + 2, 3);
        const double m1 = (o1_1 + o1_2 + o1_3 + o1_4) / 4.0;
        const double m2 = (o2_1 + o2_2 + o2_3 + o2_4) / 4.0;
        const double c1 = (abs(o1_1 - o1_2) + abs(o1_1 - o1_3) + abs(o1_1 - o1_4) + abs(o1_2 - o1_3) +
                           abs(o1_3 - o1_4) + abs(o1_2 - o1_4)) / 6.0;
        const double c2 = (abs(o2_1 - o2_2) + abs(o2_1 - o2_3) + abs(o2_1 - o2_4) + abs(o2_2 - o2_3) +
                           abs(o2_3 - o2_4) + abs(o2_2 - o2_4)) / 6.0;
        const int g2 = src.get(j, i, 3);
        if ((g2 < maximum * 0.95) && (c1 < maximum * thr) && (c2 < maximum * thr) && (m2 > 0.1)) {
            const float correction = (float)(g2 * m1 / m2);
            dst.set(j, i, 3, correction > maximum ? (unsigned short)maximum : (unsigned short)correction);
        }
    }
}

// ---- zero removal (:150-212).  The reference replaces zeros in raster order,
// so a zero can see the new value of an earlier zero of its channel in its 5x5
// window.  Zeros with no such earlier zero ("independent") are replaced in
// parallel: their windows hold no other independent zero.  The rest are
// replaced afterwards by one wavefront in raster order. ----
LRE_DEV unsigned chan(uint32_t filters, int r, int c) { return lre::bayer_native(filters, (unsigned)r, (unsigned)c); }

template <class V>
__global__ __launch_bounds__(BLOCK) void zero_classify_kernel(V img, int W, int H, uint32_t filters,
                                                             unsigned char *__restrict__ dep,
                                                             unsigned *__restrict__ rowflag) {
    const size_t n = (size_t)W * H;
    for (lre::GridWalk gw((size_t)blockIdx.x * BLOCK + threadIdx.x, n, (size_t)gridDim.x * BLOCK, W); gw.more();
         gw.next()) {
        const size_t i = gw.i;
        const int row = (int)gw.row, col = (int)gw.col;
        const unsigned ch = chan(filters, row, col);
        unsigned char d = 0;
        if (img.get(row, col, ch) == 0) {
            d = 1;  // independent zero
            for (int r = row - 2; r <= row && d == 1; r++)
                for (int c = col - 2; c <= col + 2; c++) {
                    if (r == row && c >= col) break;
                    if (r < 0 || c < 0 || c >= W) continue;
                    if (chan(filters, r, c) == ch && img.get(r, c, ch) == 0) { d = 2; break; }
                }
            if (d == 2) rowflag[row] = 1u;
        }
        dep[i] = d;
    }
}

template <class V>
LRE_DEV unsigned short window_mean(const V &img, int W, int H, uint32_t filters, int row, int col, unsigned ch,
                                   bool &any) {
    unsigned sum = 0, cnt = 0;
    for (int r = row - 2; r <= row + 2; r++)
        for (int c = col - 2; c <= col + 2; c++) {
            if (r < 0 || r >= H || c < 0 || c >= W) continue;
            const unsigned nch = chan(filters, r, c);
            const unsigned short v = img.get(r, c, nch);
            if (nch == ch && v != 0) { sum += v; cnt++; }
        }
    any = cnt > 0;
    return any ? (unsigned short)(sum / cnt) : 0;
}

template <class V>
__global__ __launch_bounds__(BLOCK) void zero_fill_independent(V img, int W, int H, uint32_t filters,
                                                              const unsigned char *__restrict__ dep) {
    const size_t n = (size_t)W * H;
    for (lre::GridWalk gw((size_t)blockIdx.x * BLOCK + threadIdx.x, n, (size_t)gridDim.x * BLOCK, W); gw.more();
         gw.next()) {
        const size_t i = gw.i;
        if (dep[i] != 1) continue;
        const int row = (int)gw.row, col = (int)gw.col;
        const unsigned ch = chan(filters, row, col);
        bool any;
        const unsigned short v = window_mean(img, W, H, filters, row, col, ch, any);
        if (any) img.set(row, col, ch, v);
    }
}

// one wavefront: rows in order; within a flagged row, 64 columns at a time,
// the flagged zeros of the chunk one after another (lowest column first)
template <class V>
__global__ __launch_bounds__(64) void zero_fill_dependent(V img, int W, int H, uint32_t filters,
                                                         const unsigned char *__restrict__ dep,
                                                         const unsigned *__restrict__ rowflag) {
    const int lane = threadIdx.x;
    for (int row = 0; row < H; row++) {
        if (!rowflag[row]) continue;
        for (int c0 = 0; c0 < W; c0 += 64) {
            const int c = c0 + lane;
            const bool mine = c < W && dep[(size_t)row * W + c] == 2;
            unsigned long long mask = __ballot(mine);
            while (mask) {
                const int k = __ffsll((long long)mask) - 1;
                mask &= mask - 1;
                const int col = c0 + k;
                if (lane == 0) {
                    const unsigned ch = chan(filters, row, col);
                    bool any;
                    const unsigned short v = window_mean(img, W, H, filters, row, col, ch, any);
                    if (any) img.set(row, col, ch, v);
                }
                // the next zero of the chunk reads this one's value
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
                __builtin_amdgcn_wave_barrier();
            }
        }
    }
}

// ---- automatic white balance of scale_colors (:621-663): exact integer sums of
// every native channel; the reference's double sums of integers are exact too ----
__global__ __launch_bounds__(BLOCK) void awb_sum_kernel(const unsigned short *__restrict__ img4, int W,
                                                       int H, uint32_t filters, XT xt,
                                                       unsigned long long *acc) {
    unsigned long long s[4] = {0, 0, 0, 0}, k[4] = {0, 0, 0, 0};
    const size_t n = (size_t)W * H;
    for (lre::GridWalk gw((size_t)blockIdx.x * BLOCK + threadIdx.x, n, (size_t)gridDim.x * BLOCK, W); gw.more();
         gw.next()) {
        const size_t i = gw.i;
        const int row = (int)gw.row, col = (int)gw.col;
        const unsigned cc = filters == 9 ? (unsigned)xt.v[((row + 6) % 6) * 6 + (col + 6) % 6]
                                         : lre::bayer_native(filters, (unsigned)row, (unsigned)col);
        const unsigned v = img4[i * 4 + cc];
#pragma unroll
        for (int c = 0; c < 4; c++) {
            s[c] += (cc == (unsigned)c) ? v : 0u;
            k[c] += (cc == (unsigned)c) ? 1u : 0u;
        }
    }
#pragma unroll
    for (int c = 0; c < 4; c++)
        for (int o = 32; o > 0; o >>= 1) {
            s[c] += __shfl_xor(s[c], o, 64);
            k[c] += __shfl_xor(k[c], o, 64);
        }
    if ((threadIdx.x & 63) == 0)
        for (int c = 0; c < 4; c++) {
            atomicAdd(&acc[c], s[c]);
            atomicAdd(&acc[4 + c], k[c]);
        }
}

__global__ void awb_finish_kernel(const unsigned long long *acc, float *pre_mul) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double norm[4], count[4];
    for (int c = 0; c < 4; c++) {
        norm[c] = (double)acc[c];
        count[c] = (double)acc[4 + c];
    }
    norm[1] += norm[3];
    count[1] += count[3];
    for (int c = 0; c < 3; c++) norm[c] = norm[c] / count[c];  // the p-norm with p = 1
    pre_mul[0] = (float)(norm[1] / norm[0]);
    pre_mul[1] = 1.0f;
    pre_mul[2] = (float)(norm[1] / norm[2]);
    pre_mul[3] = 1.0f;
}

__global__ void zero_u64(unsigned long long *p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

// ---- legacy radial chromatic aberration of scale_colors (:833-880) ----
// float -> unsigned as the reference's x86-64 build converts it: truncation to
// a 64-bit integer, low 32 bits kept (negative sources become huge, skipped)
LRE_DEV unsigned f2u_x86(float f) { return (unsigned)(long long)f; }

__global__ __launch_bounds__(BLOCK) void aber_copy_kernel(const unsigned short *__restrict__ img4,
                                                         unsigned short *__restrict__ p0,
                                                         unsigned short *__restrict__ p2, size_t n) {
    for (size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (size_t)gridDim.x * BLOCK) {
        p0[i] = img4[i * 4];
        p2[i] = img4[i * 4 + 2];
    }
}

__global__ __launch_bounds__(BLOCK) void aber_warp_kernel(unsigned short *img4, const unsigned short *plane,
                                                         int c, int W, int H, float scale) {
    const size_t n = (size_t)W * H;
    const float cy = (float)H * 0.5f, cx = (float)W * 0.5f;
    for (lre::GridWalk gw((size_t)blockIdx.x * BLOCK + threadIdx.x, n, (size_t)gridDim.x * BLOCK, W); gw.more();
         gw.next()) {
        const size_t i = gw.i;
        const int row = (int)gw.row, col = (int)gw.col;
        const float sc = ((float)col - cx) * scale + cx;
        const float sr = ((float)row - cy) * scale + cy;
        const unsigned ur = f2u_x86(sr);
        if (ur >= (unsigned)H - 1) continue;
        const float fr = sr - (float)ur;
        const unsigned uc = f2u_x86(sc);
        if (uc >= (unsigned)W - 1) continue;
        const float fc = sc - (float)uc;
        const unsigned short *p = plane + (size_t)ur * W + uc;
        const float v = ((float)p[0] * (1.0f - fc) + (float)p[1] * fc) * (1.0f - fr) +
                        ((float)p[W] * (1.0f - fc) + (float)p[W + 1] * fc) * fr;
        const float lo = (0.0f < v) ? v : 0.0f;
        const float cl = (lo < 65535.0f) ? lo : 65535.0f;
        img4[i * 4 + c] = (unsigned short)cl;
    }
}

__global__ void zero_bytes(unsigned *p, size_t n4) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

}  // namespace

namespace {

// green matching over a view; dst already holds a copy of src
template <class V>
int green_match_on(V src, V dst, int w, int h, int maximum, hipStream_t s) {
    hipLaunchKernelGGL((green_match_kernel<V>), dim3(grid_for((size_t)w * h / 4)), dim3(BLOCK), 0, s, src, dst,
                       w, h, maximum);
    LRE_HIP_CHECK(hipGetLastError());
    return LRE_OK;
}

template <class V>
int remove_zeroes_on(V img, int w, int h, uint32_t filters, void *d_scratch, hipStream_t s) {
    unsigned char *dep = static_cast<unsigned char *>(d_scratch);
    unsigned *rowflag = reinterpret_cast<unsigned *>(dep + (((size_t)w * h + 255) & ~(size_t)255));
    const size_t nflag = ((size_t)h * 4 + 255) / 4;
    hipLaunchKernelGGL(zero_bytes, dim3(grid_for(nflag)), dim3(BLOCK), 0, s, rowflag, nflag);
    const int g = grid_for((size_t)w * h);
    hipLaunchKernelGGL((zero_classify_kernel<V>), dim3(g), dim3(BLOCK), 0, s, img, w, h, filters, dep, rowflag);
    hipLaunchKernelGGL((zero_fill_independent<V>), dim3(g), dim3(BLOCK), 0, s, img, w, h, filters, dep);
    hipLaunchKernelGGL((zero_fill_dependent<V>), dim3(1), dim3(64), 0, s, img, w, h, filters, dep, rowflag);
    LRE_HIP_CHECK(hipGetLastError());
    return LRE_OK;
}

// black subtraction and clip of a one-sample-per-site plane, in place: the
// channel of each site as the fused ingest takes it (:401-491)
__global__ __launch_bounds__(BLOCK) void black_plane_kernel(unsigned short *plane, int W, int H, uint32_t filters,
                                                           XT xt, Black B) {
    const size_t n = (size_t)W * H;
    for (lre::GridWalk gw((size_t)blockIdx.x * BLOCK + threadIdx.x, n, (size_t)gridDim.x * BLOCK, W); gw.more();
         gw.next()) {
        const unsigned row = (unsigned)gw.row, col = (unsigned)gw.col;
        const unsigned c4 = filters == 9 ? (unsigned)xt.v[(row % 6) * 6 + col % 6]
                                         : lre::bayer_native(filters, row, col);
        int v = plane[gw.i];
        v -= B.any ? B.c[c4] : B.black;
        plane[gw.i] = (unsigned short)clampi(v, 0, B.maximum);
    }
}

}  // namespace

int lre_launch_green_matching(const uint16_t *d_src, uint16_t *d_dst, int w, int h, uint32_t filters,
                              int maximum, hipStream_t s) {
    if (d_src != d_dst)
        LRE_HIP_CHECK(hipMemcpyAsync(d_dst, d_src, (size_t)w * h * 8, hipMemcpyDeviceToDevice, s));
    if (filters == 9) return LRE_OK;  // Bayer only (:499-503)
    if (d_src == d_dst) {
        lre_set_error("green matching reads the unmodified image: source and destination must differ");
        return LRE_ERR_ARG;
    }
    const Img4View src{const_cast<unsigned short *>(reinterpret_cast<const unsigned short *>(d_src)), w};
    const Img4View dst{reinterpret_cast<unsigned short *>(d_dst), w};
    return green_match_on(src, dst, w, h, maximum, s);
}

int lre_launch_green_matching_plane(const uint16_t *d_src, uint16_t *d_dst, int w, int h, uint32_t filters,
                                    int maximum, hipStream_t s) {
    if (d_src == d_dst) {
        lre_set_error("green matching reads the unmodified image: source and destination must differ");
        return LRE_ERR_ARG;
    }
    LRE_HIP_CHECK(hipMemcpyAsync(d_dst, d_src, (size_t)w * h * 2, hipMemcpyDeviceToDevice, s));
    if (filters == 9) return LRE_OK;
    const PlaneView src{const_cast<unsigned short *>(reinterpret_cast<const unsigned short *>(d_src)), w, filters};
    const PlaneView dst{reinterpret_cast<unsigned short *>(d_dst), w, filters};
    return green_match_on(src, dst, w, h, maximum, s);
}

size_t lre_remove_zeroes_scratch_bytes(int w, int h) {
    return (((size_t)w * h + 255) & ~(size_t)255) + (((size_t)h * 4 + 255) & ~(size_t)255);
}

int lre_launch_remove_zeroes(uint16_t *d_img4, int w, int h, uint32_t filters, void *d_scratch,
                             hipStream_t s) {
    return remove_zeroes_on(Img4View{reinterpret_cast<unsigned short *>(d_img4), w}, w, h, filters, d_scratch, s);
}

int lre_launch_remove_zeroes_plane(uint16_t *d_plane, int w, int h, uint32_t filters, void *d_scratch,
                                   hipStream_t s) {
    return remove_zeroes_on(PlaneView{reinterpret_cast<unsigned short *>(d_plane), w, filters}, w, h, filters,
                            d_scratch, s);
}

int lre_launch_black_plane(uint16_t *d_plane, int w, int h, uint32_t filters, const signed char *xtrans,
                           const int cblack[4], int black, int maximum, hipStream_t s) {
    XT xt{};
    if (xtrans)
        for (int i = 0; i < 36; i++) xt.v[i] = xtrans[i];
    hipLaunchKernelGGL(black_plane_kernel, dim3(grid_for((size_t)w * h)), dim3(BLOCK), 0, s,
                       reinterpret_cast<unsigned short *>(d_plane), w, h, filters, xt,
                       make_black(cblack, black, maximum));
    LRE_HIP_CHECK(hipGetLastError());
    return LRE_OK;
}

int lre_launch_auto_wb(const uint16_t *d_img4, int w, int h, uint32_t filters,
                       const signed char *xtrans, float *d_pre_mul, void *d_scratch, hipStream_t s) {
    XT xt{};
    if (xtrans)
        for (int i = 0; i < 36; i++) xt.v[i] = xtrans[i];
    unsigned long long *acc = static_cast<unsigned long long *>(d_scratch);
    hipLaunchKernelGGL(zero_u64, dim3(1), dim3(64), 0, s, acc, 8);
    hipLaunchKernelGGL(awb_sum_kernel, dim3(grid_for((size_t)w * h)), dim3(BLOCK), 0, s,
                       reinterpret_cast<const unsigned short *>(d_img4), w, h, filters, xt, acc);
    hipLaunchKernelGGL(awb_finish_kernel, dim3(1), dim3(64), 0, s, acc, d_pre_mul);
    LRE_HIP_CHECK(hipGetLastError());
    return LRE_OK;
}

size_t lre_scale_aber_scratch_bytes(int w, int h) {
    return 2 * (((size_t)w * h * 2 + 255) & ~(size_t)255);
}

int lre_launch_scale_aber(uint16_t *d_img4, int w, int h, const float aber[4], void *d_scratch,
                          hipStream_t s) {
    if (w <= 1 || h <= 1) return LRE_OK;
    const size_t n = (size_t)w * h;
    unsigned short *p0 = static_cast<unsigned short *>(d_scratch);
    unsigned short *p2 = p0 + (((n * 2 + 255) & ~(size_t)255) / 2);
    hipLaunchKernelGGL(aber_copy_kernel, dim3(grid_for(n)), dim3(BLOCK), 0, s,
                       reinterpret_cast<const unsigned short *>(d_img4), p0, p2, n);
    for (int c = 0; c < 4; c += 2) {
        const float scale = aber[c];
        if (fabsf(scale - 1.0f) < 1e-6f || scale <= 0.0f) continue;
        hipLaunchKernelGGL(aber_warp_kernel, dim3(grid_for(n)), dim3(BLOCK), 0, s,
                           reinterpret_cast<unsigned short *>(d_img4), c == 0 ? p0 : p2, c, w, h, scale);
    }
    LRE_HIP_CHECK(hipGetLastError());
    return LRE_OK;
}

int lre_launch_adjust_maximum(const float *d_data_max, float maximum, float threshold,
                              float *d_maxv, hipStream_t s) {
    hipLaunchKernelGGL(adjust_maximum_kernel, dim3(1), dim3(64), 0, s, d_data_max, maximum,
                       threshold, d_maxv);
    LRE_HIP_CHECK(hipGetLastError());
    return LRE_OK;
}

int lre_launch_black_level(uint16_t *d_img4, int w, int h, const int cblack[4], const int *d_pos,
                           int pos_rows, int pos_cols, int black, int maximum, hipStream_t s) {
    Black B = make_black(cblack, black, maximum);
    B.pos_rows = pos_rows;
    B.pos_cols = pos_cols;
    hipLaunchKernelGGL(black_level_kernel, dim3(grid_for((size_t)w * h)), dim3(BLOCK), 0, s,
                       reinterpret_cast<ushort4 *>(d_img4), w, h, B,
                       (pos_rows > 0 && pos_cols > 0) ? d_pos : nullptr);
    LRE_HIP_CHECK(hipGetLastError());
    return LRE_OK;
}

int lre_launch_ingest_u16(const uint16_t *d_raw, float *d_cfa, int w, int h, uint32_t filters,
                          const signed char *xtrans, const int cblack[4], int black, int maximum,
                          const float wb[4], float *d_data_max, hipStream_t s) {
    XT xt{};
    if (xtrans)
        for (int i = 0; i < 36; i++) xt.v[i] = xtrans[i];
    unsigned *dm = reinterpret_cast<unsigned *>(d_data_max);
    if (dm) hipLaunchKernelGGL(zero_u32, dim3(1), dim3(1), 0, s, dm);
    hipLaunchKernelGGL(ingest_kernel, dim3(grid_for((size_t)w * h)), dim3(BLOCK), 0, s,
                       reinterpret_cast<const unsigned short *>(d_raw), d_cfa, w, h, filters, xt,
                       make_black(cblack, black, maximum), make_float4(wb[0], wb[1], wb[2], wb[3]), dm);
    LRE_HIP_CHECK(hipGetLastError());
    return LRE_OK;
}
