// Integer preprocessing in front of the demosaics (gfx950):
//  * black subtraction and clip on the reference's uint16 x 4 image
//    (core/libraw_wrapper.cpp:401-491);
//  * fused ingest of a uint16 CFA plane: black subtraction and clip, white
//    balance into a float CFA plane (core/cpu_accelerator.cpp:2813-2869), and
//    the data maximum that adjust_maximum compares against the white level
//    (core/libraw_wrapper.cpp:1037-1077), reduced on the device.
// Integer arithmetic and one float multiply per sample: bit-exact.
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
    for (size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (size_t)gridDim.x * BLOCK) {
        const unsigned row = (unsigned)(i / W), col = (unsigned)(i % W);
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
        atomicMax(dmax, __float_as_uint(v));
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

}  // namespace

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
