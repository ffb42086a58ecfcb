// Chromatic-aberration stages for gfx950.
//
// Axial CA (core/cpu_accelerator.cpp:3799-3928): guided filter of R and B with
// G as the guide, FP64 box means over windows clipped at the frame, edge gate
// from the box-mean of a 2-tap gradient magnitude.  The reference builds FP64
// summed-area tables; here each workgroup sums its windows directly in FP64
// (horizontal then vertical passes through LDS), which is at least as accurate;
// the float results agree except for rare last-bit differences.
//
// Lateral CA (core/cpu_accelerator.cpp:3496-3673): 2x2 box pyramid, then per
// level one workgroup per (cell, channel) iterating Lucas-Kanade with the 2x2
// normal equations reduced in FP64 across the workgroup; confidence gate and
// 3x3 cell smoothing run on the device, so no level waits on the host.
#include "lre_common.h"
#include "lre_internal.h"

namespace {

constexpr int BLOCK = 256;

LRE_DEV float mx(float a, float b) { return lre::max_std(a, b); }

__global__ void init_u32(unsigned *p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0u;
}

// zero fill (a kernel, not a memset node, so the stage can be graph-captured)
__global__ void zero_f32(float *p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.f;
}

// max(1e-6, max G) folded into stats[0] (:3823-3830)
__global__ __launch_bounds__(BLOCK) void axial_max_g(const float *__restrict__ src, size_t n,
                                                     unsigned *stats) {
    float m = lre::green_max(src, n, 1e-6f, [](float v) { return v; });
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    __shared__ float red[BLOCK / 64];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float v = red[0];
        for (int k = 1; k < BLOCK / 64; k++) v = fmaxf(v, red[k]);
        // the maximum only grows: skip the contended atomic below a stored value
        const unsigned u = __float_as_uint(v);
        if (u > __hip_atomic_load(stats, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(stats, u);
    }
}

constexpr int AX_TX = 64, AX_TY = 32, AX_RMAX = 16;
constexpr int AX_CHUNK = BLOCK * 16;  // pixels per workgroup of the multi-image max(G)
constexpr int AX_BLOCK = 512;  // one 4-row strip of one column per thread

struct AxialArgs {
    const float *src;
    float *dst;
    float *aR, *bR, *aB, *bB, *ew;
    int W, H, r;
    float eps, strength;
    const unsigned *stats;
};

// several images: the max(G) statistics reset (one workgroup per image) and
// folded over chunks of AX_CHUNK pixels
__global__ void axial_stats_reset_many(lre::Many<AxialArgs> S) {
    const lre::TileOf tl = S.tile();
    if (threadIdx.x < 4) const_cast<unsigned *>(S.args(tl).stats)[threadIdx.x] = 0u;
}
__global__ __launch_bounds__(BLOCK) void axial_max_g_many(lre::Many<AxialArgs> S) {
    const lre::TileOf tl = S.tile();
    const AxialArgs &A = S.args(tl);
    const size_t n = (size_t)A.W * A.H, b = (size_t)tl.tx * AX_CHUNK, e = b + AX_CHUNK < n ? b + AX_CHUNK : n;
    float m = lre::green_max_range(A.src, b, e, 1e-6f, [](float v) { return v; });
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    __shared__ float red[BLOCK / 64];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float v = red[0];
        for (int k = 1; k < BLOCK / 64; k++) v = fmaxf(v, red[k]);
        unsigned *st = const_cast<unsigned *>(A.stats);
        const unsigned u = __float_as_uint(v);
        if (u > __hip_atomic_load(st, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(st, u);
    }
}

// Horizontal FP64 window sums over LDS rows [0, RY) into hs[RY][AX_HP], for
// output columns x0 .. x0+AX_TX-1; region column 0 is image column x0 - r - 1
// and region values outside the image are 0.  q(ry, c) gives the summed value
// at region column c.  Each thread sums 4 neighbouring outputs of one row: the
// first window in full, the next ones by adding the entering and subtracting
// the leaving tap (FP64, so the result stays well inside the tolerance of the
// reference's FP64 summed-area tables); lanes walk down rows (odd pitches:
// distinct LDS banks).
constexpr int AX_HP = 64 + 1;  // hs pitch in doubles
constexpr int AX_RB = 4;

template <class Q>
LRE_DEV void hsum(Q q, int RY, int r, double *hs) {
    const int nt = 2 * r + 1;
    for (int t = threadIdx.x; t < RY * (AX_TX / AX_RB); t += AX_BLOCK) {
        const int ry = t % RY, cx = (t / RY) * AX_RB;
        const int c0 = cx + 1;  // region col cx+1 = image col x0+cx-r
        double s = 0.0;
        for (int k = 0; k < nt; k++) s += (double)q(ry, c0 + k);
        hs[ry * AX_HP + cx] = s;
#pragma unroll
        for (int u = 1; u < AX_RB; u++) {
            s += (double)q(ry, c0 + u - 1 + nt) - (double)q(ry, c0 + u - 1);
            hs[ry * AX_HP + cx + u] = s;
        }
    }
}

// Vertical sums of hs for output rows oy .. oy+3 of column cx (a full window,
// then entering minus leaving taps), over the true tap count of the clipped
// window: times 1 / (2r+1)^2 where the window is whole (inv_full, one FP64
// division per thread instead of one per mean; the two differ by at most an
// ulp of the double mean), divided where it is clipped.
LRE_DEV void vmean(const double *hs, int r, int cx, int oy, int x0, int y0, int W, int H, double inv_full,
                   double out[AX_RB]) {
    const int nt = 2 * r + 1;
    const double *col = hs + oy * AX_HP + cx;
    double s[AX_RB];
    double acc = 0.0;
    for (int j = 0; j < nt; j++) acc += col[j * AX_HP];
    s[0] = acc;
#pragma unroll
    for (int u = 1; u < AX_RB; u++) {
        acc += col[(u - 1 + nt) * AX_HP] - col[(u - 1) * AX_HP];
        s[u] = acc;
    }
    const int x = x0 + cx;
    const int xl = max(0, x - r), xr = min(W - 1, x + r);
    const int full = (2 * r + 1) * (2 * r + 1);
#pragma unroll
    for (int u = 0; u < AX_RB; u++) {
        const int y = y0 + oy + u;
        const int yl = max(0, y - r), yb = min(H - 1, y + r);
        const int area = (xr - xl + 1) * (yb - yl + 1);
        out[u] = area == full ? s[u] * inv_full : (area > 0 ? s[u] / (double)area : 0.0);
    }
}

static_assert(AX_TX * AX_TY == AX_BLOCK * AX_RB, "one 4-row strip per thread");

size_t axial_ab_lds(int r) {
    const int RX = AX_TX + 2 * r + 2, RY = AX_TY + 2 * r + 2;
    return sizeof(float) * (RX | 1) * (4 * RY - 2) + sizeof(double) * (RY - 2) * AX_HP;
}

// pass 1: a, b for R and B, and the edge weight.  Src: lre::One<AxialArgs>
// (one image, 2D grid) or lre::Many<AxialArgs> (the tiles of several images).
template <class Src>
__global__ __launch_bounds__(AX_BLOCK) void axial_ab(Src S) {
    extern __shared__ __align__(16) unsigned char smem[];
    const lre::TileOf tl = S.tile();
    const AxialArgs &A = S.args(tl);
    const int r = A.r;
    const double inv_full = 1.0 / (double)((2 * r + 1) * (2 * r + 1));
    const int RX = AX_TX + 2 * r + 2, RY = AX_TY + 2 * r + 2;  // one extra ring for the gradient
    const int RXP = RX | 1;                              // odd LDS pitch
    float *sR = reinterpret_cast<float *>(smem);
    float *sG = sR + RXP * RY;
    float *sB = sG + RXP * RY;
    float *sE = sB + RXP * RY;                           // gradient magnitude, rows 1 .. RY-2
    double *hs = reinterpret_cast<double *>(sE + RXP * (RY - 2));  // (RY-2) x AX_HP
    const int x0 = tl.tx * AX_TX, y0 = tl.ty * AX_TY;
    const int W = A.W, H = A.H;
    const float inv_norm = 1.f / mx(__uint_as_float(A.stats[0]), 1e-6f);
    lre::stage2d<AX_BLOCK>(
        RY, RX,
        [&](int ry, int rx) {
            const int y = y0 - r - 1 + ry, x = x0 - r - 1 + rx;
            float3 v = make_float3(0.f, 0.f, 0.f);
            if (x >= 0 && x < W && y >= 0 && y < H) {
                const float *p = A.src + ((size_t)y * W + x) * 3;
                v = make_float3(p[0], p[1], p[2]);
            }
            return v;
        },
        [&](int ry, int rx, float3 v) {
            const int t = ry * RXP + rx;
            sR[t] = v.x * inv_norm; sG[t] = v.y * inv_norm; sB[t] = v.z * inv_norm;
        });
    __syncthreads();
    const int RYi = RY - 2;  // rows of the window region (the gradient ring dropped)
    // 2-tap central-difference magnitude, 0 on the image's 1-px frame (:3876-3890)
    for (int t = threadIdx.x; t < RYi * RX; t += AX_BLOCK) {
        const int ry = t / RX + 1, rx = t % RX;
        const int c = ry * RXP + rx;
        const int y = y0 - r - 1 + ry, x = x0 - r - 1 + rx;
        float v = 0.f;
        if (x >= 1 && x < W - 1 && y >= 1 && y < H - 1) {
            const float gx = sG[c + 1] - sG[c - 1];
            const float gy = sG[c + RXP] - sG[c - RXP];
            v = sqrtf(gx * gx + gy * gy);
        }
        sE[(ry - 1) * RXP + rx] = v;
    }
    __syncthreads();
    // this thread's outputs: rows oy .. oy+3 of column cx.  Each output is
    // finished as soon as its means are known, so few FP64 means stay live.
    constexpr int PER = AX_RB;
    const int cx = threadIdx.x % AX_TX, oy = (threadIdx.x / AX_TX) * AX_RB;
    // window means of quantity qi (region row ry + 1 holds window row ry)
    auto means = [&](int qi, double out[PER]) {
        hsum(
            [&](int ry, int c) {
                const int k = (ry + 1) * RXP + c;
                switch (qi) {
                    case 0: return sG[k];
                    case 1: return sG[k] * sG[k];
                    case 2: return sR[k];
                    case 3: return sG[k] * sR[k];
                    case 4: return sB[k];
                    case 5: return sG[k] * sB[k];
                    default: return sE[ry * RXP + c];
                }
            },
            RYi, r, hs);
        __syncthreads();
        vmean(hs, r, cx, oy, x0, y0, W, H, inv_full, out);
        __syncthreads();
    };
    auto index_of = [&](int k, size_t &i) {
        const int x = x0 + cx, y = y0 + oy + k;
        i = (size_t)y * W + x;
        return x < W && y < H;
    };
    double mI[PER], var[PER];
    {
        double mII[PER];
        means(0, mI);
        means(1, mII);
#pragma unroll
        for (int k = 0; k < PER; k++) var[k] = mII[k] - mI[k] * mI[k];
    }
    const double e = (double)A.eps;
    // a = cov(I, p) / (var + eps), b = mean(p) - a mean(I) for p = R (qi 2, 3) and B (qi 4, 5)
#pragma unroll
    for (int ch = 0; ch < 2; ch++) {
        double mP[PER], mIP[PER];
        means(2 + 2 * ch, mP);
        means(3 + 2 * ch, mIP);
#pragma unroll
        for (int k = 0; k < PER; k++) {
            size_t i;
            if (!index_of(k, i)) continue;
            const double a = (mIP[k] - mI[k] * mP[k]) / (var[k] + e);
            __builtin_nontemporal_store((float)a, &(ch ? A.aB : A.aR)[i]);
            __builtin_nontemporal_store((float)(mP[k] - a * mI[k]), &(ch ? A.bB : A.bR)[i]);
        }
    }
    double mE[PER];
    means(6, mE);
#pragma unroll
    for (int k = 0; k < PER; k++) {
        size_t i;
        if (!index_of(k, i)) continue;
        const float t = lre::clamp_std(((float)mE[k] - 0.01f) / 0.05f, 0.f, 1.f);
        __builtin_nontemporal_store(t * t * (3.f - 2.f * t), &A.ew[i]);
    }
}

size_t axial_apply_lds(int r) {
    const int RX = AX_TX + 2 * r + 2, RY = AX_TY + 2 * r;
    return sizeof(float) * (RX | 1) * RY * 4 + sizeof(double) * RY * AX_HP;
}

// pass 2: box means of a, b; q = ma * G + mb; blend by strength * edge weight
template <class Src>
__global__ __launch_bounds__(AX_BLOCK) void axial_apply(Src S) {
    extern __shared__ __align__(16) unsigned char smem[];
    const lre::TileOf tl = S.tile();
    const AxialArgs &A = S.args(tl);
    const int r = A.r;
    const double inv_full = 1.0 / (double)((2 * r + 1) * (2 * r + 1));
    const int RX = AX_TX + 2 * r + 2, RY = AX_TY + 2 * r;
    const int RXP = RX | 1;
    float *q = reinterpret_cast<float *>(smem);             // 4 planes, RY x RXP each
    double *hs = reinterpret_cast<double *>(q + 4 * RXP * RY);
    const int x0 = tl.tx * AX_TX, y0 = tl.ty * AX_TY;
    const int W = A.W, H = A.H;
    const float norm_ref = mx(__uint_as_float(A.stats[0]), 1e-6f);
    const float inv_norm = 1.f / norm_ref;
    constexpr int PER = AX_RB;
    const int cx = threadIdx.x % AX_TX, oy = (threadIdx.x / AX_TX) * AX_RB;
    double m[4][PER];
    // the four planes staged together
    lre::stage2d<AX_BLOCK>(
        RY, RX,
        [&](int ry, int rx) {
            const int y = y0 - r + ry, x = x0 - r - 1 + rx;
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            if (x >= 0 && x < W && y >= 0 && y < H) {
                const size_t i = (size_t)y * W + x;
                v = make_float4(A.aR[i], A.bR[i], A.aB[i], A.bB[i]);
            }
            return v;
        },
        [&](int ry, int rx, float4 v) {
            const int t = ry * RXP + rx;
            q[t] = v.x; q[RXP * RY + t] = v.y; q[2 * RXP * RY + t] = v.z; q[3 * RXP * RY + t] = v.w;
        });
    __syncthreads();
#pragma unroll
    for (int qi = 0; qi < 4; qi++) {
        const float *pl = q + qi * RXP * RY;
        hsum([&](int ry, int c) { return pl[ry * RXP + c]; }, RY, r, hs);
        __syncthreads();
        vmean(hs, r, cx, oy, x0, y0, W, H, inv_full, m[qi]);
        __syncthreads();
    }
    for (int k = 0; k < PER; k++) {
        const int x = x0 + cx, y = y0 + oy + k;
        if (x >= W || y >= H) continue;
        const size_t i = (size_t)y * W + x;
        const float s0 = A.src[3 * i], s1 = A.src[3 * i + 1], s2 = A.src[3 * i + 2];
        const float R = s0 * inv_norm, G = s1 * inv_norm, B = s2 * inv_norm;
        const float Rf = (float)(m[0][k] * G + m[1][k]);
        const float Bf = (float)(m[2][k] * G + m[3][k]);
        const float w = A.strength * A.ew[i];
        const float r_out = R * (1.f - w) + Rf * w;
        const float b_out = B * (1.f - w) + Bf * w;
        A.dst[3 * i] = r_out * norm_ref;
        A.dst[3 * i + 1] = s1;
        A.dst[3 * i + 2] = b_out * norm_ref;
    }
}

// ---------------- lateral CA ----------------

__global__ void split_rgb(const float *__restrict__ src, float *__restrict__ R, float *__restrict__ G,
                          float *__restrict__ B, size_t n) {
    for (size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (size_t)gridDim.x * BLOCK) {
        R[i] = src[3 * i];
        G[i] = src[3 * i + 1];
        B[i] = src[3 * i + 2];
    }
}

// 2x2 box (:3369-3384) of three planes
__global__ void down2(const float *R, const float *G, const float *B, float *R2, float *G2, float *B2,
                      int W, int H, int W2, int H2) {
    const size_t n = (size_t)W2 * H2;
    for (lre::GridWalk gw((size_t)blockIdx.x * BLOCK + threadIdx.x, n, (size_t)gridDim.x * BLOCK, W2); gw.more();
         gw.next()) {
        const size_t i = gw.i;
        const int y = gw.row, x = gw.col;
        const int y0 = 2 * y, y1 = min(y0 + 1, H - 1), x0 = 2 * x, x1 = min(x0 + 1, W - 1);
        const size_t a = (size_t)y0 * W + x0, b = (size_t)y0 * W + x1, c = (size_t)y1 * W + x0, d = (size_t)y1 * W + x1;
        R2[i] = 0.25f * (R[a] + R[b] + R[c] + R[d]);
        G2[i] = 0.25f * (G[a] + G[b] + G[c] + G[d]);
        B2[i] = 0.25f * (B[a] + B[b] + B[c] + B[d]);
    }
}

LRE_DEV float bilinear(const float *plane, int W, int H, float x, float y) {  // :3328-3346
    if (x < 0.f) x = 0.f;
    if (y < 0.f) y = 0.f;
    if (x > (float)(W - 1)) x = (float)(W - 1);
    if (y > (float)(H - 1)) y = (float)(H - 1);
    const int x0 = (int)x, y0 = (int)y;
    const int x1 = min(x0 + 1, W - 1), y1 = min(y0 + 1, H - 1);
    const float fx = x - (float)x0, fy = y - (float)y0;
    const float v00 = plane[(size_t)y0 * W + x0], v01 = plane[(size_t)y0 * W + x1];
    const float v10 = plane[(size_t)y1 * W + x0], v11 = plane[(size_t)y1 * W + x1];
    return (v00 * (1.f - fx) + v01 * fx) * (1.f - fy) + (v10 * (1.f - fx) + v11 * fx) * fy;
}

struct LkArgs {
    const float *G, *T[2];
    int W, H, mw, mh, iters;
    float max_shift, down, up;
    float *dx[2], *dy[2], *conf[2];
};

// one workgroup per (cell, channel): :3390-3471
__global__ __launch_bounds__(BLOCK) void lk_level(LkArgs A) {
    const int nc = A.mw * A.mh;
    const int ch = blockIdx.x / nc, c = blockIdx.x % nc;
    const int cy = c / A.mw, cx = c % A.mw;
    const float *G = A.G, *T = A.T[ch];
    const int W = A.W, H = A.H;
    float *dxm = A.dx[ch], *dym = A.dy[ch], *cm = A.conf[ch];
    const float cw = (float)W / (float)A.mw, chh = (float)H / (float)A.mh;
    const int x0 = (int)(cx * cw), y0 = (int)(cy * chh);
    int x1 = (int)((cx + 1) * cw), y1 = (int)((cy + 1) * chh);
    x1 = min(x1, W);
    y1 = min(y1, H);
    float adx = dxm[c] * A.down, ady = dym[c] * A.down;
    __shared__ double red[5][BLOCK / 64];
    __shared__ float s_acc[2];
    __shared__ int s_stop;
    __shared__ float s_eig;
    if (!(x1 <= x0 + 4 || y1 <= y0 + 4)) {
        float min_eig = 0.f;
        const int iw = x1 - x0 - 2, ih = y1 - y0 - 2;  // interior
        const int npix = iw > 0 && ih > 0 ? iw * ih : 0;
        for (int it = 0; it < A.iters; it++) {
            double a00 = 0, a01 = 0, a11 = 0, bx = 0, by = 0;
            // pixel t of the interior at (t / iw, t % iw): the quotient and
            // remainder advance by those of the stride (no division per pixel)
            const int dq = BLOCK / iw, dr = BLOCK - dq * iw;
            int q = (int)threadIdx.x / iw, rm = (int)threadIdx.x - q * iw;
            for (int t = threadIdx.x; t < npix; t += BLOCK) {
                const int y = y0 + 1 + q, x = x0 + 1 + rm;
                q += dq;
                rm += dr;
                if (rm >= iw) {
                    rm -= iw;
                    q++;
                }
                const size_t i = (size_t)y * W + x;
                const float gx = G[i + 1] - G[i - 1];
                const float gy = G[i + W] - G[i - W];
                const float tw = bilinear(T, W, H, (float)x + adx, (float)y + ady);
                const float err = tw - G[i];
                a00 += (double)gx * gx;
                a01 += (double)gx * gy;
                a11 += (double)gy * gy;
                bx -= (double)gx * err;
                by -= (double)gy * err;
            }
            for (int o = 32; o > 0; o >>= 1) {
                a00 += __shfl_xor(a00, o, 64);
                a01 += __shfl_xor(a01, o, 64);
                a11 += __shfl_xor(a11, o, 64);
                bx += __shfl_xor(bx, o, 64);
                by += __shfl_xor(by, o, 64);
            }
            if ((threadIdx.x & 63) == 0) {
                const int wv = threadIdx.x >> 6;
                red[0][wv] = a00; red[1][wv] = a01; red[2][wv] = a11; red[3][wv] = bx; red[4][wv] = by;
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                double A00 = 0, A01 = 0, A11 = 0, BX = 0, BY = 0;
                for (int k = 0; k < BLOCK / 64; k++) {
                    A00 += red[0][k]; A01 += red[1][k]; A11 += red[2][k]; BX += red[3][k]; BY += red[4][k];
                }
                int stop = 0;
                if (npix < 32) {
                    stop = 1;
                } else {
                    const double det = A00 * A11 - A01 * A01;
                    const double tr = A00 + A11;
                    const double dd = tr * tr - 4.0 * det;
                    const double disc = dd > 0.0 ? dd : 0.0;
                    min_eig = (float)(0.5 * (tr - sqrt(disc)));
                    if (det <= 1e-12) {
                        stop = 1;
                    } else {
                        const double inv = 1.0 / det;
                        const double sx = inv * (A11 * BX - A01 * BY);
                        const double sy = inv * (-A01 * BX + A00 * BY);
                        adx += (float)sx;
                        ady += (float)sy;
                        if (adx > A.max_shift) adx = A.max_shift;
                        if (adx < -A.max_shift) adx = -A.max_shift;
                        if (ady > A.max_shift) ady = A.max_shift;
                        if (ady < -A.max_shift) ady = -A.max_shift;
                        if (fabs(sx) < 0.005 && fabs(sy) < 0.005) stop = 1;
                    }
                }
                s_acc[0] = adx; s_acc[1] = ady; s_stop = stop; s_eig = min_eig;
            }
            __syncthreads();
            adx = s_acc[0];
            ady = s_acc[1];
            min_eig = s_eig;
            const int stop = s_stop;
            __syncthreads();
            if (stop) break;
        }
        if (threadIdx.x == 0) cm[c] = mx(0.f, min_eig);
    }
    if (threadIdx.x == 0) {
        dxm[c] = adx * A.up;
        dym[c] = ady * A.up;
    }
}

// confidence gate (:3584-3596) and 3x3 smoothing (:3474-3492); one block per channel
__global__ __launch_bounds__(1024) void lk_finish(float *dx0, float *dy0, float *c0, float *dx1, float *dy1,
                                                  float *c1, float *sdx0, float *sdy0, float *sdx1,
                                                  float *sdy1, int mw, int mh, float min_conf) {
    const int nc = mw * mh;
    float *dx = blockIdx.x ? dx1 : dx0, *dy = blockIdx.x ? dy1 : dy0, *cf = blockIdx.x ? c1 : c0;
    float *ox = blockIdx.x ? sdx1 : sdx0, *oy = blockIdx.x ? sdy1 : sdy0;
    __shared__ float red[16];
    float m = 0.f;
    for (int i = threadIdx.x; i < nc; i += blockDim.x) m = mx(m, cf[i]);
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    float maxc = 0.f;
    for (int k = 0; k < (int)(blockDim.x / 64); k++) maxc = mx(maxc, red[k]);
    if (maxc > 0.f) {
        const float thr = maxc * min_conf;
        for (int i = threadIdx.x; i < nc; i += blockDim.x)
            if (cf[i] < thr) { dx[i] = 0.f; dy[i] = 0.f; }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nc; i += blockDim.x) {
        const int cy = i / mw, cx = i % mw;
        float ax = 0.f, ay = 0.f;
        int n = 0;
        for (int ddy = -1; ddy <= 1; ddy++)
            for (int ddx = -1; ddx <= 1; ddx++) {
                const int yy = cy + ddy, xx = cx + ddx;
                if (yy < 0 || yy >= mh || xx < 0 || xx >= mw) continue;
                ax += dx[yy * mw + xx];
                ay += dy[yy * mw + xx];
                n++;
            }
        ox[i] = n > 0 ? ax / n : 0.f;
        oy[i] = n > 0 ? ay / n : 0.f;
    }
}

struct WarpArgs {
    const float *R, *G, *B;
    float *dst;
    const float *dx[2], *dy[2];
    int W, H, mw, mh, cell;
    float clamp_shift;
};

LRE_DEV void sample_shift(const WarpArgs &A, int ch, float px, float py, float &odx, float &ody) {
    const float gxf = (px / (float)A.cell) - 0.5f, gyf = (py / (float)A.cell) - 0.5f;
    int gx0 = (int)floorf(gxf), gy0 = (int)floorf(gyf);
    if (gx0 < 0) gx0 = 0;
    if (gy0 < 0) gy0 = 0;
    if (gx0 >= A.mw - 1) gx0 = A.mw - 2;
    if (gy0 >= A.mh - 1) gy0 = A.mh - 2;
    if (gx0 < 0) gx0 = 0;
    if (gy0 < 0) gy0 = 0;
    const float fx = lre::clamp_std(gxf - gx0, 0.f, 1.f), fy = lre::clamp_std(gyf - gy0, 0.f, 1.f);
    const int gx1 = min(gx0 + 1, A.mw - 1), gy1 = min(gy0 + 1, A.mh - 1);
    const float *dxm = A.dx[ch], *dym = A.dy[ch];
    const int mw = A.mw;
    const float dx = (dxm[gy0 * mw + gx0] * (1.f - fx) + dxm[gy0 * mw + gx1] * fx) * (1.f - fy) +
                     (dxm[gy1 * mw + gx0] * (1.f - fx) + dxm[gy1 * mw + gx1] * fx) * fy;
    const float dy = (dym[gy0 * mw + gx0] * (1.f - fx) + dym[gy0 * mw + gx1] * fx) * (1.f - fy) +
                     (dym[gy1 * mw + gx0] * (1.f - fx) + dym[gy1 * mw + gx1] * fx) * fy;
    odx = lre::clamp_std(dx, -A.clamp_shift, A.clamp_shift);
    ody = lre::clamp_std(dy, -A.clamp_shift, A.clamp_shift);
}

__global__ __launch_bounds__(BLOCK) void lateral_warp(WarpArgs A) {
    const size_t n = (size_t)A.W * A.H;
    for (lre::GridWalk gw((size_t)blockIdx.x * BLOCK + threadIdx.x, n, (size_t)gridDim.x * BLOCK, A.W); gw.more();
         gw.next()) {
        const size_t i = gw.i;
        const int y = gw.row, x = gw.col;
        float drx, dry, dbx, dby;
        sample_shift(A, 0, (float)x, (float)y, drx, dry);
        sample_shift(A, 1, (float)x, (float)y, dbx, dby);
        const float rn = bilinear(A.R, A.W, A.H, (float)x + drx, (float)y + dry);
        const float bn = bilinear(A.B, A.W, A.H, (float)x + dbx, (float)y + dby);
        A.dst[3 * i] = rn;
        A.dst[3 * i + 1] = A.G[i];
        A.dst[3 * i + 2] = bn;
    }
}

int grid_for(size_t n) {
    size_t b = (n + BLOCK - 1) / BLOCK;
    if (b > 8192) b = 8192;
    return (int)(b > 0 ? b : 1);
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

int lateral_levels(int W0, int H0, int levels) {
    int L = levels < 1 ? 1 : levels;
    if (L > 12) L = 12;
    while (L > 1 && ((W0 >> (L - 1)) < 32 || (H0 >> (L - 1)) < 32)) L--;
    return L;
}

}  // namespace

size_t lre_axial_scratch_bytes(int w, int h) { return align256((size_t)w * h * 4) * 5 + 256; }

int lre_check_axial_params(int radius) {
    if (radius > AX_RMAX) {
        lre_set_error("axial CA radius %d above the supported %d", radius, AX_RMAX);
        return LRE_ERR_ARG;
    }
    return LRE_OK;
}

namespace {

// raise the dynamic LDS limit of both passes once per process and source type
template <class Src>
hipError_t axial_lds_attr() {
    static const hipError_t rc = [] {
        hipError_t e = hipFuncSetAttribute((const void *)axial_ab<Src>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)axial_ab_lds(AX_RMAX));
        if (e == hipSuccess)
            e = hipFuncSetAttribute((const void *)axial_apply<Src>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)axial_apply_lds(AX_RMAX));
        return e;
    }();
    return rc;
}

// the arguments of one image; planes carved from its scratch (5 planes + stats)
AxialArgs axial_args(const float *d_in, float *d_out, int w, int h, int radius, float epsilon, float strength,
                     void *d_scratch) {
    const size_t n = (size_t)w * h, plane = align256(n * 4);
    char *base = static_cast<char *>(d_scratch);
    AxialArgs A;
    A.src = d_in; A.dst = d_out;
    A.aR = reinterpret_cast<float *>(base);
    A.bR = reinterpret_cast<float *>(base + plane);
    A.aB = reinterpret_cast<float *>(base + 2 * plane);
    A.bB = reinterpret_cast<float *>(base + 3 * plane);
    A.ew = reinterpret_cast<float *>(base + 4 * plane);
    A.stats = reinterpret_cast<unsigned *>(base + 5 * plane);
    A.W = w; A.H = h; A.r = radius < 1 ? 1 : radius; A.eps = epsilon;
    A.strength = strength < 0.f ? 0.f : (1.f < strength ? 1.f : strength);
    return A;
}

}  // namespace

int lre_launch_axial_ca(const float *d_in, float *d_out, int w, int h, int radius, float epsilon,
                        float strength, void *d_scratch, hipStream_t s) {
    if (const int rc = lre_check_axial_params(radius)) return rc;
    const AxialArgs A = axial_args(d_in, d_out, w, h, radius, epsilon, strength, d_scratch);
    const size_t n = (size_t)w * h;
    unsigned *stats = const_cast<unsigned *>(A.stats);
    hipLaunchKernelGGL(init_u32, dim3(1), dim3(64), 0, s, stats, 4);
    hipLaunchKernelGGL(axial_max_g, dim3(grid_for(n)), dim3(BLOCK), 0, s, d_in, n, stats);
    dim3 grid((w + AX_TX - 1) / AX_TX, (h + AX_TY - 1) / AX_TY);
    using One = lre::One<AxialArgs>;
    LRE_HIP_CHECK(axial_lds_attr<One>());
    hipLaunchKernelGGL(axial_ab<One>, grid, dim3(AX_BLOCK), axial_ab_lds(A.r), s, One{A});
    hipLaunchKernelGGL(axial_apply<One>, grid, dim3(AX_BLOCK), axial_apply_lds(A.r), s, One{A});
    LRE_HIP_CHECK(hipGetLastError());
    return LRE_OK;
}

// Many images: three tables over the same per-image arguments (one workgroup
// per image for the statistics reset, max(G) chunks, the window tiles).
namespace {
using AxTable = lre::ManyTable<AxialArgs>;
size_t ax_part(int n) { return align256(AxTable::bytes(n)); }
}  // namespace

size_t lre_axial_multi_table_bytes(int n) { return 3 * ax_part(n); }

int lre_axial_multi_fill(const lre_post_job *jobs, int n, int radius, float epsilon, float strength,
                         void *h_table, int grids[3]) {
    if (const int rc = lre_check_axial_params(radius)) return rc;
    auto args = [&](int i) {
        return axial_args(jobs[i].d_in, jobs[i].d_out, jobs[i].w, jobs[i].h, radius, epsilon, strength,
                          jobs[i].d_scratch);
    };
    char *h = static_cast<char *>(h_table);
    grids[0] = AxTable::fill(h, n, [](int) { return make_int2(1, 1); }, args);
    grids[1] = AxTable::fill(h + ax_part(n), n, [&](int i) {
        return make_int2((int)(((size_t)jobs[i].w * jobs[i].h + AX_CHUNK - 1) / AX_CHUNK), 1);
    }, args);
    grids[2] = AxTable::fill(h + 2 * ax_part(n), n, [&](int i) {
        return make_int2((jobs[i].w + AX_TX - 1) / AX_TX, (jobs[i].h + AX_TY - 1) / AX_TY);
    }, args);
    return LRE_OK;
}

int lre_launch_axial_multi(const void *d_table, int n, int radius, const int grids[3], hipStream_t s) {
    if (n == 0) return LRE_OK;
    if (const int rc = lre_check_axial_params(radius)) return rc;
    if (radius < 1) radius = 1;
    const char *d = static_cast<const char *>(d_table);
    using Many = lre::Many<AxialArgs>;
    LRE_HIP_CHECK(axial_lds_attr<Many>());
    hipLaunchKernelGGL(axial_stats_reset_many, dim3(grids[0]), dim3(64), 0, s, AxTable::view(d, n));
    hipLaunchKernelGGL(axial_max_g_many, dim3(grids[1]), dim3(BLOCK), 0, s, AxTable::view(d + ax_part(n), n));
    const Many tiles = AxTable::view(d + 2 * ax_part(n), n);
    hipLaunchKernelGGL(axial_ab<Many>, dim3(grids[2]), dim3(AX_BLOCK), axial_ab_lds(radius), s, tiles);
    hipLaunchKernelGGL(axial_apply<Many>, dim3(grids[2]), dim3(AX_BLOCK), axial_apply_lds(radius), s, tiles);
    LRE_HIP_CHECK(hipGetLastError());
    return LRE_OK;
}

size_t lre_lateral_scratch_bytes(int w, int h, int cell_size) {
    if (cell_size < 16) cell_size = 16;
    const size_t n = (size_t)w * h;
    const int mw = std::max(2, (w + cell_size - 1) / cell_size), mh = std::max(2, (h + cell_size - 1) / cell_size);
    // planes of all levels: < 3 * n * 4/3, plus 10 cell maps
    return align256(n * 4) * 3 * 2 + align256((size_t)mw * mh * 4) * 10 + 4096;
}

int lre_launch_lateral_ca(const float *d_in, float *d_out, int w, int h, int cell_size,
                          int max_iterations, float max_shift, float min_confidence,
                          int pyramid_levels, void *d_scratch, hipStream_t s) {
    if (cell_size < 16) cell_size = 16;
    if (max_iterations < 1) max_iterations = 1;
    if (d_in == d_out) {
        // the warp reads R/B planes, never the input buffer: in place is safe
    }
    const int L = lateral_levels(w, h, pyramid_levels);
    const int mw = std::max(2, (w + cell_size - 1) / cell_size), mh = std::max(2, (h + cell_size - 1) / cell_size);
    const int nc = mw * mh;
    char *base = static_cast<char *>(d_scratch);
    size_t off = 0;
    auto take = [&](size_t bytes) { float *p = reinterpret_cast<float *>(base + off); off += align256(bytes); return p; };
    float *Rl[12], *Gl[12], *Bl[12];
    int Wl[12], Hl[12];
    Wl[0] = w; Hl[0] = h;
    for (int l = 1; l < L; l++) { Wl[l] = Wl[l - 1] / 2; Hl[l] = Hl[l - 1] / 2; }
    for (int l = 0; l < L; l++) {
        const size_t nl = (size_t)Wl[l] * Hl[l];
        Rl[l] = take(nl * 4); Gl[l] = take(nl * 4); Bl[l] = take(nl * 4);
    }
    float *dxr = take(nc * 4), *dyr = take(nc * 4), *cr = take(nc * 4);
    float *dxb = take(nc * 4), *dyb = take(nc * 4), *cb = take(nc * 4);
    float *sdxr = take(nc * 4), *sdyr = take(nc * 4), *sdxb = take(nc * 4), *sdyb = take(nc * 4);
    {
        const size_t nz = (size_t)((char *)(cb + nc) - (char *)dxr) / sizeof(float);
        hipLaunchKernelGGL(zero_f32, dim3(grid_for(nz)), dim3(BLOCK), 0, s, dxr, nz);
    }
    hipLaunchKernelGGL(split_rgb, dim3(grid_for((size_t)w * h)), dim3(BLOCK), 0, s, d_in, Rl[0], Gl[0], Bl[0], (size_t)w * h);
    for (int l = 1; l < L; l++)
        hipLaunchKernelGGL(down2, dim3(grid_for((size_t)Wl[l] * Hl[l])), dim3(BLOCK), 0, s, Rl[l - 1], Gl[l - 1],
                           Bl[l - 1], Rl[l], Gl[l], Bl[l], Wl[l - 1], Hl[l - 1], Wl[l], Hl[l]);
    for (int l = L - 1; l >= 0; l--) {
        LkArgs A;
        A.G = Gl[l]; A.T[0] = Rl[l]; A.T[1] = Bl[l];
        A.W = Wl[l]; A.H = Hl[l]; A.mw = mw; A.mh = mh; A.iters = max_iterations;
        A.down = 1.f / (float)(1 << l);
        A.up = (float)(1 << l);
        A.max_shift = max_shift * A.down + (float)(1 << l) * 0.5f;
        A.dx[0] = dxr; A.dy[0] = dyr; A.conf[0] = cr;
        A.dx[1] = dxb; A.dy[1] = dyb; A.conf[1] = cb;
        hipLaunchKernelGGL(lk_level, dim3(2 * nc), dim3(BLOCK), 0, s, A);
    }
    hipLaunchKernelGGL(lk_finish, dim3(2), dim3(1024), 0, s, dxr, dyr, cr, dxb, dyb, cb, sdxr, sdyr, sdxb, sdyb,
                       mw, mh, min_confidence);
    WarpArgs Wa;
    Wa.R = Rl[0]; Wa.G = Gl[0]; Wa.B = Bl[0]; Wa.dst = d_out;
    Wa.dx[0] = sdxr; Wa.dy[0] = sdyr; Wa.dx[1] = sdxb; Wa.dy[1] = sdyb;
    Wa.W = w; Wa.H = h; Wa.mw = mw; Wa.mh = mh; Wa.cell = cell_size;
    Wa.clamp_shift = max_shift * (float)(1 << (L - 1));
    hipLaunchKernelGGL(lateral_warp, dim3(grid_for((size_t)w * h)), dim3(BLOCK), 0, s, Wa);
    LRE_HIP_CHECK(hipGetLastError());
    return LRE_OK;
}
