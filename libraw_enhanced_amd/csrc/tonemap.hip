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
// The guide g = max(0, R, G, B), box means of radius 12 of g and g*g, the
// gain coefficients a = var / (var + eps) and b = mean (1 - a), their box
// means A and B, and per pixel base = max(1e-6, A g + B), whose ACES gain is
// blended in by a smoothstep over [0.55, 0.90] (the pixel is left as it is
// where that weight is 0).  Box means clip the window at the frame and divide
// by the true tap count (core/libraw_wrapper.cpp:45-116).
//
// One workgroup of two wavefronts develops a strip of 64 output columns over a
// segment of rows, walking down the rows with every intermediate on chip, 8
// rows (a batch) at a time:
//   wave 0  V1: vertical running sums of g and g*g over the 25 rows of the
//               window, per column of the 112 guide columns (the strip and a
//               24-column ring on each side); each lane keeps its two columns'
//               last 32 guide values in registers for the value that leaves;
//           H1: horizontal running sums of those over 25 columns, giving the
//               means and a, b on the 88 columns the second box reads (lane:
//               a run of 11 of one row of the batch);
//   wave 1  H2: horizontal running sums of a and b (lane: a run of 8 of one
//               row);
//           V2: vertical running sums of those, per output column, with its
//               last 32 rows in registers, then the output row 12 rows behind.
// Wave 1 works on the batch before the one wave 0 works on; a and b pass
// between them through two LDS buffers, one barrier per batch.  The two waves'
// rings and sums share registers.  The first box sums columns then rows and
// the second rows then columns; the reference sums rows then columns in both,
// and the two orders differ by the rounding of the running sums only (the
// reference's own runs span whole rows and columns).  A segment starts 36 rows
// above its first output row with empty sums and rings: the first three
// batches only load the guide.
//
// The guide comes from a guide plane written beforehand (safe in place: the
// kernel reads the image only at its own output pixels, each lane reading its
// pixel before writing it), or, when the output is another buffer, straight
// from the input pixels.
constexpr int DR = 12;
constexpr int DS_W = 64;               // output columns per strip, one per lane
constexpr int DS_GX = DS_W + 4 * DR;   // guide columns of a strip
constexpr int DS_MX = DS_W + 2 * DR;   // a, b columns of a strip
constexpr int DS_RB = 8;               // rows per batch
constexpr int DS_RING = 32;            // ring of past rows (>= 2 DR + 2, a multiple of DS_RB)
// LDS planes (one float per value): pitches that spread each pass's 64
// lanes over the banks, lane 8 r + k (row r, run k) in H1 and 8 k + r in H2.
// V1 means: H1 reads runs of 11 (no conflict); a, b: written in runs of 11,
// read in runs of 8 (at most two lanes per bank); H2 means: written in runs of
// 8 (no conflict).  V1 and V2 touch consecutive columns.
constexpr int DS_VP = 120, DS_AP = 90, DS_HP = 65;
static_assert(DS_VP >= DS_GX && DS_AP >= DS_MX && DS_HP >= DS_W, "LDS pitches hold a row");
static_assert(DS_RING >= 2 * DR + 2 && DS_RING == 4 * DS_RB, "a ring cycle is 4 batches");
static_assert(DS_MX == 8 * 11 && DS_W == 8 * 8 && DS_RB * 8 == 64, "a batch is 8 rows x 8 runs, one per lane");
static_assert(DS_GX - 64 <= 64, "two guide columns per lane");

// one image of a detail tone map launch
struct DtmImg {
    const float *guide;  // guide plane; null: computed from `in` (only when in != out)
    const float *in;
    float *out;
    int W, H;
    int seg;             // output rows per segment (32 k - 48)
};

LRE_DEV float dtm_gd(float r, float g, float b) { return lre::max_std(0.0f, lre::max_std(r, lre::max_std(g, b))); }

// 1 / (taps of the radius-DR window at p in [0, len)); positions outside have
// no use for the value and get 1.  A window inside the frame has 2 DR + 1 taps.
LRE_DEV float dtm_inv_count(int p, int len) {
    if (p >= DR && p + DR < len) return 1.0f / (2 * DR + 1);
    const int c = min(len - 1, p + DR) - max(0, p - DR) + 1;
    return __builtin_amdgcn_rcpf((float)(c > 0 ? c : 1));
}

// n / d from the hardware reciprocal (within 2 ulp; d > 0 here)
LRE_DEV float dtm_div(float n, float d) { return n * __builtin_amdgcn_rcpf(d); }

// LDS writes of this wavefront's lanes are seen by its later LDS reads
LRE_DEV void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// the LDS of a strip workgroup
struct DtmLds {
    float vg[DS_RB * DS_VP], vq[DS_RB * DS_VP];          // wave 0: V1 means of g, g*g
    float a[2][DS_RB * DS_AP], b[2][DS_RB * DS_AP];      // a, b from wave 0 to wave 1, by batch parity
    float ha[DS_RB * DS_HP], hb[DS_RB * DS_HP];          // wave 1: H2 row means
};

template <bool FUSED, bool EPI>
struct DtmStrip {
    // raw values in flight: wave 0 the next batch's guide samples (a pixel per
    // sample when fused), wave 1 the next batch's output pixels
    static constexpr int LD = FUSED ? 3 * 2 * DS_RB : 3 * DS_RB;
    DtmImg I;
    int x0, y0, yend, lane;
    int gx0, gx1, ox;
    bool in0, in1, oin;
    float ring0[DS_RING], ring1[DS_RING];  // wave 0: g at gx0, gx1; wave 1: row means of a, b
    float acc[4];                          // wave 0: sums of g, g*g at gx0, gx1; wave 1: of a, b
    float ld[LD];
    float gc[DS_RB][2];                    // wave 0: guide values of the batch's V1 rows

    // wave 0: raw guide samples of guide rows yb + DR .. yb + DR + 7 (0 outside
    // the frame)
    LRE_DEV void load_guide(int yb) {
#pragma unroll
        for (int r = 0; r < DS_RB; r++)
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const int y = yb + r + DR, x = c ? gx1 : gx0;
                const bool ok = (c ? in1 : in0) && y >= 0 && y < I.H;
                const size_t i = ok ? (size_t)y * I.W + x : 0;
                if (FUSED) {
                    float *v = ld + 3 * (2 * r + c);
                    v[0] = v[1] = v[2] = 0.f;
                    if (ok) {
                        v[0] = I.in[3 * i];
                        v[1] = I.in[3 * i + 1];
                        v[2] = I.in[3 * i + 2];
                    }
                } else {
                    ld[2 * r + c] = ok ? I.guide[i] : 0.f;
                }
            }
    }
    LRE_DEV void take_guide() {
#pragma unroll
        for (int r = 0; r < DS_RB; r++)
#pragma unroll
            for (int c = 0; c < 2; c++) {
                if (FUSED) {
                    const float *v = ld + 3 * (2 * r + c);
                    gc[r][c] = dtm_gd(v[0], v[1], v[2]);
                } else {
                    gc[r][c] = ld[2 * r + c];
                }
            }
    }
    // wave 1: the output pixels of rows yb - DR .. yb - DR + 7 (rows outside
    // the segment are not used: they read a clamped row)
    LRE_DEV void load_pixels(int yb) {
        const int cx = min(ox, I.W - 1);
#pragma unroll
        for (int r = 0; r < DS_RB; r++) {
            const int y = min(max(yb + r - DR, 0), I.H - 1);
            const float *p = I.in + 3 * ((size_t)y * I.W + cx);
            ld[3 * r] = __builtin_nontemporal_load(p);
            ld[3 * r + 1] = __builtin_nontemporal_load(p + 1);
            ld[3 * r + 2] = __builtin_nontemporal_load(p + 2);
        }
    }

    // wave 0, batch in ring quarter Q: a, b rows yb .. yb + 7 into ab
    template <int Q>
    LRE_DEV void first_box(int yb, bool warm, float *s_vg, float *s_vq, float *s_a, float *s_b) {
        // V1: add guide row y + DR, drop row y - DR - 1 (ring steps u + 12, u - 13)
#pragma unroll
        for (int r = 0; r < DS_RB; r++) {
            const int u = Q * DS_RB + r, y = yb + r;
            const float g0 = gc[r][0], g1 = gc[r][1];
            const float o0 = ring0[(u + DS_RING - DR - 1) % DS_RING], o1 = ring1[(u + DS_RING - DR - 1) % DS_RING];
            ring0[(u + DR) % DS_RING] = g0;
            ring1[(u + DR) % DS_RING] = g1;
            acc[0] = (acc[0] - o0) + g0;
            acc[1] = (acc[1] - o0 * o0) + g0 * g0;
            acc[2] = (acc[2] - o1) + g1;
            acc[3] = (acc[3] - o1 * o1) + g1 * g1;
            if (!warm) {
                const float ic = dtm_inv_count(y, I.H);
                s_vg[r * DS_VP + lane] = acc[0] * ic;
                s_vq[r * DS_VP + lane] = acc[1] * ic;
                if (lane < DS_GX - 64) {
                    s_vg[r * DS_VP + 64 + lane] = acc[2] * ic;
                    s_vq[r * DS_VP + 64 + lane] = acc[3] * ic;
                }
            }
        }
        // the next batch's guide rows load behind this batch's row sums
        load_guide(yb + DS_RB);
        if (!warm) {
            wave_lds_sync();
            // H1: lane (row r, run k) -> a, b at a, b columns 11 k .. 11 k + 10
            const int r = lane >> 3, j0 = 11 * (lane & 7), y = yb + r;
            const float *vg = s_vg + r * DS_VP + j0, *vq = s_vq + r * DS_VP + j0;
            float s0 = 0.f, s1 = 0.f;
#pragma unroll 5
            for (int t = 0; t <= 2 * DR; t++) {
                s0 += vg[t];
                s1 += vq[t];
            }
            const bool yin = y >= 0 && y < I.H;
            const int xa = x0 - DR + j0;  // the run's first column
            const bool inner = xa >= DR && xa + 10 + DR < I.W;
#pragma unroll 2
            for (int j = 0; j < 11; j++) {
                const int x = xa + j;
                const float ic = inner ? 1.0f / (2 * DR + 1) : dtm_inv_count(x, I.W);
                const float mean = s0 * ic, msq = s1 * ic;
                float a = 0.f, b = 0.f;  // outside the image: no taps
                if (yin && x >= 0 && x < I.W) {
                    const float var = lre::max_std(0.0f, msq - mean * mean);
                    a = dtm_div(var, var + 0.0025f);
                    b = mean * (1.0f - a);
                }
                s_a[r * DS_AP + j0 + j] = a;
                s_b[r * DS_AP + j0 + j] = b;
                if (j < 10) {
                    s0 = (s0 - vg[j]) + vg[j + 2 * DR + 1];
                    s1 = (s1 - vq[j]) + vq[j + 2 * DR + 1];
                }
            }
        }
        take_guide();
    }

    // wave 1, batch in ring quarter Q: a, b rows yb .. yb + 7 from ab, output
    // rows yb - DR .. yb - DR + 7
    template <int Q>
    LRE_DEV void second_box(int yb, const float *s_a, const float *s_b, float *s_ha, float *s_hb,
                               const lre::Epilogue &E) {
        {
            // H2: lane (run k, row r) -> row means of a, b at output columns 8 k .. 8 k + 7
            const int r = lane & 7, j0 = 8 * (lane >> 3);
            const float *va = s_a + r * DS_AP + j0, *vb = s_b + r * DS_AP + j0;
            float s0 = 0.f, s1 = 0.f;
#pragma unroll 5
            for (int t = 0; t <= 2 * DR; t++) {
                s0 += va[t];
                s1 += vb[t];
            }
            const bool inner = x0 + j0 >= DR && x0 + j0 + 7 + DR < I.W;
#pragma unroll 2
            for (int j = 0; j < 8; j++) {
                const float ic = inner ? 1.0f / (2 * DR + 1) : dtm_inv_count(x0 + j0 + j, I.W);
                s_ha[r * DS_HP + j0 + j] = s0 * ic;
                s_hb[r * DS_HP + j0 + j] = s1 * ic;
                if (j < 7) {
                    s0 = (s0 - va[j]) + va[j + 2 * DR + 1];
                    s1 = (s1 - vb[j]) + vb[j + 2 * DR + 1];
                }
            }
        }
        wave_lds_sync();
        // V2 and the output rows: add row u, drop row u - 25
        const bool inplace = I.in == I.out;
#pragma unroll
        for (int r = 0; r < DS_RB; r++) {
            const int u = Q * DS_RB + r, yo = yb + r - DR;
            const float2 h = make_float2(s_ha[r * DS_HP + lane], s_hb[r * DS_HP + lane]);
            const float oa = ring0[(u + DS_RING - 2 * DR - 1) % DS_RING];
            const float ob = ring1[(u + DS_RING - 2 * DR - 1) % DS_RING];
            ring0[u % DS_RING] = h.x;
            ring1[u % DS_RING] = h.y;
            acc[0] = (acc[0] - oa) + h.x;
            acc[1] = (acc[1] - ob) + h.y;
            if (!oin || yo < y0 || yo >= yend) continue;
            const float ic = dtm_inv_count(yo, I.H);
            const float A = acc[0] * ic, B = acc[1] * ic;
            float R = ld[3 * r], G = ld[3 * r + 1], Bl = ld[3 * r + 2];
            const float base = lre::max_std(1e-6f, A * dtm_gd(R, G, Bl) + B);
            // aces(base) / base with the common factor base cancelled
            float gain = dtm_div(2.51f * base + 0.03f, base * (2.43f * base + 0.59f) + 0.14f);
            gain = lre::clamp_std(gain, 0.0f, 1.0f);
            float w = lre::clamp_std((base - 0.55f) * (1.0f / (0.90f - 0.55f)), 0.0f, 1.0f);
            w = w * w * (3.0f - 2.0f * w);
            // where the weight is 0 the pixel stays as it is (times exactly 1)
            const float f = w > 0.0f ? 1.0f + (gain - 1.0f) * w : 1.0f;
            R *= f;
            G *= f;
            Bl *= f;
            if (EPI) lre::epilogue_rgb(R, G, Bl, E);
            if (EPI || !inplace || w > 0.0f) {  // unchanged in place: nothing to store
                float *o = I.out + 3 * ((size_t)yo * I.W + ox);
                __builtin_nontemporal_store(R, o);
                __builtin_nontemporal_store(G, o + 1);
                __builtin_nontemporal_store(Bl, o + 2);
            }
            // one row's output at a time: interleaving the rows' arithmetic
            // would hold more registers than the occupancy leaves
            __builtin_amdgcn_sched_barrier(0);
        }
        load_pixels(yb + DS_RB);
    }

    // step Q of a 32-row iteration: wave 0 on batch bn, wave 1 on batch bn - 1
    template <int Q>
    LRE_DEV void step(int bn, int nb, int ystart, DtmLds &L, const lre::Epilogue &E) {
        if (threadIdx.x < 64) {
            if (bn < nb) first_box<Q>(ystart + DS_RB * bn, bn < 3, L.vg, L.vq, L.a[Q & 1], L.b[Q & 1]);
        } else if (bn >= 4 && bn <= nb) {
            second_box<(Q + 3) % 4>(ystart + DS_RB * (bn - 1), L.a[(Q + 1) & 1], L.b[(Q + 1) & 1], L.ha, L.hb, E);
        }
        __syncthreads();
    }
};

// The strip segment of workgroup p of n.  Workgroups are dealt to the 8 XCDs
// round-robin; each XCD's workgroups take one contiguous range of the
// row-major segment list, so the strips beside each other in a segment row,
// which read 48 of each other's guide columns, run on one XCD and share its L2.
LRE_DEV int dtm_xcd_tile(int p, int n) {
    const int x = p & 7, slot = p >> 3;
    return x * (n >> 3) + min(x, n & 7) + slot;
}
LRE_DEV lre::TileOf dtm_tile_of(const lre::One<DtmImg> &s, int t) {
    const int strips = (s.a.W + DS_W - 1) / DS_W;
    return {0, t % strips, t / strips};
}
LRE_DEV lre::TileOf dtm_tile_of(const lre::Many<DtmImg> &s, int t) { return s.tile_at(t); }

// Src: lre::One<DtmImg> or lre::Many<DtmImg> (the strip segments of several
// images); a 1D grid of one workgroup of 128 threads per strip segment.
template <bool FUSED, bool EPI, class Src>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(EPI ? 2 : 3))) void dtm_strip(Src src,
                                                                                              lre::Epilogue E) {
    __shared__ DtmLds L;
    const lre::TileOf tl = dtm_tile_of(src, dtm_xcd_tile((int)blockIdx.x, (int)gridDim.x));
    DtmStrip<FUSED, EPI> S;
    S.I = src.args(tl);
    const int SH = S.I.seg;
    S.lane = (int)threadIdx.x & 63;
    S.x0 = tl.tx * DS_W;
    S.y0 = tl.ty * SH;
    S.yend = min(S.y0 + SH, S.I.H);
    S.gx0 = S.x0 - 2 * DR + S.lane;
    S.gx1 = S.gx0 + 64;
    S.ox = S.x0 + S.lane;
    S.in0 = S.gx0 >= 0 && S.gx0 < S.I.W;
    S.in1 = S.lane < DS_GX - 64 && S.gx1 >= 0 && S.gx1 < S.I.W;
    S.oin = S.ox < S.I.W;
#pragma unroll
    for (int k = 0; k < DS_RING; k++) S.ring0[k] = S.ring1[k] = 0.f;
#pragma unroll
    for (int k = 0; k < 4; k++) S.acc[k] = 0.f;
    // a, b rows from y0 - 3 DR: 24 lead-in rows fill the first box's window;
    // wave 1 starts at batch 3, the first with a, b rows
    const int ystart = S.y0 - 3 * DR, iters = (SH + 4 * DR) / DS_RING, nb = 4 * iters;
    if (threadIdx.x < 64) {
        S.load_guide(ystart);
        S.take_guide();
    } else {
        S.load_pixels(ystart + 3 * DS_RB);
    }
    for (int it = 0; it <= iters; it++) {
        const int b = 4 * it;
        S.template step<0>(b, nb, ystart, L, E);
        if (it == iters) break;  // wave 1's last batch ran in step 0
        S.template step<1>(b + 1, nb, ystart, L, E);
        S.template step<2>(b + 2, nb, ystart, L, E);
        S.template step<3>(b + 3, nb, ystart, L, E);
    }
}

// Output rows per segment: 32 k - 48 for a whole number k of 32-row
// iterations, each segment reading 48 rows of lead-in.  Chosen from the frame
// size only (the batch path must match the per-image path bit for bit): the
// height whose rounds of resident workgroups times segment length is least,
// for an MI355X holding 1536 of these workgroups at once (256 CUs x 6: three
// waves per SIMD).  On 24 MP that is one round of 272-row segments; a frame
// too small to fill a round takes the shortest segments.
constexpr int DTM_RESIDENT = 1536;
int dtm_seg_rows(int w, int h) {
    const long long strips = (w + DS_W - 1) / DS_W;
    int best = 80;
    long long best_cost = -1;
    for (int k = 4; k <= 16; k++) {
        const int sh = 32 * k - 48;
        const long long wg = strips * ((h + sh - 1) / sh);
        const long long cost = ((wg + DTM_RESIDENT - 1) / DTM_RESIDENT) * (sh + 48);
        if (best_cost < 0 || cost < best_cost) {
            best_cost = cost;
            best = sh;
        }
    }
    return best;
}

template <class Src>
int dtm_strip_launch(Src src, dim3 grid, bool fused, hipStream_t s, const lre::Epilogue *epi) {
    if (grid.x == 0 || grid.y == 0) return LRE_OK;
    const bool e = epi && epi->active();
    const lre::Epilogue E = e ? *epi : lre::Epilogue{};
    if (fused) {
        if (e) hipLaunchKernelGGL((dtm_strip<true, true, Src>), grid, dim3(128), 0, s, src, E);
        else hipLaunchKernelGGL((dtm_strip<true, false, Src>), grid, dim3(128), 0, s, src, E);
    } else {
        if (e) hipLaunchKernelGGL((dtm_strip<false, true, Src>), grid, dim3(128), 0, s, src, E);
        else hipLaunchKernelGGL((dtm_strip<false, false, Src>), grid, dim3(128), 0, s, src, E);
    }
    LRE_HIP_CHECK(hipGetLastError());
    return LRE_OK;
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

// the guide plane (used when the stage runs in place)
size_t lre_detail_tm_scratch_bytes(int w, int h) { return (size_t)w * h * sizeof(float); }

// In place, the guide plane is written first; with a separate output the strip
// kernel takes the guide from the input pixels and the plane is not used.
int lre_launch_detail_tonemap(const float *d_in, float *d_out, int w, int h, void *d_scratch,
                              hipStream_t s, const lre::Epilogue *epi) {
    if (w < 1 || h < 1) return LRE_OK;
    const bool fused = d_in != d_out;
    float *guide = nullptr;
    if (!fused) {
        guide = static_cast<float *>(d_scratch);
        const size_t n = (size_t)w * h;
        const size_t n4 = (reinterpret_cast<uintptr_t>(d_in) % 16 == 0) ? n / 4 : 0;
        hipLaunchKernelGGL(dtm_guide, dim3(grid_for(n4 > 0 ? n4 : n, 256)), dim3(256), 0, s, d_in, guide, n, n4);
    }
    const int seg = dtm_seg_rows(w, h);
    const lre::One<DtmImg> one{DtmImg{guide, d_in, d_out, w, h, seg}};
    return dtm_strip_launch(one, dim3(((w + DS_W - 1) / DS_W) * ((h + seg - 1) / seg)), fused, s, epi);
}

// Many images: two tables (guide chunks, then strip segments) over the same
// per-image arguments, each guide plane at the start of the image's scratch.
using DtmTable = lre::ManyTable<DtmImg>;
size_t lre_detail_tm_multi_table_bytes(int n) { return 2 * ((DtmTable::bytes(n) + 255) & ~(size_t)255); }

int lre_detail_tm_multi_fill(const lre_post_job *jobs, int n, void *h_table, int grids[2]) {
    const size_t half = (DtmTable::bytes(n) + 255) & ~(size_t)255;
    auto args = [&](int i) {
        return DtmImg{static_cast<const float *>(jobs[i].d_scratch), jobs[i].d_in, jobs[i].d_out, jobs[i].w,
                      jobs[i].h, dtm_seg_rows(jobs[i].w, jobs[i].h)};
    };
    grids[0] = DtmTable::fill(h_table, n, [&](int i) {
        return make_int2((int)(((size_t)jobs[i].w * jobs[i].h + DTM_CHUNK - 1) / DTM_CHUNK), 1);
    }, args);
    grids[1] = DtmTable::fill(static_cast<char *>(h_table) + half, n, [&](int i) {
        const int seg = dtm_seg_rows(jobs[i].w, jobs[i].h);
        return make_int2((jobs[i].w + DS_W - 1) / DS_W, (jobs[i].h + seg - 1) / seg);
    }, args);
    return LRE_OK;
}

int lre_launch_detail_tm_multi(const void *d_table, int n, const int grids[2], hipStream_t s,
                               const lre::Epilogue *epi) {
    if (n == 0 || grids[1] == 0) return LRE_OK;
    const size_t half = (DtmTable::bytes(n) + 255) & ~(size_t)255;
    hipLaunchKernelGGL(dtm_guide_many, dim3(grids[0]), dim3(256), 0, s, DtmTable::view(d_table, n));
    LRE_HIP_CHECK(hipGetLastError());
    return dtm_strip_launch(DtmTable::view(static_cast<const char *>(d_table) + half, n), dim3(grids[1]), false, s,
                            epi);
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
