// AMaZE Bayer demosaic for gfx950.
//
// Semantics: the reference's tiled AMaZE (core/cpu_accelerator.cpp:1145-1630)
// followed by its 4-px border interpolation (:2984-3028).  The reference's output
// depends on its tile grid (160x160 tiles, step 128, 16-px reflect-101 border),
// on in-place raster-order loops, and on several scratch arrays that alias each
// other, so each workgroup here processes exactly one reference tile at a time
// in a private arena with the reference's offsets:
//
//   * loops without intra-loop dependencies run across the whole workgroup;
//   * the variance test of stage 3 (:1376-1397) updates hcd left-to-right and
//     vcd top-to-bottom, each reading its already-updated predecessor two sites
//     back: it runs as one independent scan per (row, parity) for hcd and per
//     (column, parity) for vcd;
//   * the hvwt smoothing (:1446-1448) and the pmwt smoothing (:1537-1539) read
//     their updated upper-diagonal neighbours: they run row by row, in parallel
//     within a row;
//   * the arena is cleared before each tile, so every tile is a pure function
//     of its input (the reference's per-thread arena carries values from the
//     previous tile of the same OpenMP thread into a few unfilled cells).
//
// The kernel is persistent: a grid of `slots` workgroups takes tiles from a
// queue (one atomic counter), each workgroup owning one arena (1.41 MiB).  All arithmetic is FP32 with the operand
// order of the reference (built with -ffp-contract=off).
#include "lre_common.h"
#include "lre_internal.h"

namespace {

constexpr int TS = 160, TSH = 80, N2 = TS * TS;
constexpr int BLOCK = 1024;  // one tile per workgroup of 16 waves (measured best)
constexpr float EPS = 1e-5f, EPSSQ = 1e-10f, ARTHRESH = 0.75f;

// Arena layout in floats (reference :1222-1240).
constexpr int O_GREEN = 0, O_DSQ = N2, O_WV = 2 * N2, O_WH = 3 * N2, O_VCD = 4 * N2,
              O_HCD = 5 * N2, O_VCDALT = 6 * N2, O_HCDALT = 7 * N2, O_CDSQ = 8 * N2,
              O_HVWT = 9 * N2, O_DGV = 9 * N2 + TS * TSH, O_DGH = 10 * N2 + TS * TSH,
              O_Q1M = 11 * N2 + TS * TSH, O_Q1P = 11 * N2 + 2 * TS * TSH,
              O_CFA = 11 * N2 + 3 * TS * TSH, O_NYQ = 12 * N2 + 3 * TS * TSH,
              O_NYQTEST = O_NYQ + TS * TSH / 4, ARENA_FLOATS = O_NYQTEST + TS * TSH;
static_assert(ARENA_FLOATS % 64 == 0, "arena stride keeps 256-byte alignment");

struct Args {
    const float *cfa_plane;
    float *rgb;
    int W, H;
    unsigned cfw;  // CFA colour of (r & 1, c & 1) in bits 16 (r & 1) + 8 (c & 1)
    int ex, ey;
    float clip_pt, clip_pt8, scale;
    const float *d_maxv;  // when set, the demosaic maximum is read from the device
    int ntc, ntiles;
    int tile_base;        // first tile of this image in a multi-image tile list
    float *arenas;
};

__device__ const float GODD[4] = {0.14659727707323927f, 0.103592713382435f, 0.0732036125103057f,
                                  0.0365543548389495f};
__device__ const float GGRAD[6] = {0.5f * 0.07384411893421103f, 0.5f * 0.06207511968171489f,
                                   0.5f * 0.0521818194747806f, 0.5f * 0.03687419286733595f,
                                   0.5f * 0.03099732204057846f, 0.5f * 0.018413194161458882f};
__device__ const float GEVEN[2] = {0.13719494435797422f, 0.05640252782101291f};
__device__ const float GQUINC[4] = {0.169917f, 0.108947f, 0.069855f, 0.0287182f};

LRE_DEV float half(float x) { return x * 0.5f; }
LRE_DEV float twice(float x) { return x + x; }
LRE_DEV int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// CFA colour at row r, column c of the image or of a tile (tile origins are
// even).  The four colours sit in one word: an indexed array in the kernel
// arguments would cost a memory load per lookup.
LRE_DEV unsigned fct(const Args &A, int r, int c) { return (A.cfw >> (((r & 1) << 4) | ((c & 1) << 3))) & 0xffu; }

// Index t = t0, t0 + step, ... < n of a row-major range with rows of `cols`
// entries, as (row, column): one division at the start, then the quotient and
// remainder advance by those of the step (an integer division per iteration
// costs some thirty VALU instructions).
struct RowWalk {
    int t, n, cols, q, r, dq, dr, step;
    LRE_DEV RowWalk(int t0, int n_, int cols_, int step_) : t(t0), n(n_), cols(cols_), step(step_) {
        q = t0 / cols_;
        r = t0 - q * cols_;
        dq = step_ / cols_;
        dr = step_ - dq * cols_;
    }
    LRE_DEV bool more() const { return t < n; }
    LRE_DEV void next() {
        t += step;
        q += dq;
        r += dr;
        if (r >= cols) {
            r -= cols;
            q++;
        }
    }
};

template <class F>
LRE_DEV void for_rect(int r0, int r1, int c0, int c1, F f) {
    const int nc = c1 - c0, nr = r1 - r0;
    if (nc <= 0 || nr <= 0) return;
    for (RowWalk w((int)threadIdx.x, nr * nc, nc, BLOCK); w.more(); w.next()) f(r0 + w.q, c0 + w.r);
}

// Non-green ("site") positions of rows [r0,r1): cc = c0 + parity(row) + 2j < c1,
// spread over the threads [tid0, tid0 + nthr) of the workgroup.
template <class F>
LRE_DEV void for_sites_on(const Args &A, int r0, int r1, int c0, int c1, int tid0, int nthr, F f) {
    const int ns = (c1 - c0 + 1) / 2, nr = r1 - r0;
    const int t0 = (int)threadIdx.x - tid0;
    if (ns <= 0 || nr <= 0 || t0 < 0) return;
    for (RowWalk w(t0, nr * ns, ns, nthr); w.more(); w.next()) {
        const int rr = r0 + w.q;
        const int cc = c0 + (int)(fct(A, rr, 0) & 1) + 2 * w.r;
        if (cc < c1) f(rr, cc);
    }
}
template <class F>
LRE_DEV void for_sites(const Args &A, int r0, int r1, int c0, int c1, F f) {
    for_sites_on(A, r0, r1, c0, c1, 0, BLOCK, f);
}


// for_sites with the work split into compute(rr, cc) -> R and store(rr, cc, R),
// one site per iteration.
template <class C, class S>
LRE_DEV void for_sites_split(const Args &A, int r0, int r1, int c0, int c1, C compute, S store) {
    for_sites(A, r0, r1, c0, c1, [&](int rr, int cc) { store(rr, cc, compute(rr, cc)); });
}

// The same sites as for_sites, two per thread: (rr, cc) and (rr + 2, cc) share
// a column and a colour.  Each site's work is load(rr, cc) -> L, which only
// loads, then finish(rr, cc, L) -> R; both sites' loads are issued before
// either finish runs (and may branch), and both results are stored after both
// computations, so the loads of the two sites overlap and the loads they have
// in common are issued once.  Rows go in groups of four: pairs (r, r+2) and
// (r+1, r+3).
template <class LD, class C, class S>
LRE_DEV void for_site_pairs_ld(const Args &A, int r0, int r1, int c0, int c1, LD load, C finish, S store) {
    const int ns = (c1 - c0 + 1) / 2, nr = r1 - r0;
    if (ns <= 0 || nr <= 0) return;
    const int n = ((nr + 3) / 4) * 2 * ns;
    for (RowWalk w((int)threadIdx.x, n, 2 * ns, BLOCK); w.more(); w.next()) {
        const int g = w.q, rem = w.r;
        const int p = rem >= ns ? 1 : 0, j = rem - p * ns;
        const int ra = r0 + 4 * g + p, rb = ra + 2;
        const int cc = c0 + (int)(fct(A, ra, 0) & 1) + 2 * j;
        if (ra >= r1 || cc >= c1) continue;
        const bool vb = rb < r1;
        const int rb2 = vb ? rb : ra;  // a row past the range repeats the first site
        const auto La = load(ra, cc);
        const auto Lb = load(rb2, cc);
        const auto Ra = finish(ra, cc, La);
        const auto Rb = finish(rb2, cc, Lb);
        store(ra, cc, Ra);
        if (vb) store(rb, cc, Rb);
    }
}
// for_site_pairs_ld with compute(rr, cc) -> R doing both parts
template <class C, class S>
LRE_DEV void for_site_pairs(const Args &A, int r0, int r1, int c0, int c1, C compute, S store) {
    for_site_pairs_ld(A, r0, r1, c0, c1, [](int, int) { return 0; },
                      [&](int rr, int cc, int) { return compute(rr, cc); }, store);
}

// The sites of for_sites whose bit is set in the half-index mask `flags`, packed
// 64 to a wavefront; f(i) runs once per such site (i = rr * TS + cc).  The
// waves [wave0, wave0 + nwaves) walk the range 64 sites at a time, interleaved.
// A chunk's flagged sites move to the lanes behind those still pending from
// earlier chunks: lane l sends its site to lane (fill + rank) mod 64, flagged
// sites ranked among the flagged, the others after them, so the sends form a
// permutation of the 64 lanes.  Each full set of 64 runs at once and the rest
// at the end.  Flagged sites are sparse and clustered (the Nyquist flags), so
// this runs their long bodies on full waves instead of on every wave that
// holds one.  Called by whole waves, with all their lanes active.
template <class F>
LRE_DEV void for_flagged_sites(const Args &A, const unsigned *flags, int r0, int r1, int c0, int c1, int wave0,
                               int nwaves, F f) {
    const int ns = (c1 - c0 + 1) / 2, nr = r1 - r0;
    const int wave = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
    if (ns <= 0 || nr <= 0 || wave < wave0 || wave >= wave0 + nwaves) return;
    const int n = nr * ns, stride = 64 * nwaves;
    const unsigned long long below = (1ull << lane) - 1ull;
    int pend = 0, fill = 0;  // lanes [0, fill) of pend hold sites not yet run
    RowWalk w((wave - wave0) * 64 + lane, n, ns, stride);
    for (int base = (wave - wave0) * 64; base < n; base += stride, w.next()) {
        const int rr = r0 + w.q, cc = c0 + (int)(fct(A, rr, 0) & 1) + 2 * w.r;
        const int i = rr * TS + cc;
        const bool on = w.more() && cc < c1 && ((flags[i >> 6] >> ((i >> 1) & 31)) & 1u);
        const unsigned long long m = __ballot(on);
        if (m == 0ull) continue;
        const int cnt = __popcll(m);
        const int dst = on ? fill + __popcll(m & below) : fill + cnt + __popcll(~m & below);
        const int got = __builtin_amdgcn_ds_permute((dst & 63) << 2, i);
        if (fill + cnt < 64) {
            if (lane >= fill && lane < fill + cnt) pend = got;
            fill += cnt;
        } else {
            f(lane < fill ? pend : got);
            pend = got;
            fill += cnt - 64;
        }
    }
    if (lane < fill) f(pend);
}

// Lane exchange inside one wavefront: lane l receives lane l-1's value
// (wave_shr:1) or lane l+1's (wave_shl:1); the lane without a source gets 0.
LRE_DEV float from_lane_below(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x138, 0xF, 0xF, false));
}
LRE_DEV float from_lane_above(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x130, 0xF, 0xF, false));
}

// Row-sequential smoothing of a half-index weight plane held in LDS (hvwt
// :1446-1448, pmwt :1537-1539): for the sites of rows [b, rr1-b) and columns
// [b, cc1-b), in raster order, a weight is replaced by the mean of its four
// diagonal neighbours (upper-left, upper-right, lower-left, lower-right, summed
// in that order) when that mean lies further from 0.5; the upper neighbours
// are already replaced.
//
// In half-index columns h = cc >> 1 a site of row rr sits at 2h + parity(rr)
// and its diagonal neighbours at h-1, h (even parity) or h, h+1 (odd parity)
// of rows rr-1 and rr+1.  One wavefront runs the recurrence: lane l owns the
// half-columns h0 + 2l and h0 + 2l + 1 (h0 = b/2 - 2) and keeps the row above
// in registers, taking the neighbouring lanes' values by DPP; the lower row's
// original values are read from LDS one row ahead.  Every lane stores both of
// its cells back each row: a cell that is not a site of the range keeps its
// value, and lanes past the row end all address its last cell, which is never
// a site.  Sites of one row do not depend on each other.
//
// The other waves meanwhile run `others` (work that neither reads nor writes
// the half plane); the smoothing wave raises its issue priority so that its
// serial chain keeps its pace beside them.
template <class F>
LRE_DEV void smooth_rows(const Args &A, float *sm, int b, int rr1, int cc1, F others) {
    if (threadIdx.x < 64) {
        __builtin_amdgcn_s_setprio(2);
        const int h = b / 2 - 2 + 2 * (int)threadIdx.x, h1 = h + 1;
        const int last = rr1 - b;  // last row read (lower neighbours of the last row)
        const int om = min(h - 1, TSH - 1), o0 = min(h, TSH - 1), o1 = min(h1, TSH - 1), o2 = min(h + 2, TSH - 1);
        const int pe = (int)(fct(A, 0, 0) & 1), po = (int)(fct(A, 1, 0) & 1);
        const bool in0 = h >= b / 2, in1 = h1 >= b / 2;
        const bool ok0e = in0 && 2 * h + pe < cc1 - b, ok1e = in1 && 2 * h1 + pe < cc1 - b;
        const bool ok0o = in0 && 2 * h + po < cc1 - b, ok1o = in1 && 2 * h1 + po < cc1 - b;
        const float *row = sm + (b - 1) * TSH;
        float s0 = row[o0], s1 = row[o1];  // row above, final values
        row += TSH;
        float x0 = row[o0], x1 = row[o1];  // this row, original values
        row += TSH;
        float lm = row[om], l0 = row[o0], l1 = row[o1], l2 = row[o2];  // row below
        // one row; rows alternate parity, starting with an even row (b is even)
        auto step = [&](int rr, int par, bool ok0, bool ok1) {
            const float *nx = sm + min(rr + 2, last) * TSH;  // row after next
            const float nm = nx[om], n0 = nx[o0], n1 = nx[o1], n2 = nx[o2];
            const float mid = s0 + s1 + l0 + l1;
            float a0, a1;
            if (par == 0) {
                a0 = from_lane_below(s1) + s0 + lm + l0;
                a1 = mid;
            } else {
                a0 = mid;
                a1 = s1 + from_lane_above(s0) + l1 + l2;
            }
            a0 = 0.25f * a0;
            a1 = 0.25f * a1;
            const bool t0 = ok0 & (fabsf(0.5f - x0) < fabsf(0.5f - a0));
            const bool t1 = ok1 & (fabsf(0.5f - x1) < fabsf(0.5f - a1));
            const float y0 = t0 ? a0 : x0, y1 = t1 ? a1 : x1;
            float *cur = sm + rr * TSH;
            cur[o0] = y0;
            cur[o1] = y1;
            s0 = y0; s1 = y1;
            x0 = l0; x1 = l1;
            lm = nm; l0 = n0; l1 = n1; l2 = n2;
        };
        int rr = b;
        for (; rr + 1 < last; rr += 2) {
            step(rr, pe, ok0e, ok1e);
            step(rr + 1, po, ok0o, ok1o);
        }
        if (rr < last) step(rr, pe, ok0e, ok1e);
        __builtin_amdgcn_s_setprio(0);
    } else {
        others();
    }
    __syncthreads();
}

// LDS: the tile's CFA plane, and one half-index plane for the row-sequential
// smoothing passes
constexpr int LDS_CFA = N2, LDS_HALF = TS * TSH;
// and the two Nyquist flag planes as bit masks (in place of the arena's byte
// planes, which are not written)
constexpr int NYQ_WORDS = TS * TSH / 32;
constexpr size_t LDS_BYTES = sizeof(float) * (LDS_CFA + LDS_HALF) + 2 * sizeof(unsigned) * NYQ_WORDS;
static_assert((TS * TSH) % 32 == 0, "whole mask words");

__device__ void amaze_tile(const Args &A, float *ar, int top, int left) {
    constexpr int v1 = TS, v2 = 2 * TS, v3 = 3 * TS;
    constexpr int p1 = -TS + 1, p2 = -2 * TS + 2, p3 = -3 * TS + 3;
    constexpr int m1 = TS + 1, m2 = 2 * TS + 2, m3 = 3 * TS + 3;
    const int W = A.W, H = A.H;
    extern __shared__ __align__(16) float lds[];  // 16-byte aligned behind the static LDS
    // cfa lives in LDS.  The reference's tile fill also writes every value into
    // the green plane and past the end of its CFA array, into the Nyquist flags,
    // which stage 5 clears before use.  Here the green plane is written only at
    // the sites (by the green passes): at green sites it holds the CFA value,
    // which its readers take from LDS.  The writes past the end reach no
    // output and are dropped.
    float *cfa = lds, *green = ar + O_GREEN;
    float *sm = lds + LDS_CFA;
    unsigned *nqm = reinterpret_cast<unsigned *>(sm + LDS_HALF), *nqm2 = nqm + NYQ_WORDS;
    auto nbit = [](const unsigned *m, int k) { return (m[k >> 5] >> (k & 31)) & 1u; };
    float *wh = ar + O_WH, *dsq = ar + O_DSQ;
    float *vcd = ar + O_VCD, *hcd = ar + O_HCD, *vcdalt = ar + O_VCDALT, *hcdalt = ar + O_HCDALT;
    float *dgv = ar + O_DGV, *dgh = ar + O_DGH, *hvwt = ar + O_HVWT;
    float *nyqtest = ar + O_NYQTEST;

    const int bottom = min(top + TS, H + 16), right = min(left + TS, W + 16);
    const int rr1 = bottom - top, cc1 = right - left;
    const int rrmin = top < 0 ? 16 : 0, ccmin = left < 0 ? 16 : 0;
    const int rrmax = bottom > H ? H - top : rr1;
    const int ccmax = right > W ? W - left : cc1;
    const float scale = A.d_maxv ? 1.0f / *A.d_maxv : A.scale;

    // Clean-tile contract: every cell whose pre-tile value can reach the output
    // reads as zero.  Those are the CFA plane (the fill leaves the corners
    // unwritten) and the four colour-difference planes outside
    // [4, rr1-4) x [4, cc1-4), which stage 2 fills; the Nyquist flags (LDS
    // masks) are cleared here, where the reference clears them after stage 5:
    // nothing reads them before.  The other planes are written
    // before they are read, or are read only for cells that do not reach the
    // output (checked against zeroed scratch with poisoned scratch on the CPU).
    {
        float4 *c4 = reinterpret_cast<float4 *>(cfa);
        const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
        for (int t = threadIdx.x; t < N2 / 4; t += BLOCK) c4[t] = z;
        // Nyquist flags start at zero (:1432; nothing reads them before)
        for (int t = threadIdx.x; t < 2 * NYQ_WORDS; t += BLOCK) nqm[t] = 0u;
        float *cd[4] = {ar + O_VCD, ar + O_HCD, ar + O_VCDALT, ar + O_HCDALT};
        for (int t = threadIdx.x; t < N2; t += BLOCK) {
            const int rr = t / TS, cc = t - rr * TS;
            if (rr < 4 || rr >= rr1 - 4 || cc < 4 || cc >= cc1 - 4)
#pragma unroll
                for (int k = 0; k < 4; k++) cd[k][t] = 0.f;
        }
    }
    __syncthreads();

    // tile fill (:1256-1325), replaying the reference's writes by linear index:
    // the right border of a row starting past column 144 runs into the first
    // columns of the next row, and the lower border of a 160-row tile runs past
    // the end of the tile arrays; the right border is written last and wins.
    // The four border corners stay empty.
    {
        const int nc = ccmax - ccmin, nr = rrmax - rrmin;
        const int n_up = rrmin > 0 ? 16 * nc : 0;
        const int n_in = nr * nc;
        const int n_lo = rrmax < rr1 ? 16 * nc : 0;
        const int n_le = ccmin > 0 ? 16 * nr : 0;
        const int n_all = n_up + n_in + n_lo + n_le;
        // source pixel and arena position of fill write t (quotients by nc
        // through its float reciprocal: exact after one correction, t < 2^16)
        const float inv_nc = nc > 0 ? 1.0f / (float)nc : 0.f;
        auto qr = [&](int t, int &q, int &r) { lre::divmod(t, nc, inv_nc, q, r); };
        auto locate = [&](int t, int &row, int &col, int &L) {
            if (t < n_up) {
                int rr, cm;
                qr(t, rr, cm);
                const int cc = ccmin + cm;
                row = clampi(32 - rr + top, 0, H - 1); col = clampi(cc + left, 0, W - 1);
                L = rr * TS + cc;
            } else if (t < n_up + n_in) {
                int q, cm;
                qr(t - n_up, q, cm);
                const int rr = rrmin + q, cc = ccmin + cm;
                row = rr + top; col = cc + left;
                L = rr * TS + cc;
            } else if (t < n_up + n_in + n_lo) {
                int k, cm;
                qr(t - n_up - n_in, k, cm);
                const int cc = ccmin + cm;
                row = clampi(H - k - 2, 0, H - 1); col = clampi(left + cc, 0, W - 1);
                L = (rrmax + k) * TS + cc;
            } else {
                const int u = t - n_up - n_in - n_lo, rr = rrmin + u / 16, cc = u % 16;
                row = clampi(rr + top, 0, H - 1); col = clampi(32 - cc + left, 0, W - 1);
                L = rr * TS + cc;
            }
        };
        // FU loads in flight per thread before the stores
        constexpr int FU = 8;
        for (int base = threadIdx.x; base < n_all; base += BLOCK * FU) {
            float v[FU];
            int Ls[FU];
#pragma unroll
            for (int u = 0; u < FU; u++) {
                const int t = base + u * BLOCK;
                Ls[u] = -1;
                if (t < n_all) {
                    int row, col;
                    locate(t, row, col, Ls[u]);
                    v[u] = A.cfa_plane[(size_t)row * W + col];
                }
            }
#pragma unroll
            for (int u = 0; u < FU; u++) {
                const int L = Ls[u];
                if (L < 0) continue;
                if (L < N2) cfa[L] = v[u] * scale;
            }
        }
        __syncthreads();
        if (ccmax < cc1)
            for (int t = threadIdx.x; t < 16 * nr; t += BLOCK) {
                const int rr = rrmin + t / 16, k = t % 16;
                const int row = clampi(top + rr, 0, H - 1), col = clampi(W - k - 2, 0, W - 1);
                const float v = A.cfa_plane[(size_t)row * W + col] * scale;
                const int L = rr * TS + ccmax + k;
                if (L < N2) cfa[L] = v;
            }
    }
    __syncthreads();

    // stage 1 (:1331-1339).  Its planes are functions of the LDS CFA plane, so
    // every reader recomputes the values from LDS (bit-identical) and the stage
    // has no pass.  The R/B refinement reads dirwts1 after pmwt (stored over its
    // first half by stages 8 and 9) has replaced part of it: it takes those
    // cells from the arena and the others from LDS.  The dsq cells the
    // refinement reads as rbint are all written before.
    auto f_dh = [&](int i) { return fabsf(cfa[i + 1] - cfa[i - 1]); };
    auto f_dv = [&](int i) { return fabsf(cfa[i + v1] - cfa[i - v1]); };
    auto f_wv = [&](int i) { return EPS + fabsf(cfa[i + v2] - cfa[i]) + fabsf(cfa[i] - cfa[i - v2]) + f_dv(i); };
    auto f_wh = [&](int i) { return EPS + fabsf(cfa[i + 2] - cfa[i]) + fabsf(cfa[i] - cfa[i - 2]) + f_dh(i); };
    auto f_dsq = [&](int i) {
        const float dh = f_dh(i), dv = f_dv(i);
        return lre::sq(dh) + lre::sq(dv);
    };

    // stage 2 (:1342-1373)
    for_rect(4, rr1 - 4, 4, cc1 - 4, [&](int rr, int cc) {
        int i = rr * TS + cc;
        float sgn = (fct(A, rr, cc) & 1) ? -1.f : 1.f;
        float c0 = cfa[i];
        float ru = cfa[i - v1] * (f_wv(i - v2) + f_wv(i)) / (f_wv(i - v2) * (EPS + c0) + f_wv(i) * (EPS + cfa[i - v2]));
        float rd = cfa[i + v1] * (f_wv(i + v2) + f_wv(i)) / (f_wv(i + v2) * (EPS + c0) + f_wv(i) * (EPS + cfa[i + v2]));
        float rl = cfa[i - 1] * (f_wh(i - 2) + f_wh(i)) / (f_wh(i - 2) * (EPS + c0) + f_wh(i) * (EPS + cfa[i - 2]));
        float rr_ = cfa[i + 1] * (f_wh(i + 2) + f_wh(i)) / (f_wh(i + 2) * (EPS + c0) + f_wh(i) * (EPS + cfa[i + 2]));
        float gu_h = cfa[i - v1] + half(c0 - cfa[i - v2]);
        float gd_h = cfa[i + v1] + half(c0 - cfa[i + v2]);
        float gl_h = cfa[i - 1] + half(c0 - cfa[i - 2]);
        float gr_h = cfa[i + 1] + half(c0 - cfa[i + 2]);
        float gu_a = fabsf(1.f - ru) < ARTHRESH ? c0 * ru : gu_h;
        float gd_a = fabsf(1.f - rd) < ARTHRESH ? c0 * rd : gd_h;
        float gl_a = fabsf(1.f - rl) < ARTHRESH ? c0 * rl : gl_h;
        float gr_a = fabsf(1.f - rr_) < ARTHRESH ? c0 * rr_ : gr_h;
        float hw = f_wh(i - 1) / (f_wh(i - 1) + f_wh(i + 1));
        float vw = f_wv(i - v1) / (f_wv(i + v1) + f_wv(i - v1));
        float gv_h = vw * gd_h + (1.f - vw) * gu_h;
        float gh_h = hw * gr_h + (1.f - hw) * gl_h;
        float va = sgn * (gv_h - c0), ha = sgn * (gh_h - c0);
        vcdalt[i] = va;
        hcdalt[i] = ha;
        if (c0 > A.clip_pt8 || gv_h > A.clip_pt8 || gh_h > A.clip_pt8) {
            vcd[i] = va;
            hcd[i] = ha;
            gu_a = gu_h; gd_a = gd_h; gl_a = gl_h; gr_a = gr_h;
        } else {
            vcd[i] = sgn * (vw * gd_a + (1.f - vw) * gu_a - c0);
            hcd[i] = sgn * (hw * gr_a + (1.f - hw) * gl_a - c0);
        }
        dgv[i] = lre::min_std(lre::sq(gu_h - gd_h), lre::sq(gu_a - gd_a));
        dgh[i] = lre::min_std(lre::sq(gl_h - gr_h), lre::sq(gl_a - gr_a));
    });
    __syncthreads();

    // stage 3 (:1376-1397): hcd scans along rows, vcd scans down columns.  Each
    // scan reads its predecessor two sites back after that one was updated; the
    // values ahead of the scan are not yet updated, so they are fetched in
    // chunks of SK sites before the chunk is processed.
    // The waves without a scan meanwhile take the Gaussian-gradient sum of the
    // Nyquist test (:1403-1409, the subtrahend), a function of the LDS CFA
    // plane only, and store it in the nyqtest cells; stage 4-5 subtracts it.
    auto ggrad_sum = [&](int i) {
        return GGRAD[0] * f_dsq(i) + GGRAD[1] * (f_dsq(i - v1) + f_dsq(i + 1) + f_dsq(i - 1) + f_dsq(i + v1)) +
               GGRAD[2] * (f_dsq(i - m1) + f_dsq(i + p1) + f_dsq(i - p1) + f_dsq(i + m1)) +
               GGRAD[3] * (f_dsq(i - v2) + f_dsq(i - 2) + f_dsq(i + 2) + f_dsq(i + v2)) +
               GGRAD[4] * (f_dsq(i - v2 - 1) + f_dsq(i - v2 + 1) + f_dsq(i - TS - 2) + f_dsq(i - TS + 2) +
                           f_dsq(i + TS - 2) + f_dsq(i + TS + 2) + f_dsq(i + v2 - 1) + f_dsq(i + v2 + 1)) +
               GGRAD[5] * (f_dsq(i - m2) + f_dsq(i + p2) + f_dsq(i - p2) + f_dsq(i + m2));
    };
    {
        constexpr int SK = 4;
        const int nh = 2 * max(rr1 - 8, 0), nv = 2 * max(cc1 - 8, 0);
        const int nscan = ((nh + nv + 63) / 64) * 64;  // whole waves scan (at most 640 threads)
        static_assert(2 * 2 * (TS - 8) <= BLOCK - 5 * 64, "scans leave waves for the gradient sums");
        if ((int)threadIdx.x >= nscan)
            for_sites_on(A, 6, rr1 - 6, 6, cc1 - 6, nscan, BLOCK - nscan, [&](int rr, int cc) {
                const int i = rr * TS + cc;
                nyqtest[i >> 1] = ggrad_sum(i);
            });
        for (int t = threadIdx.x; t < nh + nv; t += BLOCK) {
            const bool horiz = t < nh;
            int base, first, end, step;  // positions base + k * step, k in [first, end)
            if (horiz) {
                const int rr = 4 + t / 2, par = t & 1;
                base = rr * TS; step = 1; first = 4 + par; end = cc1 - 4;
            } else {
                const int u = t - nh, cc = 4 + u / 2, par = u & 1;
                base = cc; step = TS; first = 4 + par; end = rr1 - 4;
            }
            float *cd = horiz ? hcd : vcd;
            const float *alt = horiz ? hcdalt : vcdalt;
            const int s1 = step, s2 = 2 * step;
            float prev = cd[base + (first - 2) * step];
            float a0 = alt[base + (first - 2) * step];
            float cur = 0.f, a1 = 0.f;
            if (first < end) {
                cur = cd[base + first * step];
                a1 = alt[base + first * step];
            }
            for (int p = first; p < end; p += 2 * SK) {
                float nx[SK], al[SK];
#pragma unroll
                for (int u = 0; u < SK; u++) {
                    const int i = base + (p + 2 * u) * step + s2;
                    nx[u] = cd[i];
                    al[u] = alt[i];
                }
#pragma unroll
                for (int u = 0; u < SK; u++) {
                    const int pos = p + 2 * u;
                    if (pos >= end) break;
                    const int i = base + pos * step;
                    const float c0 = cfa[i], nxt = nx[u], a2 = al[u];
                    float hv = 3.f * (lre::sq(prev) + lre::sq(cur) + lre::sq(nxt)) - lre::sq(prev + cur + nxt);
                    float hav = 3.f * (lre::sq(a0) + lre::sq(a1) + lre::sq(a2)) - lre::sq(a0 + a1 + a2);
                    float h = hav < hv ? a1 : cur;
                    const int rr = horiz ? base / TS : pos, cc = horiz ? pos : base;
                    const float lf = cfa[i - s1], rt = cfa[i + s1];
                    // the green and the R/B forms of the reference differ only in
                    // signs: with sg = -1 at greens and +1 elsewhere each operation
                    // below is the reference's own (-x is exact, and a + (-b) and
                    // a - b round alike), so one body serves both colours
                    const float sg = (fct(A, rr, cc) & 1) ? -1.f : 1.f;
                    const float gh = sg * h + c0;  // -h + c0 / h + c0
                    const float mg = lre::med3(gh, lf, rt);
                    const float hm = sg * mg - sg * c0;  // -mg + c0 / mg - c0
                    if (sg * h < 0.f) {  // h > 0 / h < 0
                        if (sg * (3.f * h) < -(gh + c0)) h = hm;  // 3h > gh + c0 / 3h < -(gh + c0)
                        else {
                            const float w2 = 1.f + sg * (3.f * h / (EPS + gh + c0));  // 1 - q / 1 + q
                            h = w2 * h + (1.f - w2) * hm;
                        }
                    }
                    if (gh > A.clip_pt) h = sg * (mg - c0);
                    cd[i] = h;
                    prev = h;
                    a0 = a1; a1 = a2; cur = nxt;
                }
            }
        }
    }
    __syncthreads();
    // cddiffsq; its only reader, the Nyquist test, reads it at sites, so the
    // sites' values go to the LDS half plane (free until stage 4-5 fills it with
    // hvwt) and the arena plane is not written
    for_sites(A, 4, rr1 - 4, 4, cc1 - 4, [&](int rr, int cc) {
        int i = rr * TS + cc;
        sm[i >> 1] = lre::sq(vcd[i] - hcd[i]);
    });
    __syncthreads();

    // stages 4-5 (:1400-1428); the gradient sums came with stage 3.  A site's
    // Nyquist flag (:1432-1434) is set as soon as its test value is known
    int local_nyq = 0;
    for_sites_split(A, 6, rr1 - 6, 6, cc1 - 6, [&](int rr, int cc) {
        const int i = rr * TS + cc;
        const float *d = sm;  // cddiffsq at sites, half-index
        return
            (GODD[0] * d[(i) >> 1] + GODD[1] * (d[(i - m1) >> 1] + d[(i + p1) >> 1] + d[(i - p1) >> 1] + d[(i + m1) >> 1]) +
             GODD[2] * (d[(i - v2) >> 1] + d[(i - 2) >> 1] + d[(i + 2) >> 1] + d[(i + v2) >> 1]) +
             GODD[3] * (d[(i - m2) >> 1] + d[(i + p2) >> 1] + d[(i - p2) >> 1] + d[(i + m2) >> 1])) -
            nyqtest[i >> 1];
    }, [&](int rr, int cc, float nt) {
        if (nt > 0.f) {
            const int k = (rr * TS + cc) >> 1;
            atomicOr(&nqm[k >> 5], 1u << (k & 31));
            local_nyq = 1;
        }
    });
    // (also: the hvwt pass below overwrites the LDS half plane)
    const int any_nyq = __syncthreads_or(local_nyq);
    for_site_pairs(A, 6, rr1 - 6, 6, cc1 - 6, [&](int rr, int cc) {
        const int i = rr * TS + cc;
        float uave = vcd[i] + vcd[i - v1] + vcd[i - v2] + vcd[i - v3];
        float dave = vcd[i] + vcd[i + v1] + vcd[i + v2] + vcd[i + v3];
        float lave = hcd[i] + hcd[i - 1] + hcd[i - 2] + hcd[i - 3];
        float rave = hcd[i] + hcd[i + 1] + hcd[i + 2] + hcd[i + 3];
        using lre::sq;
        float vu = sq(vcd[i] - uave) + sq(vcd[i - v1] - uave) + sq(vcd[i - v2] - uave) + sq(vcd[i - v3] - uave);
        float vd = sq(vcd[i] - dave) + sq(vcd[i + v1] - dave) + sq(vcd[i + v2] - dave) + sq(vcd[i + v3] - dave);
        float hl = sq(hcd[i] - lave) + sq(hcd[i - 1] - lave) + sq(hcd[i - 2] - lave) + sq(hcd[i - 3] - lave);
        float hr = sq(hcd[i] - rave) + sq(hcd[i + 1] - rave) + sq(hcd[i + 2] - rave) + sq(hcd[i + 3] - rave);
        float hw = f_wh(i - 1) / (f_wh(i - 1) + f_wh(i + 1));
        float vw = f_wv(i - v1) / (f_wv(i + v1) + f_wv(i - v1));
        float vvar = EPSSQ + vw * vd + (1.f - vw) * vu;
        float hvar = EPSSQ + hw * hr + (1.f - hw) * hl;
        float vu2 = dgv[i] + dgv[i - v1] + dgv[i - v2];
        float vd2 = dgv[i] + dgv[i + v1] + dgv[i + v2];
        float hl2 = dgh[i] + dgh[i - 1] + dgh[i - 2];
        float hr2 = dgh[i] + dgh[i + 1] + dgh[i + 2];
        float vvar1 = EPSSQ + vw * vd2 + (1.f - vw) * vu2;
        float hvar1 = EPSSQ + hw * hr2 + (1.f - hw) * hl2;
        float varwt = hvar / (vvar + hvar);
        float diffwt = hvar1 / (vvar1 + hvar1);
        return ((0.5f - varwt) * (0.5f - diffwt) > 0.f && fabsf(0.5f - diffwt) < fabsf(0.5f - varwt)) ? varwt : diffwt;
    }, [&](int rr, int cc, float w) {
        // the smoothing reads the LDS copy; the green pass stores the smoothed
        // weight of every site that is read later ([8, rr1-8)), so the arena
        // cell is not written here
        sm[(rr * TS + cc) >> 1] = w;
    });
    __syncthreads();

    // stage 6 (:1436-1481)
    if (any_nyq) {
        for_sites(A, 8, rr1 - 8, 8, cc1 - 8, [&](int rr, int cc) {
            int i = rr * TS + cc;
            const unsigned s = nbit(nqm, (i - v2) >> 1) + nbit(nqm, (i - m1) >> 1) + nbit(nqm, (i + p1) >> 1) +
                               nbit(nqm, (i - 2) >> 1) + nbit(nqm, (i + 2) >> 1) + nbit(nqm, (i - p1) >> 1) +
                               nbit(nqm, (i + m1) >> 1) + nbit(nqm, (i + v2) >> 1);
            const unsigned f = s > 4 ? 1u : (s < 4 ? 0u : nbit(nqm, i >> 1));
            const int k = i >> 1;
            if (f) atomicOr(&nqm2[k >> 5], 1u << (k & 31));
        });
        __syncthreads();
        for_flagged_sites(A, nqm2, 8, rr1 - 8, 8, cc1 - 8, 0, BLOCK / 64, [&](int i) {
            float scfa = 0, sh = 0, sv = 0, sqh = 0, sqv = 0, area = 0;
            for (int a = -6; a < 7; a += 2)
                for (int b = -6; b < 7; b += 2) {
                    int j = i + a * TS + b;
                    if (j >= 0 && j < N2 && nbit(nqm2, j >> 1)) {
                        float t = cfa[j];
                        scfa += t;
                        sh += (cfa[j - 1] + cfa[j + 1]);
                        sv += (cfa[j - v1] + cfa[j + v1]);
                        sqh += lre::sq(t - cfa[j - 1]) + lre::sq(t - cfa[j + 1]);
                        sqv += lre::sq(t - cfa[j - v1]) + lre::sq(t - cfa[j + v1]);
                        area += 1;
                    }
                }
            sh = scfa - half(sh);
            sv = scfa - half(sv);
            area = half(area);
            float hv2 = EPSSQ + fabsf(area * sqh - sh * sh), vv2 = EPSSQ + fabsf(area * sqv - sv * sv);
            sm[i >> 1] = hv2 / (vv2 + hv2);  // the arena cell: as in the hvwt pass
        });
        __syncthreads();
    }
    float *dg0 = vcdalt;        // Dgrb0 over vcdalt (half-index)
    float *d2 = ar + O_Q1M;     // Dgrb2 {h,v} pairs over Dgrbsq1m/Dgrbsq1p
    // stage 7 (:1484-1502) computes the diagonal gradients and q1 sums that
    // stage 8 reads; stage 8 recomputes them from LDS, and no later stage reads
    // those cells, so the stage has no pass of its own.

    // stage 8 (:1505-1532).  The stage-7 cells it reads are functions of the
    // LDS CFA plane and lie inside stage 7's range: the diagonal gradients of a
    // site s, and the q1 sums at a green n (the diagonal neighbours of a site
    // are sites, its axial and knight neighbours greens).  They are computed
    // here (bit-identical to the stored cells).
    auto f_delp = [&](int s) { return fabsf(cfa[s + p1] - cfa[s - p1]); };
    auto f_delm = [&](int s) { return fabsf(cfa[s + m1] - cfa[s - m1]); };
    auto f_q1p = [&](int n) { return lre::sq(cfa[n] - cfa[n - p1]) + lre::sq(cfa[n] - cfa[n + p1]); };
    auto f_q1m = [&](int n) { return lre::sq(cfa[n] - cfa[n - m1]) + lre::sq(cfa[n] - cfa[n + m1]); };
    // rbp over hcdalt and pmwt over dirwts1, as in the reference.  rbm lives in
    // the second half of the delhvsqsum plane (rbint fills only the first)
    // instead of over vcd: after this stage only the rbint pass reads it,
    // while the green passes, which here run after stage 8, read vcd.
    float *rbm = dsq + N2 / 2, *rbp = hcdalt, *pmwt = wh;
    auto stage8 = [&](int rr, int cc) {
        const int i = rr * TS + cc, k = i >> 1;
        float c0 = cfa[i];
        float rse = twice(cfa[i + m1]) / (EPS + c0 + cfa[i + m2]);
        float rnw = twice(cfa[i - m1]) / (EPS + c0 + cfa[i - m2]);
        float rne = twice(cfa[i + p1]) / (EPS + c0 + cfa[i + p2]);
        float rsw = twice(cfa[i - p1]) / (EPS + c0 + cfa[i - p2]);
        float bse = fabsf(1.f - rse) < ARTHRESH ? c0 * rse : (cfa[i + m1]) + half(c0 - cfa[i + m2]);
        float bnw = fabsf(1.f - rnw) < ARTHRESH ? c0 * rnw : (cfa[i - m1]) + half(c0 - cfa[i - m2]);
        float bne = fabsf(1.f - rne) < ARTHRESH ? c0 * rne : (cfa[i + p1]) + half(c0 - cfa[i + p2]);
        float bsw = fabsf(1.f - rsw) < ARTHRESH ? c0 * rsw : (cfa[i - p1]) + half(c0 - cfa[i - p2]);
        float wse = EPS + f_delm(i) + f_delm(i + m1) + f_delm(i + m2);
        float wnw = EPS + f_delm(i) + f_delm(i - m1) + f_delm(i - m2);
        float wne = EPS + f_delp(i) + f_delp(i + p1) + f_delp(i + p2);
        float wsw = EPS + f_delp(i) + f_delp(i - p1) + f_delp(i - p2);
        float m = (wse * bnw + wnw * bse) / (wse + wnw);
        float p = (wne * bsw + wsw * bne) / (wne + wsw);
        const float cp = cfa[i - p1], cp2 = cfa[i + p1], cm = cfa[i - m1], cm2 = cfa[i + m1];
        if (p < c0) {
            if (twice(p) < c0) p = lre::med3(p, cp, cp2);
            else {
                float w = twice(c0 - p) / (EPS + p + c0);
                p = w * p + (1.f - w) * lre::med3(p, cp, cp2);
            }
        }
        if (m < c0) {
            if (twice(m) < c0) m = lre::med3(m, cm, cm2);
            else {
                float w = twice(c0 - m) / (EPS + m + c0);
                m = w * m + (1.f - w) * lre::med3(m, cm, cm2);
            }
        }
        if (p > A.clip_pt) p = lre::med3(p, cp, cp2);
        if (m > A.clip_pt) m = lre::med3(m, cm, cm2);
        rbp[k] = p;
        rbm[k] = m;
#define QS(f, o) f(i + (o))
        float varm = EPSSQ + (GEVEN[0] * (QS(f_q1m, -v1) + QS(f_q1m, -1) + QS(f_q1m, 1) + QS(f_q1m, v1)) +
                              GEVEN[1] * (QS(f_q1m, -v2 - 1) + QS(f_q1m, -v2 + 1) + QS(f_q1m, -2 - v1) + QS(f_q1m, 2 - v1) +
                                          QS(f_q1m, -2 + v1) + QS(f_q1m, 2 + v1) + QS(f_q1m, v2 - 1) + QS(f_q1m, v2 + 1)));
        float varp = EPSSQ + (GEVEN[0] * (QS(f_q1p, -v1) + QS(f_q1p, -1) + QS(f_q1p, 1) + QS(f_q1p, v1)) +
                              GEVEN[1] * (QS(f_q1p, -v2 - 1) + QS(f_q1p, -v2 + 1) + QS(f_q1p, -2 - v1) + QS(f_q1p, 2 - v1) +
                                          QS(f_q1p, -2 + v1) + QS(f_q1p, 2 + v1) + QS(f_q1p, v2 - 1) + QS(f_q1p, v2 + 1)));
#undef QS
        const float w = varm / (varp + varm);
        pmwt[k] = w;
    };
    // green refinement at Nyquist sites (:1449-1481)
    auto nyq_refine = [&](int i) {
        const int k = i >> 1;
#define D2H(o) d2[2 * ((i + (o)) >> 1)]
#define D2V(o) d2[2 * ((i + (o)) >> 1) + 1]
        float gvh = EPSSQ + (GQUINC[0] * D2H(0) + GQUINC[1] * (D2H(-m1) + D2H(p1) + D2H(-p1) + D2H(m1)) +
                             GQUINC[2] * (D2H(-v2) + D2H(-2) + D2H(2) + D2H(v2)) +
                             GQUINC[3] * (D2H(-m2) + D2H(p2) + D2H(-p2) + D2H(m2)));
        float gvv = EPSSQ + (GQUINC[0] * D2V(0) + GQUINC[1] * (D2V(-m1) + D2V(p1) + D2V(-p1) + D2V(m1)) +
                             GQUINC[2] * (D2V(-v2) + D2V(-2) + D2V(2) + D2V(v2)) +
                             GQUINC[3] * (D2V(-m2) + D2V(p2) + D2V(-p2) + D2V(m2)));
#undef D2H
#undef D2V
        float d0 = (hcd[i] * gvv + vcd[i] * gvh) / (gvv + gvh);
        dg0[k] = d0;
        green[i] = cfa[i] + d0;
    };
    // hvwt smoothing (:1446-1448) on the LDS copy that stages 4-6 wrote beside
    // hvwt, while the other waves run stage 8 (the LDS CFA plane in, planes
    // that nothing reads before the R/B interpolation out); each smoothed
    // weight is stored back by the green pass, which also moves pmwt into the
    // LDS plane
    smooth_rows(A, sm, 8, rr1, cc1, [&] { for_sites_on(A, 8, rr1 - 8, 8, cc1 - 8, 64, BLOCK - 64, stage8); });
    struct SG {
        float hk, d0, g, dh, dv, pw;
    };
    for_site_pairs(A, 8, rr1 - 8, 8, cc1 - 8, [&](int rr, int cc) {
        const int i = rr * TS + cc, k = i >> 1;
        SG R;
        R.hk = sm[k];
        R.pw = pmwt[k];
        R.d0 = R.hk * vcd[i] + (1.f - R.hk) * hcd[i];
        R.g = cfa[i] + R.d0;
        const bool q = nbit(nqm2, k) != 0;
        // green at green sites is the CFA value (the reference's fill puts it
        // there and no pass writes those cells), so it is read from LDS
        R.dh = q ? lre::sq(R.g - half(cfa[i - 1] + cfa[i + 1])) : 0.f;
        R.dv = q ? lre::sq(R.g - half(cfa[i - v1] + cfa[i + v1])) : 0.f;
        return R;
    }, [&](int rr, int cc, const SG &R) {
        const int i = rr * TS + cc, k = i >> 1;
        hvwt[k] = R.hk;
        sm[k] = R.pw;  // the LDS plane now serves the pmwt smoothing
        dg0[k] = R.d0;
        green[i] = R.g;
        if (any_nyq) {  // Dgrb2 has no reader but the Nyquist refinement
            d2[2 * k] = R.dh;
            d2[2 * k + 1] = R.dv;
        }
    });
    __syncthreads();
    // stage 9 (:1535-1540): pmwt smoothing reads updated upper-diagonal neighbours
    float *rbint = dsq;
    // same recurrence for pmwt; rbint of a site needs only that site's final
    // pmwt, so it is computed in a parallel pass afterwards
    smooth_rows(A, sm, 10, rr1, cc1, [&] {
        // the Nyquist refinement reads and writes neither the LDS plane nor the stage-8 planes
        if (any_nyq) for_flagged_sites(A, nqm2, 8, rr1 - 8, 8, cc1 - 8, 1, BLOCK / 64 - 1, nyq_refine);
    });
    for_sites(A, 10, rr1 - 10, 10, cc1 - 10, [&](int rr, int cc) {
        const int i = rr * TS + cc, k = i >> 1;
        const float pk = sm[k];
        pmwt[k] = pk;
        rbint[k] = half(cfa[i] + rbm[k] * (1.f - pk) + rbp[k] * pk);
    });
    __syncthreads();
    // (:1541-1566)
    // dirwts1 of green cell j as the refinement reads it: where the cell lies in
    // the plane's first half and its pair's site in stage 8's range, pmwt has
    // replaced it (arena); elsewhere it holds the stage-1 value
    // (the arena cell is loaded whether or not it is taken, so that the load
    // does not wait behind a branch)
    auto wh_at = [&](int j) {
        const float a = wh[j];
        bool arena = false;
        if (j < N2 / 2) {
            const int e = 2 * j, er = e / TS, ec = e - er * TS;
            const int sc = (fct(A, er, ec) & 1) ? ec + 1 : ec;  // the site of the pair (same row)
            arena = er >= 8 && er < rr1 - 8 && sc >= 8 && sc < cc1 - 8;
        }
        return arena ? a : f_wh(j);
    };
    struct SR {
        bool on;
        float g, d0;
    };
    // the arena values of a site, loaded ahead of the gate (sm holds the
    // smoothed pmwt)
    struct RL {
        float pk, hk, rb, ru, rd, rl, rr, wl, wr;
    };
    for_site_pairs_ld(A, 12, rr1 - 12, 12, cc1 - 12, [&](int rr, int cc) {
        const int i = rr * TS + cc, k = i >> 1;
        return RL{sm[k], hvwt[k], rbint[k], rbint[k - v1], rbint[k + v1], rbint[k - 1], rbint[k + 1],
                  wh_at(i - 1), wh_at(i + 1)};
    }, [&](int rr, int cc, const RL &L) {
        const int i = rr * TS + cc;
        SR R{false, 0.f, 0.f};
        if (fabsf(0.5f - L.pk) >= fabsf(0.5f - L.hk)) return R;
        const float rb = L.rb;
        float ru = twice(cfa[i - v1]) / (EPS + rb + L.ru);
        float rd = twice(cfa[i + v1]) / (EPS + rb + L.rd);
        float rl = twice(cfa[i - 1]) / (EPS + rb + L.rl);
        float rr_ = twice(cfa[i + 1]) / (EPS + rb + L.rr);
        float gu = fabsf(1.f - ru) < ARTHRESH ? rb * ru : cfa[i - v1] + half(rb - L.ru);
        float gd = fabsf(1.f - rd) < ARTHRESH ? rb * rd : cfa[i + v1] + half(rb - L.rd);
        float gl = fabsf(1.f - rl) < ARTHRESH ? rb * rl : cfa[i - 1] + half(rb - L.rl);
        float gr = fabsf(1.f - rr_) < ARTHRESH ? rb * rr_ : cfa[i + 1] + half(rb - L.rr);
        float gv = (f_wv(i - v1) * gd + f_wv(i + v1) * gu) / (f_wv(i + v1) + f_wv(i - v1));
        const float wl = L.wl, wr = L.wr;
        float gh = (wl * gr + wr * gl) / (wl + wr);
        const float up = cfa[i - v1], dn = cfa[i + v1], lf = cfa[i - 1], rt = cfa[i + 1];
        if (gv < rb) {
            if (twice(gv) < rb) gv = lre::med3(gv, up, dn);
            else {
                float w = twice(rb - gv) / (EPS + gv + rb);
                gv = w * gv + (1.f - w) * lre::med3(gv, up, dn);
            }
        }
        if (gh < rb) {
            if (twice(gh) < rb) gh = lre::med3(gh, lf, rt);
            else {
                float w = twice(rb - gh) / (EPS + gh + rb);
                gh = w * gh + (1.f - w) * lre::med3(gh, lf, rt);
            }
        }
        if (gh > A.clip_pt) gh = lre::med3(gh, lf, rt);
        if (gv > A.clip_pt) gv = lre::med3(gv, up, dn);
        const float hk = L.hk;
        float g = gh * (1.f - hk) + gv * hk;
        R.on = true;
        R.g = g;
        R.d0 = g - cfa[i];
        return R;
    }, [&](int rr, int cc, const SR &R) {
        if (!R.on) return;
        const int i = rr * TS + cc;
        green[i] = R.g;
        dg0[i >> 1] = R.d0;
    });
    __syncthreads();

    // second chroma difference (:1571-1595).  Rows r0 + 2m (the B rows) move
    // their Dgrb0 into Dgrb1 and clear it; call the values before the move
    // (Dgrb0 of every site) X.  X goes to the LDS half plane for rows
    // 11 .. rr1-12, which covers every read below: the pass reads only X, and
    // stage 10 reads X there and the pass's output from the arena.  Every cell
    // stage 10 reads in the arena is written by the pass, so the move itself
    // (Dgrb1 = X, Dgrb0 = 0 on B rows) leaves nothing that is read and is not
    // stored.
    float *dg1 = hcd;  // Dgrb1 over hcd (half-index)
    const int r0b = 13 - A.ey;
    {
        const int n = max(rr1 - 22, 0) * TSH;
        constexpr int U = 2;
        for (int t0 = threadIdx.x; t0 < n; t0 += U * BLOCK) {
            float v[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int t = t0 + u * BLOCK;
                v[u] = t < n ? dg0[11 * TSH + t] : 0.f;
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int t = t0 + u * BLOCK;
                if (t < n) sm[11 * TSH + t] = v[u];
            }
        }
    }
    __syncthreads();
    for_sites(A, 14, rr1 - 14, 14, cc1 - 14, [&](int rr, int cc) {
        int i = rr * TS + cc;
        float *D = (1 - (int)fct(A, rr, cc) / 2) ? dg1 : dg0;
#define DD(o) sm[(i + (o)) >> 1]
        float wnw = 1.f / (EPS + fabsf(DD(-m1) - DD(m1)) + fabsf(DD(-m1) - DD(-m3)) + fabsf(DD(m1) - DD(-m3)));
        float wne = 1.f / (EPS + fabsf(DD(p1) - DD(-p1)) + fabsf(DD(p1) - DD(p3)) + fabsf(DD(-p1) - DD(p3)));
        float wsw = 1.f / (EPS + fabsf(DD(-p1) - DD(p1)) + fabsf(DD(-p1) - DD(m3)) + fabsf(DD(p1) - DD(-p3)));
        float wse = 1.f / (EPS + fabsf(DD(m1) - DD(-m1)) + fabsf(DD(m1) - DD(-p3)) + fabsf(DD(-m1) - DD(m3)));
        D[i >> 1] = (wnw * (1.325f * DD(-m1) - 0.175f * DD(-m3) - 0.075f * (DD(-m1 - 2) + DD(-m1 - v2))) +
                     wne * (1.325f * DD(p1) - 0.175f * DD(p3) - 0.075f * (DD(p1 + 2) + DD(p1 + v2))) +
                     wsw * (1.325f * DD(-p1) - 0.175f * DD(-p3) - 0.075f * (DD(-p1 - 2) + DD(-p1 - v2))) +
                     wse * (1.325f * DD(m1) - 0.175f * DD(m3) - 0.075f * (DD(m1 + 2) + DD(m1 + v2)))) /
                    (wnw + wne + wsw + wse);
#undef DD
    });
    __syncthreads();

    // stage 10 (:1597-1625), two horizontally adjacent pixels per thread.  In a
    // Bayer row one of the two is green and the other is not, so each thread
    // runs each of the two interpolations once, where a wavefront of single
    // pixels would run both for every pixel; other pairs (the last column of
    // an odd range) take each pixel on its own.
    struct SO {
        float r, g, b;
    };
    // Dgrb0 / Dgrb1 of a site cell: on B rows (rr - r0b even) X is Dgrb1 and the
    // pass's output Dgrb0, on R rows the other way round.  X is in the LDS half
    // plane, the output in the arena plane g_row(rr); every value is loaded
    // unconditionally and picked by the row parity, so that the loads of a pixel
    // are all in flight together.
    auto even_row = [&](int rr) { return ((rr - r0b) & 1) == 0; };
    auto g_row = [&](int rr) { return even_row(rr) ? dg0 : dg1; };
    auto at_green = [&](int rr, int t) {
        const float g = cfa[t];
        const int ku = (t - v1) >> 1, kr = (t + 1) >> 1, kl = (t - 1) >> 1, kd = (t + v1) >> 1;
        const float hu = hvwt[ku], hr = hvwt[kr], hl = hvwt[kl], hd = hvwt[kd];
        const float *gv = g_row(rr + 1), *gh = g_row(rr);  // rows rr - 1 and rr + 1 share a parity
        const float au = gv[ku], ad = gv[kd], ar_ = gh[kr], al = gh[kl];
        const float xu = sm[ku], xd = sm[kd], xr = sm[kr], xl = sm[kl];
        const bool e = even_row(rr);
        // d0 / d1 of the four neighbours: on an even row the vertical ones' d0
        // is X and the horizontal ones' d0 the arena value; the reverse for d1
        const float u0 = e ? xu : au, d0_ = e ? xd : ad, r0 = e ? ar_ : xr, l0 = e ? al : xl;
        const float u1 = e ? au : xu, d1_ = e ? ad : xd, r1 = e ? xr : ar_, l1 = e ? xl : al;
        float inv = 1.0f / (hu + 2.f - hr - hl + hd);
        float rd = (hu * u0 + (1.f - hr) * r0 + (1.f - hl) * l0 + hd * d0_) * inv;
        float bd = (hu * u1 + (1.f - hr) * r1 + (1.f - hl) * l1 + hd * d1_) * inv;
        return SO{g - rd, g, g - bd};
    };
    auto at_site = [&](int rr, int t) {
        const float g = green[t];
        const int k = t >> 1;
        const float a = g_row(rr)[k], x = sm[k];
        const bool e = even_row(rr);
        return SO{g - (e ? a : x), g, g - (e ? x : a)};
    };
    auto store = [&](int rr, int cc, const SO &R) {
        const int row = rr + top, col = cc + left;
        if (row < 0 || row >= H || col < 0 || col >= W) return;
        float *o = A.rgb + ((size_t)row * W + col) * 3;
        o[0] = (0.f < R.r) ? R.r : 0.f;
        o[1] = (0.f < R.g) ? R.g : 0.f;
        o[2] = (0.f < R.b) ? R.b : 0.f;
    };
    {
        const int c1 = cc1 - 16, nr = rr1 - 32, np = (c1 - 16 + 1) / 2;
        if (nr > 0 && np > 0)
            for (RowWalk w((int)threadIdx.x, nr * np, np, BLOCK); w.more(); w.next()) {
                const int rr = 16 + w.q, cc = 16 + 2 * w.r, t = rr * TS + cc;
                const bool two = cc + 1 < c1;
                const bool ga = fct(A, rr + top, cc + left) == 1, gb = fct(A, rr + top, cc + 1 + left) == 1;
                if (two && ga != gb) {
                    const SO G = at_green(rr, ga ? t : t + 1), S = at_site(rr, ga ? t + 1 : t);
                    store(rr, cc, ga ? G : S);
                    store(rr, cc + 1, ga ? S : G);
                } else {
                    store(rr, cc, ga ? at_green(rr, t) : at_site(rr, t));
                    if (two) store(rr, cc + 1, gb ? at_green(rr, t + 1) : at_site(rr, t + 1));
                }
            }
    }
    __syncthreads();
}

// The tile queue's counter sits behind the arenas and the spare one.
LRE_DEV unsigned *queue_of(float *arenas, int slots) {
    return reinterpret_cast<unsigned *>(arenas + (size_t)(slots + 1) * ARENA_FLOATS);
}

__global__ void queue_reset(float *arenas, int slots) {
    if (threadIdx.x == 0) *queue_of(arenas, slots) = 0u;
}

__global__ __launch_bounds__(BLOCK) void amaze_kernel(Args A) {
    float *ar = A.arenas + (size_t)blockIdx.x * ARENA_FLOATS;
    lre::tile_queue(queue_of(A.arenas, gridDim.x), A.ntiles, [&](int t) {
        const int tr = t / A.ntc, tc = t % A.ntc;
        amaze_tile(A, ar, -16 + tr * (TS - 32), -16 + tc * (TS - 32));
    });
}

// One tile queue over several images: tile t belongs to the last image whose
// tile_base is <= t.
__global__ __launch_bounds__(BLOCK) void amaze_multi_kernel(const Args *__restrict__ tab, int nimg,
                                                            int total_tiles, int slots, float *arenas) {
    float *ar = arenas + (size_t)blockIdx.x * ARENA_FLOATS;
    lre::tile_queue(queue_of(arenas, slots), total_tiles, [&](int t) {
        int lo = 0, hi = nimg - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (tab[mid].tile_base <= t) lo = mid;
            else hi = mid - 1;
        }
        const Args &A = tab[lo];
        const int lt = t - A.tile_base;
        const int tr = lt / A.ntc, tc = lt % A.ntc;
        amaze_tile(A, ar, -16 + tr * (TS - 32), -16 + tc * (TS - 32));
    });
}

// LibRaw-style border interpolation (:2984-3028): frame pixels get the 3x3 mean
// of each non-native colour.  Reads only native channels, writes only
// non-native ones, so it runs in place.
__global__ void border_kernel(float *img, int w, int h, uint32_t filters, int border) {
    const int total = lre::frame_pixels(w, h, border);
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        int row, col;
        lre::frame_pixel(t, w, h, border, row, col);
        if (col >= border && col < w - border && row >= border && row < h - border) continue;
        float sum[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int y = row > 0 ? row - 1 : 0; y < row + 2 && y < h; y++)
            for (int x = col > 0 ? col - 1 : 0; x < col + 2 && x < w; x++) {
                unsigned f = lre::bayer3(filters, (unsigned)y, (unsigned)x);
                sum[f] += img[((size_t)y * w + x) * 3 + f];
                sum[f + 4] += 1.f;
            }
        unsigned f = lre::bayer3(filters, (unsigned)row, (unsigned)col);
        for (unsigned c = 0; c < 3; c++)
            if (c != f && sum[c + 4] > 0) img[((size_t)row * w + col) * 3 + c] = sum[c] / sum[c + 4];
    }
}

int tiles_for(int n) {
    int k = 0;
    for (int o = -16; o < n; o += TS - 32) k++;
    return k;
}

}  // namespace

// raise the dynamic LDS limit of both kernels once per process
static hipError_t amaze_lds_attr() {
    static const hipError_t rc = [] {
        hipError_t e = hipFuncSetAttribute((const void *)amaze_kernel,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)LDS_BYTES);
        if (e == hipSuccess)
            e = hipFuncSetAttribute((const void *)amaze_multi_kernel,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)LDS_BYTES);
        return e;
    }();
    return rc;
}

// One workgroup fits a CU (LDS), and the tile queue balances the tiles, so
// the grid is one workgroup per CU.
static int amaze_slots(lre_ctx *ctx, int w, int h) {
    const int ntiles = tiles_for(w) * tiles_for(h);
    const int cap = ctx->num_cus > 0 ? ctx->num_cus : 256;
    return ntiles < cap ? (ntiles > 0 ? ntiles : 1) : cap;
}

// the slots' arenas, one spare arena of padding behind the last, the queue counter
static constexpr size_t arenas_bytes(int slots) {
    return (size_t)(slots + 1) * ARENA_FLOATS * sizeof(float) + 256;
}

size_t lre_amaze_scratch_bytes(lre_ctx *ctx, int w, int h) { return arenas_bytes(amaze_slots(ctx, w, h)); }

static void amaze_args(Args &A, const float *d_cfa, float *d_rgb, int w, int h, uint32_t filters,
                       const float cam_mul[4], float maximum_value, const float *d_maxv) {
    A.cfa_plane = d_cfa;
    A.rgb = d_rgb;
    A.W = w;
    A.H = h;
    float mx = cam_mul[0] < cam_mul[2] ? cam_mul[2] : cam_mul[0];
    float mn = cam_mul[2] < cam_mul[0] ? cam_mul[2] : cam_mul[0];
    float gain = (mn > 1e-6f) ? (mx / mn) : 1.0f;
    A.clip_pt = 1.0f / gain;
    A.clip_pt8 = 0.8f / gain;
    A.scale = 1.0f / maximum_value;
    unsigned cf[2][2];
    A.cfw = 0;
    for (int r = 0; r < 2; r++)
        for (int c = 0; c < 2; c++) {
            cf[r][c] = lre::bayer3(filters, r, c);
            A.cfw |= cf[r][c] << (16 * r + 8 * c);
        }
    if (cf[0][0] == 1) {
        if (cf[0][1] == 0) { A.ey = 0; A.ex = 1; } else { A.ey = 1; A.ex = 0; }
    } else {
        if (cf[0][0] == 0) { A.ey = 0; A.ex = 0; } else { A.ey = 1; A.ex = 1; }
    }
    A.d_maxv = d_maxv;
    A.ntc = tiles_for(w);
    A.ntiles = A.ntc * tiles_for(h);
    A.tile_base = 0;
    A.arenas = nullptr;
}

int lre_launch_amaze(lre_ctx *ctx, const float *d_cfa, float *d_rgb, int w, int h,
                     uint32_t filters, const float cam_mul[4], float maximum_value,
                     void *d_scratch, hipStream_t s, const float *d_maxv) {
    Args A;
    amaze_args(A, d_cfa, d_rgb, w, h, filters, cam_mul, maximum_value, d_maxv);
    A.arenas = static_cast<float *>(d_scratch);
    const int slots = amaze_slots(ctx, w, h);
    LRE_HIP_CHECK(amaze_lds_attr());
    hipLaunchKernelGGL(queue_reset, dim3(1), dim3(64), 0, s, A.arenas, slots);
    hipLaunchKernelGGL(amaze_kernel, dim3(slots), dim3(BLOCK), LDS_BYTES, s, A);
    LRE_HIP_CHECK(hipGetLastError());
    return lre_launch_border_interpolate(nullptr, d_rgb, w, h, filters, 4, s);
}

// ---- the variant that allocates nothing -------------------------------------
// Its tile arenas and queue counter live in this fixed device array (one copy
// per GPU, held until the process ends): LRE_AMAZE_NOALLOC_SLOTS arenas, the
// spare arena behind them and the counter, laid out as arenas_bytes() lays out
// an allocated scratch.  Every call on a GPU shares that GPU's copy.
static_assert(arenas_bytes(LRE_AMAZE_NOALLOC_SLOTS) <= LRE_AMAZE_NOALLOC_BYTES, "no-alloc arena size");
static __device__ __align__(256) float scratch[LRE_AMAZE_NOALLOC_BYTES / sizeof(float)];

int lre_amaze_noalloc_setup(void **d_arenas) {
    // the device copy of `scratch` (the host-side name is not a device address)
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        *d_arenas = nullptr;
        lre_set_error("device address of the no-alloc AMaZE arenas: %s", hipGetErrorString(e));
        return LRE_ERR_HIP;
    }
    *d_arenas = ptr;
    return LRE_OK;
}

int lre_launch_amaze_noalloc(const float *d_cfa, float *d_rgb, int w, int h, uint32_t filters,
                             const float cam_mul[4], float maximum_value, void *d_arenas, hipStream_t s) {
    Args A;
    amaze_args(A, d_cfa, d_rgb, w, h, filters, cam_mul, maximum_value, nullptr);
    A.arenas = static_cast<float *>(d_arenas);
    const int slots = A.ntiles < LRE_AMAZE_NOALLOC_SLOTS ? (A.ntiles > 0 ? A.ntiles : 1) : LRE_AMAZE_NOALLOC_SLOTS;
    LRE_HIP_CHECK(amaze_lds_attr());
    // the counter is cleared by a kernel (no memset node if the caller captures)
    hipLaunchKernelGGL(queue_reset, dim3(1), dim3(64), 0, s, A.arenas, slots);
    hipLaunchKernelGGL(amaze_kernel, dim3(slots), dim3(BLOCK), LDS_BYTES, s, A);
    LRE_HIP_CHECK(hipGetLastError());
    return lre_launch_border_interpolate(nullptr, d_rgb, w, h, filters, 4, s);
}

size_t lre_amaze_multi_table_bytes(int n) { return sizeof(Args) * (size_t)n; }

size_t lre_amaze_multi_scratch_bytes(int slots) { return arenas_bytes(slots); }

int lre_amaze_tiles(int w, int h) { return tiles_for(w) * tiles_for(h); }

int lre_amaze_multi_fill(const lre_demosaic_job *jobs, int n, void *h_table) {
    Args *tab = static_cast<Args *>(h_table);
    int total = 0;
    for (int k = 0; k < n; k++) {
        const lre_demosaic_job &j = jobs[k];
        amaze_args(tab[k], j.d_cfa, j.d_rgb, j.w, j.h, j.filters, j.cam_mul, j.maximum_value,
                   j.d_maxv);
        tab[k].tile_base = total;
        total += tab[k].ntiles;
    }
    return total;
}

int lre_launch_amaze_multi(const lre_demosaic_job *jobs, int n, const void *d_table, int total_tiles,
                           int slots, void *d_scratch, hipStream_t s) {
    if (n == 0) return LRE_OK;
    if (total_tiles > 0) {
        const int grid = slots < total_tiles ? slots : total_tiles;
        LRE_HIP_CHECK(amaze_lds_attr());
        hipLaunchKernelGGL(queue_reset, dim3(1), dim3(64), 0, s, static_cast<float *>(d_scratch), slots);
        hipLaunchKernelGGL(amaze_multi_kernel, dim3(grid), dim3(BLOCK), LDS_BYTES, s,
                           static_cast<const Args *>(d_table), n, total_tiles, slots,
                           static_cast<float *>(d_scratch));
        LRE_HIP_CHECK(hipGetLastError());
    }
    for (int k = 0; k < n; k++) {
        int rc = lre_launch_border_interpolate(nullptr, jobs[k].d_rgb, jobs[k].w, jobs[k].h,
                                               jobs[k].filters, 4, s);
        if (rc) return rc;
    }
    return LRE_OK;
}

int lre_launch_border_interpolate(const float *, float *d_rgb, int w, int h, uint32_t filters,
                                  int border, hipStream_t s) {
    const int total = 2 * border * w + 2 * border * h;
    const int blocks = (total + 255) / 256;
    hipLaunchKernelGGL(border_kernel, dim3(blocks > 0 ? blocks : 1), dim3(256), 0, s, d_rgb, w, h,
                       filters, border);
    LRE_HIP_CHECK(hipGetLastError());
    return LRE_OK;
}
