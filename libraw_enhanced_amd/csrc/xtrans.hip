// X-Trans 3-pass (Markesteijn) demosaic for gfx950
// (core/cpu_accelerator.cpp:1870-2676, border :1806-1861).
//
// Each workgroup processes one reference tile (114 x 114, step 98, first tile
// at (3,3)) in a private arena holding the 8 direction buffers, the derivative
// maps and the green bounds; the homogeneity maps live in LDS.  Every step of the reference is a
// loop over one class of CFA sites whose iterations neither read nor write each
// other's cells (checked on the X-Trans pattern), so each step runs as one
// parallel pass over its sites with a barrier after it.  Site positions along a
// row follow the reference's column walks in closed form: a stride of 3, or
// strides alternating between 1 and 2.  The green bounds keep the reference's
// per-row walk (one thread per row), because pairs of columns share a slot and
// the last write wins.  The homogeneity map is cleared per tile, so rows the
// tile does not write read as zero (the reference's per-thread buffers keep
// the previous tile's values there).
//
// Every interpolation step reads and writes one direction buffer at a time,
// and buffers 4-7 start as copies of 0-3 after the first pass.  So the tile
// runs direction by direction: buffer j is built in LDS (156 KB), goes through
// the first pass, is stored as buffer j, then goes through the second and
// third passes and is stored as buffer j+4; the derivative maps of both are
// computed from LDS before each store.
#include <cfloat>

#include "lre_common.h"
#include "lre_internal.h"

namespace {

constexpr int TS = 114;
constexpr int BLOCK = 1024;  // one tile per workgroup of 16 waves (measured best)
constexpr int LW = TS - 8, DW = TS - 10;
constexpr size_t DIRSZ = (size_t)TS * TS * 3;  // floats per direction buffer

// arena layout (bytes)
constexpr size_t A_RGB = 0;
constexpr size_t A_DRV = A_RGB + 8 * DIRSZ * 4;
constexpr size_t A_GM = A_DRV + (size_t)8 * DW * DW * 4;
constexpr size_t ARENA = ((A_GM + (size_t)TS * (TS / 2) * 8) + 255) & ~(size_t)255;

struct XArgs {
    const float *cfa;  // CFA plane (one sample per site)
    float *out;
    int W, H;
    float maxv;
    signed char xt[6][6];
    short hex[2][3][3][8];
    int sgrow, sgcol;
    int rshift[3];
    int ntc, ntiles;
    int tile_base;  // first tile of this image in a multi-image tile list
    char *arenas;
};

LRE_DEV int fcolx(const XArgs &A, int r, int c) { return A.xt[(r + 6) % 6][(c + 6) % 6]; }
LRE_DEV int isgreen(const XArgs &A, int r, int c) { return A.xt[r % 3][c % 3] & 1; }
// raw sample / maximum (division, as the reference), at a linear offset
LRE_DEV float rawv(const XArgs &A, int r, int c) { return A.cfa[(size_t)r * A.W + c] / A.maxv; }
LRE_DEV float rawhex(const XArgs &A, int r, int c, int hex) {
    return A.cfa[(size_t)(r * A.W + c + hex)] / A.maxv;
}
LRE_DEV float limf(float v, float lo, float hi) {
    float m = (hi < v) ? hi : v;
    return (lo < m) ? m : lo;
}
LRE_DEV float sq(float x) { return x * x; }

struct MM { float min, max; };

// first column >= c0 (and < lim) that is not green
LRE_DEV int first_nongreen(const XArgs &A, int row, int c0, int lim) {
    int c = c0;
    for (; c < lim; c++)
        if (!isgreen(A, row, c)) break;
    return c;
}

// k-th site of a walk starting at ls with stride pattern: fixed 3, or
// alternating c0, c0^3, ...
LRE_DEV int walk_col(int ls, int coff0, int k) {
    if (coff0 == 3) return ls + 3 * k;
    return ls + 3 * (k >> 1) + ((k & 1) ? coff0 : 0);
}

// LDS: one direction buffer, [row][col][rgb] like a buffer of the arena, then
// the tile's tables: the hexagon offsets of both strides (2 x 9 x 8 shorts),
// the 6x6 pattern, and one column walk per row for three of the steps
constexpr size_t LDS_BUF = DIRSZ * sizeof(float);
constexpr size_t LDS_HEX = LDS_BUF, LDS_XT = LDS_HEX + 2 * 9 * 8 * sizeof(short);
constexpr size_t LDS_WALK = LDS_XT + 48;
constexpr size_t LDS_BYTES = LDS_WALK + 3 * TS * sizeof(uint2);
static_assert(LDS_BUF % 16 == 0 && LDS_WALK % 8 == 0, "table alignment");

// One tile row's column walk, packed: first site ls (bits 0-15), stride
// pattern coff (16-19: 3, or alternating coff and 3 - coff), hexagon table
// index (row % 3) * 3 + col % 3 of even and odd steps (20-23, 24-27); and the
// colour of step k at byte k & 3 of the second word (the colours repeat with
// period 4 along a walk).
LRE_DEV uint2 make_walk(const XArgs &A, int row, int c0, int lim) {
    const int ls = first_nongreen(A, row, c0, lim);
    const int coff = A.rshift[row % 3] == 1 ? 3 : 1 + (fcolx(A, row, ls + 1) & 1);
    const int hx0 = (row % 3) * 3 + ls % 3, hx1 = (row % 3) * 3 + (ls + (coff == 3 ? 0 : coff)) % 3;
    uint2 w;
    w.x = (unsigned)ls | ((unsigned)coff << 16) | ((unsigned)hx0 << 20) | ((unsigned)hx1 << 24);
    w.y = 0;
    for (int q = 0; q < 4; q++) w.y |= (unsigned)fcolx(A, row, walk_col(ls, coff, q)) << (8 * q);
    return w;
}
// bound on the number of sites a column walk visits in one tile row, and in
// a row whose walk has a stride of 3
constexpr int KW = (2 * TS) / 3 + 2, K3 = TS / 3 + 1;

// Slots of a loop over the column walks of rows r0 .. r1-1: KW per row, but
// K3 for a row with a stride-3 walk (two rows of three in the X-Trans
// pattern), laid out in groups of three rows, so fewer lanes idle on slots
// past the end of a short walk.  Every site still has exactly one slot.
struct WalkSlots {
    int r0, n, g, c1, c2;
    float inv_g;
};
LRE_DEV WalkSlots walk_slots(const XArgs &A, int r0, int r1) {
    WalkSlots s;
    const int k0 = A.rshift[r0 % 3] == 1 ? K3 : KW, k1 = A.rshift[(r0 + 1) % 3] == 1 ? K3 : KW,
              k2 = A.rshift[(r0 + 2) % 3] == 1 ? K3 : KW;
    const int rows = r1 > r0 ? r1 - r0 : 0, rem = rows % 3;
    s.r0 = r0;
    s.g = k0 + k1 + k2;
    s.c1 = k0;
    s.c2 = k0 + k1;
    s.n = rows / 3 * s.g + (rem > 0 ? k0 : 0) + (rem > 1 ? k1 : 0);
    s.inv_g = 1.0f / (float)s.g;
    return s;
}
// slot t -> (row, k-th site of the row's walk)
LRE_DEV void walk_slot(const WalkSlots &s, int t, int &row, int &k) {
    int q, o;
    lre::divmod(t, s.g, s.inv_g, q, o);
    const int i = o >= s.c2 ? 2 : (o >= s.c1 ? 1 : 0);
    row = s.r0 + 3 * q + i;
    k = o - (i == 2 ? s.c2 : (i == 1 ? s.c1 : 0));
}

__device__ void xtrans_tile(const XArgs &A, char *ar, int top, int left) {
    const int W = A.W, H = A.H;
    extern __shared__ __align__(16) float lds[];  // 16-byte aligned behind the static LDS
    short *hexl = reinterpret_cast<short *>(reinterpret_cast<char *>(lds) + LDS_HEX);
    signed char *xtl = reinterpret_cast<signed char *>(lds) + LDS_XT;
    uint2 *walks = reinterpret_cast<uint2 *>(reinterpret_cast<char *>(lds) + LDS_WALK);
    float *rgb0 = reinterpret_cast<float *>(ar + A_RGB);
    float *drv = reinterpret_cast<float *>(ar + A_DRV);
    MM *gm = reinterpret_cast<MM *>(ar + A_GM);
    const int mrow0 = min(top + TS, H - 3), mcol0 = min(left + TS, W - 3);
    const int nr = mrow0 - top, nc = mcol0 - left;
#define RGB(base, d, r, c) ((base) + ((size_t)(d) * DIRSZ + ((size_t)(r) * TS + (c)) * 3))
#define PX(r, c) (lds + ((r) * TS + (c)) * 3)
#define GM(r, c) gm[(size_t)(r) * (TS / 2) + (c)]

    // tables: walks of green4 (from left), the pass green step (from left + 2)
    // and the R-at-B step (from left + 3)
    if (threadIdx.x < 2 * 9 * 8) hexl[threadIdx.x] = (&A.hex[0][0][0][0])[threadIdx.x];
    if (threadIdx.x < 36) xtl[threadIdx.x] = A.xt[threadIdx.x / 6][threadIdx.x % 6];
    for (int t = threadIdx.x; t < 3 * TS; t += BLOCK) {
        const int step = t / TS, row = top + t % TS;
        if (row >= mrow0) continue;
        walks[t] = step == 0 ? make_walk(A, row, left, mcol0)
                 : step == 1 ? make_walk(A, row, left + 2, mcol0 - 2)
                             : make_walk(A, row, left + 3, mcol0 - 1);
    }
    __syncthreads();
    auto walk_ls = [](uint2 w) { return (int)(w.x & 0xffffu); };
    auto walk_coff = [](uint2 w) { return (int)((w.x >> 16) & 15u); };
    auto walk_hex = [&](int set, uint2 w, int k) {
        return hexl + (set * 9 + (int)((w.x >> ((k & 1) ? 24 : 20)) & 15u)) * 8;
    };
    auto walk_colour = [](uint2 w, int k) { return (int)((w.y >> (8 * (k & 3))) & 255u); };

    // green min/max (:1973-2059).  The reference walks each row: on rows with
    // one non-green site per three columns every site fills its own slot
    // (slots are column pairs); on the other rows a first lone site may come,
    // then each pair of adjacent sites fills both its slots with the bounds of
    // its first site, and a last lone site may follow.  No two units of a
    // row share a slot with different values, so the units run in parallel.
    const WalkSlots sl_all = walk_slots(A, top, top + nr);
    for (int t = threadIdx.x; t < sl_all.n; t += BLOCK) {
        int row, u;
        walk_slot(sl_all, t, row, u);
        // the row's first site and stride, as its green-step walk holds them
        const uint2 w = walks[row - top];
        const int ls = walk_ls(w), coff = walk_coff(w);
        auto mmhex = [&](int col, const short *hx) {
            float lo = FLT_MAX, hi = 0.f;
            for (int c = 0; c < 6; c++) {
                const float v = rawhex(A, row, col, hx[c]);
                lo = lo < v ? lo : v;
                hi = hi > v ? hi : v;
            }
            return MM{lo, hi};
        };
        if (coff == 3) {
            const int col = ls + 3 * u;
            if (col < mcol0) GM(row - top, (col - left) >> 1) = mmhex(col, A.hex[0][row % 3][ls % 3]);
            continue;
        }
        int col;
        if (coff == 2) {
            if (u == 0) {
                if (ls < mcol0) GM(row - top, (ls - left) >> 1) = mmhex(ls, A.hex[0][row % 3][ls % 3]);
                continue;
            }
            col = ls + 2 + 3 * (u - 1);
        } else {
            col = ls + 3 * u;
        }
        if (col >= mcol0) continue;
        const MM m = mmhex(col, A.hex[0][row % 3][col % 3]);
        GM(row - top, (col - left) >> 1) = m;
        if (col < mcol0 - 1) GM(row - top, (col + 1 - left) >> 1) = m;
    }
    __syncthreads();  // the green bounds, before the green step reads them

    // rows of the R/B steps for 2x2 green blocks: those with (row - sgrow) % 3 != 0
    int topstart = top + 2;
    for (; topstart < mrow0 - 2; topstart++)
        if ((topstart - A.sgrow) % 3) break;
    int ls22 = left + 2;
    for (; ls22 < mcol0 - 2; ls22++)
        if ((ls22 - A.sgcol) % 3) break;
    const int coff22 = 2 - (fcolx(A, topstart, ls22 + 1) & 1);
    const int adj22 = ((topstart - A.sgrow) % 3 + 3) % 3 == 2 ? 1 : 0;  // q-th valid row: topstart + (3q + adj) / 2
    int nrow22 = 0;
    while (topstart + (3 * nrow22 + adj22) / 2 < mrow0 - 2) nrow22++;

    const WalkSlots sl_pg = walk_slots(A, top + 2, mrow0 - 2), sl_rb = walk_slots(A, top + 3, mrow0 - 3);

    // green along four directions (:2075-2128), the value that lands in buffer j
    auto green4 = [&](int j) {
        for (int t = threadIdx.x; t < sl_all.n; t += BLOCK) {
            int row, k;
            walk_slot(sl_all, t, row, k);
            const uint2 w = walks[row - top];
            const int coff0 = walk_coff(w);
            const int col = walk_col(walk_ls(w), coff0, k);
            if (col >= mcol0) continue;
            const short *hx = walk_hex(0, w, k);
            const int c = j ^ (coff0 == 3 ? 0 : 1);
            float color;
            if (c == 0)
                color = 0.6796875f * (rawhex(A, row, col, hx[1]) + rawhex(A, row, col, hx[0])) -
                        0.1796875f * (rawhex(A, row, col, 2 * hx[1]) + rawhex(A, row, col, 2 * hx[0]));
            else if (c == 1)
                color = 0.87109375f * rawhex(A, row, col, hx[3]) + rawhex(A, row, col, hx[2]) * 0.12890625f +
                        0.359375f * (rawhex(A, row, col, 0) - rawhex(A, row, col, -hx[2]));
            else
                color = 0.640625f * rawhex(A, row, col, hx[2 + c]) + 0.359375f * rawhex(A, row, col, -2 * hx[2 + c]) +
                        0.12890625f * (2.f * rawhex(A, row, col, 0) - rawhex(A, row, col, 3 * hx[2 + c]) -
                                       rawhex(A, row, col, -3 * hx[2 + c]));
            const MM m = GM(row - top, (col - left) >> 1);
            PX(row - top, col - left)[1] = limf(color, m.min, m.max);
        }
    };

    // green from closer pixels (:2141-2182), passes 1 and 2: the d of 3..5
    // whose buffer is j
    auto pass_green = [&](int j) {
        for (int t = threadIdx.x; t < sl_pg.n; t += BLOCK) {
            int row, k;
            walk_slot(sl_pg, t, row, k);
            const uint2 w = walks[TS + row - top];
            const int coff0 = walk_coff(w);
            const int d = coff0 == 3 ? j + 2 : (j ^ 1) + 2;
            if (d < 3 || d > 5) continue;
            const int col = walk_col(walk_ls(w), coff0, k);
            if (col >= mcol0 - 2) continue;
            const int f = walk_colour(w, k);
            const short *hx = walk_hex(1, w, k);
            const MM m = GM(row - top, (col - left) >> 1);
            float *rix = PX(row - top, col - left);
            const int o1 = 3 * hx[d], o2 = 3 * (-2 * hx[d]);
            const float val = 0.33333333f * (rix[o2 + 1] + 2 * (rix[o1 + 1] - rix[o1 + f]) - rix[o2 + f]) + rix[f];
            rix[1] = limf(val, m.min, m.max);
        }
    };

    // R and B at solitary green sites (:2184-2222).  The reference walks d = 0..5
    // with one buffer pointer: buffer 0 takes d = 0, buffer 1 d = 1, buffer 2
    // the better of d = 2, 3 and buffer 3 the better of d = 4, 5.
    auto solitary = [&](int j) {
        const int rstart = (top - A.sgrow + 4) / 3 * 3 + A.sgrow;
        const int cstart = (left - A.sgcol + 4) / 3 * 3 + A.sgcol;
        const int nrs = mrow0 - 2 > rstart ? (mrow0 - 2 - rstart + 2) / 3 : 0;
        const int ncs = mcol0 - 2 > cstart ? (mcol0 - 2 - cstart + 2) / 3 : 0;
        const int dlo = j < 2 ? j : 2 * j - 2, dhi = j < 2 ? j : 2 * j - 1;
        for (int t = threadIdx.x; t < nrs * ncs; t += BLOCK) {
            const int row = rstart + 3 * (t / ncs), col = cstart + 3 * (t % ncs);
            const int h0 = fcolx(A, row, cstart + 1) ^ (((t % ncs) & 1) ? 2 : 0);
            float *rix = PX(row - top, col - left);
            float cr[2] = {0.f, 0.f}, cb[2] = {0.f, 0.f}, diff[2] = {0.f, 0.f};  // colours 0 and 2 of d = dlo + q
#pragma unroll
            for (int q = 0; q < 2; q++) {
                const int d = dlo + q;
                if (d > dhi) break;
                const int i = (d & 1) ? TS : 1;
                int h = h0 ^ ((d & 1) ? 2 : 0);
                float col[2];
#pragma unroll
                for (int c = 0; c < 2; c++, h ^= 2) {
                    const int o = i << c;
                    const float g = rix[1] + rix[1] - rix[3 * o + 1] - rix[3 * -o + 1];
                    col[c] = g + rix[3 * o + h] + rix[3 * -o + h];
                    if (d > 1) diff[q] += sq(rix[3 * o + 1] - rix[3 * -o + 1] - rix[3 * o + h] + rix[3 * -o + h]) + sq(g);
                }
                const bool h_is_0 = (h0 ^ ((d & 1) ? 2 : 0)) == 0;  // colour of the c = 0 step
                cr[q] = h_is_0 ? col[0] : col[1];
                cb[q] = h_is_0 ? col[1] : col[0];
            }
            // d = dhi, or d = dhi - 1 of a pair when its difference is smaller
            const bool first = dhi > 2 && diff[0] < diff[1];
            const int last = dhi - dlo;
            rix[0] = 0.5f * (first ? cr[0] : (last ? cr[1] : cr[0]));
            rix[2] = 0.5f * (first ? cb[0] : (last ? cb[1] : cb[0]));
        }
    };

    // R at B sites and B at R sites (:2224-2266), buffer j of the pass
    auto r_at_b = [&](int j) {
        for (int t = threadIdx.x; t < sl_rb.n; t += BLOCK) {
            int row, k;
            walk_slot(sl_rb, t, row, k);
            const uint2 w = walks[2 * TS + row - top];
            const int col = walk_col(walk_ls(w), walk_coff(w), k);
            if (col >= mcol0 - 3) continue;
            const int c = ((row - A.sgrow) % 3) ? TS : 1;
            const int h = 3 * (c ^ TS ^ 1);
            const int f = 2 - walk_colour(w, k);
            float *rix = PX(row - top, col - left);
            const int i = j > 1 || ((j ^ c) & 1) ||
                                  ((fabsf(rix[1] - rix[3 * c + 1]) + fabsf(rix[1] - rix[3 * -c + 1])) <
                                   2.f * (fabsf(rix[1] - rix[3 * h + 1]) + fabsf(rix[1] - rix[3 * -h + 1])))
                              ? c : h;
            rix[f] = rix[1] + 0.5f * (rix[3 * i + f] + rix[3 * -i + f] - rix[3 * i + 1] - rix[3 * -i + 1]);
        }
    };

    // R and B for 2x2 green blocks (:2268-2314), buffer j of the pass (d = 2j)
    auto two_by_two = [&](int j) {
        const int d = 2 * j;
        for (int t = threadIdx.x; t < nrow22 * KW; t += BLOCK) {
            const int q = t / KW, k = t % KW;
            const int row = topstart + (3 * q + adj22) / 2;
            const int col = walk_col(ls22, coff22, k);
            if (col >= mcol0 - 2) continue;
            const short *hx = hexl + (9 + (row % 3) * 3 + ((k & 1) ? ls22 + coff22 : ls22) % 3) * 8;
            float *rix = PX(row - top, col - left);
            if (hx[d] + hx[d + 1]) {
                const float g = 3 * rix[1] - 2 * rix[3 * hx[d] + 1] - rix[3 * hx[d + 1] + 1];
                for (int c = 0; c < 4; c += 2)
                    rix[c] = (g + 2 * rix[3 * hx[d] + c] + rix[3 * hx[d + 1] + c]) * 0.33333333f;
            } else {
                const float g = 2 * rix[1] - rix[3 * hx[d] + 1] - rix[3 * hx[d + 1] + 1];
                for (int c = 0; c < 4; c += 2) rix[c] = (g + rix[3 * hx[d] + c] + rix[3 * hx[d + 1] + c]) * 0.5f;
            }
        }
    };

    // YPbPr derivatives (:2346-2419) of buffer d, luma/chroma evaluated where needed
    const int mrow_d = mrow0 - top, mcol_d = mcol0 - left;
    auto derivatives = [&](int d) {
        const int rows = max(mrow_d - 10, 0), cols = max(mcol_d - 10, 0);
        const int fo = (d & 3) == 0 ? 1 : (d & 3) == 1 ? TS : (d & 3) == 2 ? TS + 1 : TS - 1;  // :1878
        // f offsets in the reference index a (TS-8)-wide plane; here rows are TS wide
        const int dr = fo == 1 ? 0 : 1, dc = fo == 1 ? 1 : fo - TS;
        for (int t = threadIdx.x; t < rows * cols; t += BLOCK) {
            const int row = 5 + t / cols, col = 5 + t % cols;
            auto ypp = [&](int r, int c, float &y, float &pb, float &pr) {
                const float *p = PX(r, c);
                y = 0.2627f * p[0] + 0.6780f * p[1] + 0.0593f * p[2];
                pb = (p[2] - y) * 0.56433f;
                pr = (p[0] - y) * 0.67815f;
            };
            float y0, u0, v0, y1, u1, v1, y2, u2, v2;
            ypp(row, col, y0, u0, v0);
            ypp(row + dr, col + dc, y1, u1, v1);
            ypp(row - dr, col - dc, y2, u2, v2);
            drv[(size_t)d * DW * DW + (size_t)(row - 5) * DW + col - 5] =
                sq(2 * y0 - y1 - y2) + sq(2 * u0 - u1 - u2) + sq(2 * v0 - v1 - v2);
        }
    };

    // LDS buffer -> arena buffer d.  Only the rows the final average reads go
    // out: [rs, re) with the bounds of the last pass below.
    const int row_s = top < 8 ? top : 8;
    const int row_e = max(((H - top < TS + 4) ? H - top + 2 : mrow_d) - 8, row_s);
    const int q0 = row_s * TS * 3 / 4, q1 = (row_e * TS * 3 + 3) / 4;  // float4 range
    auto store_buffer = [&](int d) {
        const float4 *src = reinterpret_cast<const float4 *>(lds);
        float4 *dst = reinterpret_cast<float4 *>(RGB(rgb0, d, 0, 0));
        typedef float f4 __attribute__((ext_vector_type(4)));
        const f4 *s4 = reinterpret_cast<const f4 *>(src);
        f4 *d4 = reinterpret_cast<f4 *>(dst);
        for (int t = q0 + threadIdx.x; t < q1; t += BLOCK) __builtin_nontemporal_store(s4[t], &d4[t]);
    };

    for (int j = 0; j < 4; j++) {
        // buffer j: the raw sample at its colour, zero elsewhere; the reference
        // copies buffer 0 into 1-3 one float short, so their last float is 0
        for (int t = threadIdx.x; t < TS * TS; t += BLOCK) {
            const int r = t / TS, c = t % TS;
            float v[3] = {0.f, 0.f, 0.f};
            if (r < nr && c < nc) v[xtl[((r + top) % 6) * 6 + (c + left) % 6]] = rawv(A, r + top, c + left);
            float *p = lds + 3 * t;
            p[0] = v[0];
            p[1] = v[1];
            p[2] = (j > 0 && t == TS * TS - 1) ? 0.f : v[2];
        }
        __syncthreads();
        green4(j);
        __syncthreads();
        for (int pass = 0; pass < 3; pass++) {
            if (pass == 1) {
                // buffers 0-3 are copied into 4-7 one float short: buffer 7's last float is 0
                store_buffer(j);
                derivatives(j);
                __syncthreads();
                if (j == 3 && threadIdx.x == 0) lds[DIRSZ - 1] = 0.f;
                __syncthreads();
            }
            if (pass) {
                pass_green(j);
                __syncthreads();
            }
            solitary(j);
            __syncthreads();
            r_at_b(j);
            __syncthreads();
            two_by_two(j);
            __syncthreads();
        }
        store_buffer(j + 4);
        derivatives(j + 4);
        __syncthreads();
    }

    int mrow = mrow_d, mcol = mcol_d;
#define DRV(d, r, c) drv[(size_t)(d) * DW * DW + (size_t)(r) * DW + (c)]
    // homogeneity maps in LDS: directions 0-3 and 4-7 as the bytes of one word
    // per pixel (a 5x5 sum of counts up to 9 stays below 256 in every byte);
    // cleared per tile, so rows the tile does not write read as zero
    unsigned *hw = reinterpret_cast<unsigned *>(lds);
    unsigned char *hb = reinterpret_cast<unsigned char *>(lds);
    static_assert(2 * TS * TS * sizeof(unsigned) <= LDS_BUF, "homogeneity maps fit the LDS buffer");
    for (int t = threadIdx.x; t < 2 * TS * TS; t += BLOCK) hw[t] = 0u;
    __syncthreads();
    {
        // four rows per thread share the 3x3 windows of each direction
        const int rows = max(mrow - 12, 0), cols = max(mcol - 12, 0);
        const int ng = (rows + 3) / 4;
        for (int t = threadIdx.x; t < ng * cols; t += BLOCK) {
            const int row0 = 6 + 4 * (t / cols), col = 6 + t % cols;
            const int nq = min(4, 6 + rows - row0);
            float tr[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int r = row0 + min(q, nq - 1) - 5;
                float m = DRV(0, r, col - 5) < DRV(1, r, col - 5) ? DRV(0, r, col - 5) : DRV(1, r, col - 5);
                for (int d = 2; d < 8; d++) m = DRV(d, r, col - 5) < m ? DRV(d, r, col - 5) : m;
                tr[q] = m * 8;
            }
            for (int d = 0; d < 8; d++) {
                float w[6][3];  // rows row0-1 .. row0+4, columns col-1 .. col+1
#pragma unroll
                for (int a = 0; a < 6; a++) {
                    const int r = min(row0 - 6 + a, DW - 1);  // rows past the last group are not used
#pragma unroll
                    for (int b = 0; b < 3; b++) w[a][b] = DRV(d, r, col - 6 + b);
                }
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    if (q >= nq) break;
                    unsigned s = 0;
#pragma unroll
                    for (int a = q; a < q + 3; a++)
#pragma unroll
                        for (int b = 0; b < 3; b++) s += w[a][b] <= tr[q] ? 1u : 0u;
                    hb[(((d >> 2) * TS + row0 + q) * TS + col) * 4 + (d & 3)] = (unsigned char)s;
                }
            }
        }
    }
    __syncthreads();
    if (H - top < TS + 4) mrow = H - top + 2;
    if (W - left < TS + 4) mcol = W - left + 2;
    const int rs = top < 8 ? top : 8, cs = left < 8 ? left : 8;
    {
        // 5x5 sums of the maps (exact integers: the running form of the
        // reference gives the same), their maximum less an eighth, and the
        // average of the most homogeneous directions, per pixel
        const int rows = max(mrow - 8 - rs, 0), cols = max(mcol - 8 - cs, 0);
        // AVU pixels per thread: the selected directions' samples of both are
        // loaded before any is summed (the sums keep the direction order)
        constexpr int AVU = 2;
        for (int base = threadIdx.x; base < rows * cols; base += BLOCK * AVU) {
            unsigned sel[AVU];
            int prow[AVU], pcol[AVU];
            float p[AVU][8][3];
#pragma unroll
            for (int u = 0; u < AVU; u++) {
                const int t = base + u * BLOCK;
                sel[u] = 0u;
                prow[u] = pcol[u] = 0;
                if (t >= rows * cols) continue;
                const int row = rs + t / cols, col = cs + t % cols;
                prow[u] = row;
                pcol[u] = col;
                unsigned s0 = 0, s1 = 0;
                for (int v = -2; v <= 2; v++)
#pragma unroll
                    for (int h = -2; h <= 2; h++) {
                        const int k = (row + v) * TS + col + h;
                        s0 += hw[k];
                        s1 += hw[TS * TS + k];
                    }
                unsigned char hm[8];
#pragma unroll
                for (int d = 0; d < 4; d++) {
                    hm[d] = (unsigned char)(s0 >> (8 * d));
                    hm[d + 4] = (unsigned char)(s1 >> (8 * d));
                }
                unsigned char mx = hm[0];
#pragma unroll
                for (int d = 1; d < 8; d++) mx = mx < hm[d] ? hm[d] : mx;
                mx -= mx >> 3;
#pragma unroll
                for (int d = 4; d < 8; d++) {
                    if (hm[d - 4] < hm[d]) hm[d - 4] = 0;
                    else if (hm[d - 4] > hm[d]) hm[d] = 0;
                }
#pragma unroll
                for (int d = 0; d < 8; d++) sel[u] |= (hm[d] >= mx ? 1u : 0u) << d;
            }
#pragma unroll
            for (int u = 0; u < AVU; u++)
#pragma unroll
                for (int d = 0; d < 8; d++)
                    if ((sel[u] >> d) & 1u) {
                        const float *q = RGB(rgb0, d, prow[u], pcol[u]);
#pragma unroll
                        for (int c = 0; c < 3; c++) p[u][d][c] = __builtin_nontemporal_load(q + c);
                    }
#pragma unroll
            for (int u = 0; u < AVU; u++) {
                if (base + u * BLOCK >= rows * cols) continue;
                float avg[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int d = 0; d < 8; d++)
                    if ((sel[u] >> d) & 1u) {
#pragma unroll
                        for (int c = 0; c < 3; c++) avg[c] += p[u][d][c];
                        avg[3] += 1.f;
                    }
                float *o = A.out + ((size_t)(prow[u] + top) * W + pcol[u] + left) * 3;
                const float a0 = avg[0] / avg[3], a1 = avg[1] / avg[3], a2 = avg[2] / avg[3];
                o[0] = (0.f < a0) ? a0 : 0.f;
                o[1] = (0.f < a1) ? a1 : 0.f;
                o[2] = (0.f < a2) ? a2 : 0.f;
            }
        }
    }
    __syncthreads();
#undef DRV
#undef RGB
#undef PX
#undef GM
}

// The tile queue's counter sits behind the arenas and the spare one.
LRE_DEV unsigned *queue_of(char *arenas, int slots) {
    return reinterpret_cast<unsigned *>(arenas + (size_t)(slots + 1) * ARENA);
}

__global__ void queue_reset(char *arenas, int slots) {
    if (threadIdx.x == 0) *queue_of(arenas, slots) = 0u;
}

__global__ __launch_bounds__(BLOCK) void xtrans_kernel(XArgs A) {
    char *ar = A.arenas + (size_t)blockIdx.x * ARENA;
    lre::tile_queue(queue_of(A.arenas, gridDim.x), A.ntiles, [&](int t) {
        xtrans_tile(A, ar, 3 + (t / A.ntc) * (TS - 16), 3 + (t % A.ntc) * (TS - 16));
    });
}

// One tile queue over several images (table in device memory): tile t belongs
// to the last image whose tile_base is <= t.
__global__ __launch_bounds__(BLOCK) void xtrans_multi_kernel(const XArgs *__restrict__ tab, int nimg,
                                                             int total_tiles, int slots, char *arenas) {
    char *ar = arenas + (size_t)blockIdx.x * ARENA;
    lre::tile_queue(queue_of(arenas, slots), total_tiles, [&](int t) {
        int lo = 0, hi = nimg - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (tab[mid].tile_base <= t) lo = mid;
            else hi = mid - 1;
        }
        const XArgs &A = tab[lo];
        const int lt = t - A.tile_base;
        xtrans_tile(A, ar, 3 + (lt / A.ntc) * (TS - 16), 3 + (lt % A.ntc) * (TS - 16));
    });
}

// copies a demosaic maximum produced on the device into an image's table entry
__global__ void xtrans_set_maxv(XArgs *entry, const float *d_maxv) {
    if (threadIdx.x == 0 && blockIdx.x == 0) entry->maxv = *d_maxv;
}

// weighted 3x3 border interpolation (:1806-1861) on the 8-px frame
__global__ void xtrans_border(XArgs A, int border, const float *d_maxv) {
    if (d_maxv) A.maxv = *d_maxv;
    const int W = A.W, H = A.H;
    const float wt[3][3] = {{0.25f, 0.5f, 0.25f}, {0.5f, 0.f, 0.5f}, {0.25f, 0.5f, 0.25f}};
    const int total = lre::frame_pixels(W, H, border);
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        int row, col;
        lre::frame_pixel(t, W, H, border, row, col);
        const size_t i = (size_t)row * W + col;
        float sum[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (int y = row > 0 ? row - 1 : 0, v = row == 0 ? 0 : -1; y < row + 2 && y < H; y++, v++)
            for (int x = col > 0 ? col - 1 : 0, h = col == 0 ? 0 : -1; x < col + 2 && x < W; x++, h++) {
                const int f = fcolx(A, y, x);
                sum[f] += rawv(A, y, x) * wt[v + 1][h + 1];
                sum[f + 3] += wt[v + 1][h + 1];
            }
        float *o = A.out + i * 3;
        switch (fcolx(A, row, col)) {
        case 0:
            o[0] = rawv(A, row, col);
            o[1] = sum[1] / sum[4];
            o[2] = sum[2] / sum[5];
            break;
        case 1:
            if (sum[3] == 0.f) {
                const float v = rawv(A, row, col);
                o[0] = o[1] = o[2] = v;
            } else {
                o[0] = sum[0] / sum[3];
                o[1] = rawv(A, row, col);
                o[2] = sum[2] / sum[5];
            }
            break;
        default:
            o[0] = sum[0] / sum[3];
            o[1] = sum[1] / sum[4];
            o[2] = rawv(A, row, col);
        }
    }
}

int tiles_for(int n) {
    int k = 0;
    for (int o = 3; o < n - 19; o += TS - 16) k++;
    return k;
}

void setup(XArgs &A, const signed char *xt36) {
    for (int i = 0; i < 36; i++) A.xt[i / 6][i % 6] = xt36[i];
    auto isg = [&](int r, int c) { return A.xt[r % 3][c % 3] & 1; };
    static const short orth[12] = {1, 0, 0, 1, -1, 0, 0, -1, 1, 0, 0, 1};
    static const short patt[2][16] = {{0, 1, 0, -1, 2, 0, -1, 0, 1, 1, 1, -1, 0, 0, 0, 0},
                                      {0, 1, 0, -2, 1, 0, -2, 0, 1, 1, -2, -2, 1, -1, -1, 1}};
    A.sgrow = A.sgcol = 0;
    for (int a = 0; a < 2; a++)
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++)
                for (int k = 0; k < 8; k++) A.hex[a][r][c][k] = 0;
    for (int row = 0; row < 3; row++)
        for (int col = 0; col < 3; col++) {
            const int gint = isg(row, col);
            int ng = 0;
            for (int d = 0; d < 10; d += 2) {
                if (isg(row + orth[d] + 6, col + orth[d + 2] + 6)) ng = 0;
                else ng++;
                if (ng == 4) { A.sgrow = row; A.sgcol = col; }
                if (ng == gint + 1)
                    for (int c = 0; c < 8; c++) {
                        const int v = orth[d] * patt[gint][c * 2] + orth[d + 1] * patt[gint][c * 2 + 1];
                        const int h = orth[d + 2] * patt[gint][c * 2] + orth[d + 3] * patt[gint][c * 2 + 1];
                        A.hex[0][row][col][c ^ (gint * 2 & d)] = (short)(h + v * A.W);
                        A.hex[1][row][col][c ^ (gint * 2 & d)] = (short)(h + v * TS);
                    }
            }
        }
    for (int row = 0; row < 3; row++) {
        int gc = 0;
        for (int col = 0; col < 3; col++) gc += isg(row, col);
        A.rshift[row] = gc == 2;
    }
}

}  // namespace

// raise the dynamic LDS limit of both tile kernels once per process
static hipError_t xtrans_lds_attr() {
    static const hipError_t rc = [] {
        hipError_t e = hipFuncSetAttribute((const void *)xtrans_kernel,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)LDS_BYTES);
        if (e == hipSuccess)
            e = hipFuncSetAttribute((const void *)xtrans_multi_kernel,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)LDS_BYTES);
        return e;
    }();
    return rc;
}

// One workgroup fits a CU (LDS), and the tile queue balances the tiles, so
// the grid is one workgroup per CU.
static int xtrans_slots(lre_ctx *ctx, int w, int h) {
    const int ntiles = tiles_for(w) * tiles_for(h);
    const int cap = ctx->num_cus > 0 ? ctx->num_cus : 256;
    return ntiles < cap ? (ntiles > 0 ? ntiles : 1) : cap;
}

// the slots' arenas, one spare arena behind the last, the queue counter
static size_t arenas_bytes(int slots) { return (size_t)(slots + 1) * ARENA + 256; }

size_t lre_xtrans_scratch_bytes(lre_ctx *ctx, int w, int h) { return arenas_bytes(xtrans_slots(ctx, w, h)); }

int lre_launch_xtrans3(lre_ctx *ctx, const float *d_cfa, float *d_rgb, int w, int h,
                       const signed char xtrans36[36], float maximum_value, void *d_scratch,
                       hipStream_t s) {
    if (2 * w + 2 > 32767 || h < 1) {
        lre_set_error("X-Trans width %d exceeds the 16-bit hexagon offsets", w);
        return LRE_ERR_ARG;
    }
    XArgs A;
    A.cfa = d_cfa;
    A.out = d_rgb;
    A.W = w;
    A.H = h;
    A.maxv = maximum_value;
    setup(A, xtrans36);
    A.ntc = tiles_for(w);
    A.ntiles = A.ntc * tiles_for(h);
    A.tile_base = 0;
    A.arenas = static_cast<char *>(d_scratch);
    if (A.ntiles > 0) {
        LRE_HIP_CHECK(xtrans_lds_attr());
        const int slots = xtrans_slots(ctx, w, h);
        hipLaunchKernelGGL(queue_reset, dim3(1), dim3(64), 0, s, A.arenas, slots);
        hipLaunchKernelGGL(xtrans_kernel, dim3(slots), dim3(BLOCK), LDS_BYTES, s, A);
        LRE_HIP_CHECK(hipGetLastError());
    }
    const size_t frame = (size_t)lre::frame_pixels(w, h, 8);
    size_t blocks = (frame + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(xtrans_border, dim3((unsigned)blocks), dim3(256), 0, s, A, 8,
                       static_cast<const float *>(nullptr));
    LRE_HIP_CHECK(hipGetLastError());
    return LRE_OK;
}

size_t lre_xtrans_multi_table_bytes(int n) { return sizeof(XArgs) * (size_t)n; }

size_t lre_xtrans_multi_scratch_bytes(int slots) { return arenas_bytes(slots); }

int lre_xtrans_tiles(int w, int h) { return tiles_for(w) * tiles_for(h); }

int lre_xtrans_multi_fill(const lre_demosaic_job *jobs, int n, void *h_table) {
    XArgs *tab = static_cast<XArgs *>(h_table);
    int total = 0;
    for (int k = 0; k < n; k++) {
        const lre_demosaic_job &j = jobs[k];
        if (2 * j.w + 2 > 32767 || j.h < 1) {
            lre_set_error("X-Trans width %d exceeds the 16-bit hexagon offsets", j.w);
            return LRE_ERR_ARG;
        }
        XArgs &A = tab[k];
        A.cfa = j.d_cfa;
        A.out = j.d_rgb;
        A.W = j.w;
        A.H = j.h;
        A.maxv = j.maximum_value;
        setup(A, j.xtrans);
        A.ntc = tiles_for(j.w);
        A.ntiles = A.ntc * tiles_for(j.h);
        A.tile_base = total;
        A.arenas = nullptr;
        total += A.ntiles;
    }
    return total;
}

int lre_launch_xtrans_multi(const lre_demosaic_job *jobs, int n, const void *h_table, void *d_table,
                            int total_tiles, int slots, void *d_scratch, hipStream_t s) {
    if (n == 0) return LRE_OK;
    const XArgs *tab = static_cast<const XArgs *>(h_table);
    XArgs *dtab = static_cast<XArgs *>(d_table);
    for (int k = 0; k < n; k++)
        if (jobs[k].d_maxv)
            hipLaunchKernelGGL(xtrans_set_maxv, dim3(1), dim3(64), 0, s, dtab + k, jobs[k].d_maxv);
    if (total_tiles > 0) {
        const int grid = slots < total_tiles ? slots : total_tiles;
        LRE_HIP_CHECK(xtrans_lds_attr());
        hipLaunchKernelGGL(queue_reset, dim3(1), dim3(64), 0, s, static_cast<char *>(d_scratch), slots);
        hipLaunchKernelGGL(xtrans_multi_kernel, dim3(grid), dim3(BLOCK), LDS_BYTES, s,
                           static_cast<const XArgs *>(dtab), n, total_tiles, slots,
                           static_cast<char *>(d_scratch));
        LRE_HIP_CHECK(hipGetLastError());
    }
    for (int k = 0; k < n; k++) {
        const size_t frame = (size_t)lre::frame_pixels(jobs[k].w, jobs[k].h, 8);
        size_t blocks = (frame + 255) / 256;
        if (blocks > 8192) blocks = 8192;
        if (blocks < 1) blocks = 1;
        hipLaunchKernelGGL(xtrans_border, dim3((unsigned)blocks), dim3(256), 0, s, tab[k], 8,
                           jobs[k].d_maxv);
        LRE_HIP_CHECK(hipGetLastError());
    }
    return LRE_OK;
}
