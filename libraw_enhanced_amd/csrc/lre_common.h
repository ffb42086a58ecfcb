// Shared device/host helpers of the HIP develop path (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>

#define LRE_DEV __device__ __forceinline__

namespace lre {

// Colour of a Bayer site from LibRaw's 32-bit `filters` word (0 R, 1 G, 2 B; the
// second green, 3, folds to 1).  Same bit formula as the reference's
// fcol_bayer_native / fcol_bayer (core/metal/shader_common.h:25-31).
LRE_DEV __host__ unsigned bayer_native(uint32_t filters, unsigned row, unsigned col) {
    return filters >> ((((row << 1) & 14) | (col & 1)) << 1) & 3;
}
LRE_DEV __host__ unsigned bayer3(uint32_t filters, unsigned row, unsigned col) {
    unsigned n = bayer_native(filters, row, col);
    return n == 3 ? 1 : n;
}

LRE_DEV float sq(float x) { return x * x; }
// std::min / std::max operand semantics (NaN handling follows the comparison).
LRE_DEV float min_std(float a, float b) { return (b < a) ? b : a; }
LRE_DEV float max_std(float a, float b) { return (a < b) ? b : a; }
LRE_DEV float clamp_std(float v, float lo, float hi) { return v < lo ? lo : (hi < v ? hi : v); }
// Median of three by the reference's three compare-swaps (shader_common.h:55-61).
LRE_DEV float med3(float a, float b, float c) {
    float t;
    if (a > b) { t = a; a = b; b = t; }
    if (b > c) { t = b; b = c; c = t; }
    if (a > b) { t = a; a = b; b = t; }
    return b;
}

// q = t / d, r = t % d for 0 <= t < 2^24, 0 < d < 2^12, from a float estimate
// of the quotient (off by at most one) and one correction step.
LRE_DEV void divmod(int t, int d, float inv_d, int &q, int &r) {
    q = (int)((float)t * inv_d);
    r = t - q * d;
    if (r < 0) { q--; r += d; }
    else if (r >= d) { q++; r -= d; }
}

// A grid-stride walk over the n = rows * cols entries of a row-major image:
// i = first, first + stride, ... < n, with (row, col) = (i / cols, i % cols)
// advanced by the stride's quotient and remainder (one 64-bit division per
// thread instead of one per entry).
struct GridWalk {
    size_t i, n, stride;
    int row, col, cols, drow, dcol;
    LRE_DEV GridWalk(size_t first, size_t n_, size_t stride_, int cols_) : i(first), n(n_), stride(stride_), cols(cols_) {
        const size_t q = first / (size_t)cols_, dq = stride_ / (size_t)cols_;
        row = (int)q;
        col = (int)(first - q * (size_t)cols_);
        drow = (int)dq;
        dcol = (int)(stride_ - dq * (size_t)cols_);
    }
    LRE_DEV bool more() const { return i < n; }
    LRE_DEV void next() {
        i += stride;
        row += drow;
        col += dcol;
        if (col >= cols) {
            col -= cols;
            row++;
        }
    }
};

// Stages a rows x cols window into LDS: element (r, c) = load(r, c), written by
// store(r, c, value).  Each thread issues U independent global loads before
// its first LDS store, so their latencies overlap.
// When a row fits the workgroup, each thread keeps one column and walks the
// rows (neighbouring threads take neighbouring columns), so the column's part
// of the address and of any clamping is computed once per thread.
template <int NT, int U = 8, class Load, class Store>
LRE_DEV void stage2d(int rows, int cols, Load load, Store store) {
    if (cols <= NT) {
        using V = decltype(load(0, 0));
        const int per = NT / cols;  // rows per pass
        const int c = (int)threadIdx.x % cols, r0 = (int)threadIdx.x / cols;
        if (r0 >= per) return;
        for (int base = r0; base < rows; base += per * U) {
            V v[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int r = base + u * per;
                if (r < rows) v[u] = load(r, c);
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int r = base + u * per;
                if (r < rows) store(r, c, v[u]);
            }
        }
        return;
    }
    const int n = rows * cols;
    const float inv = 1.0f / (float)cols;
    using V = decltype(load(0, 0));
    for (int base = (int)threadIdx.x; base < n; base += NT * U) {
        V v[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = base + u * NT;
            if (t < n) {
                int r, c;
                divmod(t, cols, inv, r, c);
                v[u] = load(r, c);
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = base + u * NT;
            if (t < n) {
                int r, c;
                divmod(t, cols, inv, r, c);
                store(r, c, v[u]);
            }
        }
    }
}

LRE_DEV float aces(float x) {
    const float a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f;
    return (x * (a * x + b)) / (x * (c * x + d) + e);
}

// Grid-stride maximum over the green samples of an interleaved RGB image:
// m = max(m, g(G)) for every pixel, g applied per sample.  Aligned images are
// read as three float4 per four pixels, two groups per iteration.
template <class G>
LRE_DEV float green_max(const float *src, size_t n, float m, G g) {
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (size_t)gridDim.x * blockDim.x;
    size_t done = 0;
    if ((reinterpret_cast<uintptr_t>(src) & 15) == 0) {
        const float4 *s4 = reinterpret_cast<const float4 *>(src);
        const size_t n4 = n / 4;  // groups of four pixels
        for (size_t q = tid; q < n4; q += 2 * stride) {
            const size_t q2 = q + stride < n4 ? q + stride : q;
            const float4 a = s4[3 * q], b = s4[3 * q + 1], c = s4[3 * q + 2];
            const float4 d = s4[3 * q2], e = s4[3 * q2 + 1], f = s4[3 * q2 + 2];
            const float v[8] = {a.y, b.x, b.w, c.z, d.y, e.x, e.w, f.z};  // G of pixels 0-3 of both groups
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const float x = g(v[k]);
                m = x > m ? x : m;
            }
        }
        done = n4 * 4;
    }
    for (size_t i = done + tid; i < n; i += stride) {
        const float x = g(src[3 * i + 1]);
        m = x > m ? x : m;
    }
    return m;
}

// Maximum over the green samples of pixels [b, e) of an interleaved RGB
// image, by the threads of one workgroup: m = max(m, g(G)).
template <class G>
LRE_DEV float green_max_range(const float *src, size_t b, size_t e, float m, G g) {
    for (size_t i = b + threadIdx.x; i < e; i += blockDim.x) {
        const float x = g(src[3 * i + 1]);
        m = x > m ? x : m;
    }
    return m;
}

// Pixels of the frame of width `border` around a w x h image, enumerated as
// the full rows at the top, then at the bottom, then the left and right strips
// of the rows between (on images narrower than two strips a pixel may appear
// twice).  frame_pixels gives the count, frame_pixel the t-th position.
LRE_DEV __host__ int frame_pixels(int w, int h, int border) {
    const int top = border < h ? border : h;
    const int bot0 = border < h - border ? border : h - border, bot = bot0 > 0 ? bot0 : 0;
    const int mid = h - 2 * border > 0 ? h - 2 * border : 0, strip = border < w ? border : w;
    return (top + bot) * w + mid * 2 * strip;
}
LRE_DEV void frame_pixel(int t, int w, int h, int border, int &row, int &col) {
    const int top_n = min(border, h) * w;
    const int bot_rows = max(0, min(border, h - border)), bot_n = bot_rows * w;
    const int strip = min(border, w);
    if (t < top_n) {
        row = t / w;
        col = t % w;
    } else if (t < top_n + bot_n) {
        const int u = t - top_n;
        row = h - bot_rows + u / w;
        col = u % w;
    } else {
        const int u = t - top_n - bot_n, s = u % (2 * strip);
        row = border + u / (2 * strip);
        col = s < strip ? s : w - strip + (s - strip);
    }
}

// ---- one launch over many images ------------------------------------------
// A tiled kernel finds its image and tile through a source S: S.tile() gives
// (image, tx, ty) and S.args(t) the image's arguments.  One<A>: a single
// image, its arguments passed by value and the tile from the 2D grid.  Many<A>:
// a flat grid over several images; workgroup b belongs to the last image whose
// first workgroup first[i] is <= b (binary search; first[n] is the grid size)
// and takes tile b - first[i] of that image's row-major grid of tiles_x[i]
// columns.  The per-image arguments and both index arrays live in one device
// table.
struct TileOf {
    int image, tx, ty;
};
template <class A>
struct One {
    A a;
    LRE_DEV TileOf tile() const { return {0, (int)blockIdx.x, (int)blockIdx.y}; }
    LRE_DEV const A &args(const TileOf &) const { return a; }
};
template <class A>
struct Many {
    const A *a;
    const int *first;
    const int *tiles_x;
    int n;
    LRE_DEV TileOf tile() const { return tile_at((int)blockIdx.x); }
    // tile b of the flat list (persistent grids walk it themselves)
    LRE_DEV TileOf tile_at(int b) const {
        int lo = 0, hi = n - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (first[mid] <= b) lo = mid;
            else hi = mid - 1;
        }
        const int lt = b - first[lo], tx = tiles_x[lo];
        return {lo, lt % tx, lt / tx};
    }
    LRE_DEV const A &args(const TileOf &t) const { return a[t.image]; }
};

// Host side of a Many<A> table: the per-image arguments, then first[n + 1] and
// tiles_x[n], in one buffer.  fill() writes the host copy and returns the grid
// size; the device copy (same layout) is what view() points into.
template <class A>
struct ManyTable {
    static size_t bytes(int n) {
        return ((sizeof(A) * (size_t)n + 15) & ~(size_t)15) + sizeof(int) * (2 * (size_t)n + 1);
    }
    // tiles(i) -> (tiles_x, tiles_y) of image i; args(i) -> its arguments
    template <class TilesF, class ArgsF>
    static int fill(void *h, int n, TilesF tiles, ArgsF args) {
        A *a = static_cast<A *>(h);
        int *first = reinterpret_cast<int *>(static_cast<char *>(h) + ((sizeof(A) * (size_t)n + 15) & ~(size_t)15));
        int *tx = first + n + 1;
        int total = 0;
        for (int i = 0; i < n; i++) {
            a[i] = args(i);
            const int2 t = tiles(i);
            first[i] = total;
            tx[i] = t.x;
            total += t.x * t.y;
        }
        first[n] = total;
        return total;
    }
    static Many<A> view(const void *d, int n) {
        const char *c = static_cast<const char *>(d);
        const int *first = reinterpret_cast<const int *>(c + ((sizeof(A) * (size_t)n + 15) & ~(size_t)15));
        return Many<A>{reinterpret_cast<const A *>(c), first, first + n + 1, n};
    }
};

// Tile queue of a persistent grid: thread 0 takes the next tile with one
// atomic add on *counter (zeroed before the launch) and the workgroup reads it
// from LDS, so a workgroup that finishes early takes more tiles.  Each
// workgroup leaves once the counter has passed n; workgroups never wait for
// each other, so the grid may exceed what is resident at once.
template <class F>
LRE_DEV void tile_queue(unsigned *counter, int n, F f) {
    __shared__ int next;
    for (;;) {
        if (threadIdx.x == 0) next = (int)atomicAdd(counter, 1u);
        __syncthreads();
        const int t = next;
        __syncthreads();  // every wave has read `next` before thread 0 replaces it
        if (t >= n) break;
        f(t);
    }
}

}  // namespace lre

// Error plumbing of the C-ABI: every HIP call is checked, the message is kept
// per thread and the entry point returns a negative status.
void lre_set_error(const char *fmt, ...);

#define LRE_HIP_CHECK(expr)                                                                   \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess) {                                                               \
            lre_set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, \
                          __LINE__);                                                          \
            return LRE_ERR_HIP;                                                               \
        }                                                                                     \
    } while (0)
