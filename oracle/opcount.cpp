// TEST INFRASTRUCTURE ONLY.  FP32 operation counts of the restated demosaics
// (built with -fpermissive: the C files convert malloc results implicitly):
// amaze.c and xtrans.c (the oracle's restatements of
// core/cpu_accelerator.cpp:1145-1653 and :1870-2676) compiled as C++ with
// every float replaced by a type that counts the operations done on it.  The
// counts are per output pixel of a seeded scene; they are the algorithmic
// FP32 work of the two stages (no index or address arithmetic), reported next
// to the kernels' measured VALU instruction rates.  Single-threaded.
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

namespace opc {
struct Counts {
    unsigned long long add = 0, mul = 0, div = 0, cmp = 0, abs = 0;
};
Counts n;

struct F {
    float v;
    F() = default;
    constexpr F(float x) : v(x) {}
    constexpr F(double x) : v((float)x) {}
    constexpr F(int x) : v((float)x) {}
    constexpr F(unsigned x) : v((float)x) {}
    explicit operator int() const { return (int)v; }
    explicit operator unsigned() const { return (unsigned)v; }
    explicit operator double() const { return (double)v; }
    F &operator+=(F o) { n.add++; v += o.v; return *this; }
    F &operator-=(F o) { n.add++; v -= o.v; return *this; }
    F &operator*=(F o) { n.mul++; v *= o.v; return *this; }
    F &operator/=(F o) { n.div++; v /= o.v; return *this; }
    F operator-() const { return F(-v); }
    F &operator++() { n.add++; v += 1.f; return *this; }
    F operator++(int) { F t = *this; n.add++; v += 1.f; return t; }
};
inline F operator+(F a, F b) { n.add++; return F(a.v + b.v); }
inline F operator-(F a, F b) { n.add++; return F(a.v - b.v); }
inline F operator*(F a, F b) { n.mul++; return F(a.v * b.v); }
inline F operator/(F a, F b) { n.div++; return F(a.v / b.v); }
inline bool operator<(F a, F b) { n.cmp++; return a.v < b.v; }
inline bool operator>(F a, F b) { n.cmp++; return a.v > b.v; }
inline bool operator<=(F a, F b) { n.cmp++; return a.v <= b.v; }
inline bool operator>=(F a, F b) { n.cmp++; return a.v >= b.v; }
inline bool operator==(F a, F b) { n.cmp++; return a.v == b.v; }
inline bool operator!=(F a, F b) { n.cmp++; return a.v != b.v; }
inline F fabsf(F a) { n.abs++; return F(::fabsf(a.v)); }
}  // namespace opc
using opc::fabsf;

#define float opc::F
namespace amz {
#include "amaze.c"
}
#undef LRE_ORACLE_H
namespace xtr {
#include "xtrans.c"
}
#undef float

namespace {
void scene_cfa(float *raw3, int w, int h, uint64_t seed, bool xtrans, const signed char *xt) {
    // smooth gradient, a few hard-edged discs and a zone plate, sampled through the CFA
    uint64_t s = seed;
    auto rnd = [&]() {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return (double)((z ^ (z >> 31)) >> 11) / 9007199254740992.0;
    };
    double a[3], b[3], cx[20], cy[20], rr[20], col[20][3];
    for (int c = 0; c < 3; c++) { a[c] = rnd(); b[c] = rnd(); }
    for (int k = 0; k < 20; k++) {
        cx[k] = rnd() * w; cy[k] = rnd() * h; rr[k] = (0.02 + 0.1 * rnd()) * (w < h ? w : h);
        for (int c = 0; c < 3; c++) col[k][c] = 0.05 + 0.9 * rnd();
    }
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            double v[3];
            for (int c = 0; c < 3; c++) v[c] = 0.15 + 0.6 * (a[c] * x / w + b[c] * y / h) * 0.5;
            for (int k = 0; k < 20; k++)
                if ((x - cx[k]) * (x - cx[k]) + (y - cy[k]) * (y - cy[k]) < rr[k] * rr[k])
                    for (int c = 0; c < 3; c++) v[c] = col[k][c];
            const double dx = x - w / 2.0, dy = y - h / 2.0, r2 = dx * dx + dy * dy, R = (w < h ? w : h) / 2.0;
            if (r2 < R * R)
                for (int c = 0; c < 3; c++) v[c] = 0.5 * v[c] + 0.25 * (1 + cos(M_PI * 0.45 * r2 / R));
            const int f = xtrans ? xt[(y % 6) * 6 + x % 6] : ((y & 1) ? ((x & 1) ? 2 : 1) : ((x & 1) ? 1 : 0));
            float *p = raw3 + 3 * ((size_t)y * w + x);
            p[0] = p[1] = p[2] = 0.f;
            p[f] = (float)(v[f] + 0.002 * (rnd() - 0.5));
        }
}

void report(const char *name, int w, int h) {
    const double px = (double)w * h;
    const opc::Counts &c = opc::n;
    printf("{\"stage\": \"%s\", \"width\": %d, \"height\": %d, \"add_sub\": %.1f, \"mul\": %.1f, \"div\": %.1f, "
           "\"compare\": %.1f, \"abs\": %.1f, \"fp32_arith_per_px\": %.1f, \"fp32_ops_incl_compare_per_px\": %.1f}\n",
           name, w, h, c.add / px, c.mul / px, c.div / px, c.cmp / px, c.abs / px, (c.add + c.mul + c.div) / px,
           (c.add + c.mul + c.div + c.cmp + c.abs) / px);
}
}  // namespace

int main(int argc, char **argv) {
    const int w = argc > 2 ? atoi(argv[2]) : 1536, h = argc > 3 ? atoi(argv[3]) : 1024;
    omp_set_num_threads(1);
    float *raw3 = (float *)malloc(sizeof(float) * 3 * (size_t)w * h);
    opc::F *in = (opc::F *)raw3;
    opc::F *out = (opc::F *)calloc((size_t)w * h * 3, sizeof(float));
    static const signed char XT[36] = {1, 1, 0, 1, 1, 2, 1, 1, 2, 1, 1, 0, 2, 0, 1, 0, 2, 1,
                                       1, 1, 2, 1, 1, 0, 1, 1, 0, 1, 1, 2, 0, 2, 1, 2, 0, 1};
    const bool want_xtrans = argc > 1 && argv[1][0] == 'x';
    scene_cfa(raw3, w, h, 20261203ull, want_xtrans, XT);
    opc::F cm[4] = {1.f, 1.f, 1.f, 1.f};
    opc::n = opc::Counts();
    if (want_xtrans) {
        if (xtr::oracle_xtrans3(in, out, w, h, XT, opc::F(1.0f), 0) != 0) return 1;
        report("xtrans3", w, h);
    } else {
        if (amz::oracle_amaze(in, out, w, h, 0x94949494u, cm, opc::F(1.0f), 0) != 0) return 1;
        report("amaze", w, h);
    }
    free(raw3);
    free(out);
    return 0;
}
