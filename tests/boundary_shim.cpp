// TEST INFRASTRUCTURE ONLY.  The reference's facade pattern around the HIP
// backend, for the boundary tests: a HIPAccelerator (the class a maintainer
// adds to core/) beside the reference's own CPUAccelerator, and per method the
// selection of core/accelerator.cpp:141-157 (try the GPU backend; on false, run
// the CPU backend).  Built by oracle/boundary_build.sh against the reference's
// headers and its compiled CPU code (oracle/_ref); loaded by tests/ only.
//
// Each boundary_<method> returns 1 when the HIP backend served the call, 0
// when it returned false and the CPU backend did, -1 when both failed.
#include <cstdint>
#include <iostream>

#include "cpu_accelerator.h"
#include "hip_accelerator.h"

using namespace libraw_enhanced;

namespace {
struct Facade {
    HIPAccelerator hip;
    CPUAccelerator cpu;
    bool hip_ok;
    explicit Facade(int device) : hip(device) {
        hip_ok = hip.initialize();
        cpu.initialize();
    }
};
ImageBufferFloat view(const float *p, int w, int h) {
    ImageBufferFloat b;
    b.image = reinterpret_cast<float (*)[3]>(const_cast<float *>(p));
    b.width = (size_t)w;
    b.height = (size_t)h;
    b.channels = 3;
    return b;
}
template <class G, class C>
int select(Facade *f, G gpu, C cpu) {
    if (f->hip_ok && gpu()) return 1;
    return cpu() ? 0 : -1;
}
}  // namespace

extern "C" {

void *boundary_create(int device) {
    std::cout.setstate(std::ios::failbit);  // the reference's CPU code logs every stage
    return new Facade(device);
}
void boundary_destroy(void *f) { delete static_cast<Facade *>(f); }
int boundary_hip_ok(void *f) { return static_cast<Facade *>(f)->hip_ok ? 1 : 0; }

int boundary_amaze(void *fp, const float *raw3, float *rgb3, int w, int h, uint32_t filters, const float *cm4,
                   float maximum_value) {
    auto *f = static_cast<Facade *>(fp);
    ImageBufferFloat in = view(raw3, w, h), out = view(rgb3, w, h);
    const float cm[4] = {cm4[0], cm4[1], cm4[2], cm4[3]};
    return select(f, [&] { return f->hip.demosaic_bayer_amaze(in, out, filters, cm, maximum_value); },
                  [&] { return f->cpu.demosaic_bayer_amaze(in, out, filters, cm, maximum_value); });
}

int boundary_defringe(void *fp, const float *in3, float *out3, int w, int h, float radius, float strength,
                      int green) {
    auto *f = static_cast<Facade *>(fp);
    ImageBufferFloat in = view(in3, w, h), out = view(out3, w, h);
    return select(f, [&] { return f->hip.defringe(in, out, radius, 0.1f, 0.15f, strength, green != 0, 0.3f); },
                  [&] { return f->cpu.defringe(in, out, radius, 0.1f, 0.15f, strength, green != 0, 0.3f); });
}

int boundary_axial(void *fp, const float *in3, float *out3, int w, int h, int radius, float eps, float strength) {
    auto *f = static_cast<Facade *>(fp);
    ImageBufferFloat in = view(in3, w, h), out = view(out3, w, h);
    return select(f, [&] { return f->hip.ca_axial_cleanup(in, out, radius, eps, strength); },
                  [&] { return f->cpu.ca_axial_cleanup(in, out, radius, eps, strength); });
}

int boundary_tone_mapping(void *fp, const float *in3, float *out3, int w, int h, float after_scale) {
    auto *f = static_cast<Facade *>(fp);
    ImageBufferFloat in = view(in3, w, h), out = view(out3, w, h);
    return select(f, [&] { return f->hip.tone_mapping(in, out, after_scale); },
                  [&] { return f->cpu.tone_mapping(in, out, after_scale); });
}

}  // extern "C"
