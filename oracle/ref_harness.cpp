// TEST INFRASTRUCTURE ONLY.  C-ABI glue around the reference's own CPU code
// (core/cpu_accelerator.cpp, built from the reference tree by ref_build.sh into
// oracle/_ref/).  Only tests/, the smoke check and bench.py's CPU-baseline leg
// may load the resulting library, and only as a checker or a timed baseline.
//
// Each entry point forwards to the CPUAccelerator method named beside it;
// buffers are caller-owned interleaved float RGB (H x W x 3) or uint16 x 4.
#include "cpu_accelerator.h"

#include <cstdint>
#include <cstring>
#include <iostream>
#include <omp.h>

using namespace libraw_enhanced;

namespace {
ImageBufferFloat view(const float *p, int w, int h) {
    ImageBufferFloat b;
    b.image = reinterpret_cast<float (*)[3]>(const_cast<float *>(p));
    b.width = (size_t)w;
    b.height = (size_t)h;
    b.channels = 3;
    return b;
}
CPUAccelerator &cpu() {
    static CPUAccelerator *acc = [] {
        std::cout.setstate(std::ios::failbit);  // the reference logs every stage
        auto *a = new CPUAccelerator();
        a->initialize();
        return a;
    }();
    return *acc;
}
}  // namespace

// Detail-preserving tone map (core/libraw_wrapper.cpp:1398-1471), compiled by
// ref_build.sh from the reference's own lines into ref_detail.o.
extern "C" void ref_detail_tm_impl(float *rgb, int w, int h);

extern "C" {

void ref_set_threads(int n) { omp_set_num_threads(n); }

// core/cpu_accelerator.cpp:1635 demosaic_bayer_amaze
int ref_amaze(const float *raw3, float *rgb3, int w, int h, uint32_t filters,
              const float *cam_mul4, float maximum_value) {
    float cm[4] = {cam_mul4[0], cam_mul4[1], cam_mul4[2], cam_mul4[3]};
    ImageBufferFloat in = view(raw3, w, h), out = view(rgb3, w, h);
    return cpu().demosaic_bayer_amaze(in, out, filters, cm, maximum_value) ? 0 : -1;
}

// core/cpu_accelerator.cpp:2776 demosaic_xtrans_3pass
int ref_xtrans3(const float *raw3, float *rgb3, int w, int h, const signed char *xtrans36,
                float maximum_value) {
    char xt[6][6];
    for (int i = 0; i < 36; i++) xt[i / 6][i % 6] = (char)xtrans36[i];
    float cm[3][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}};
    ImageBufferFloat in = view(raw3, w, h), out = view(rgb3, w, h);
    return cpu().demosaic_xtrans_3pass(in, out, xt, cm, maximum_value) ? 0 : -1;
}

// core/cpu_accelerator.cpp:2789 apply_white_balance
int ref_white_balance(const uint16_t *img4, float *rgb3, int w, int h, const float *wb4,
                      uint32_t filters, const signed char *xtrans36) {
    char xt[6][6] = {};
    if (xtrans36)
        for (int i = 0; i < 36; i++) xt[i / 6][i % 6] = (char)xtrans36[i];
    ImageBuffer in;
    in.image = reinterpret_cast<uint16_t (*)[4]>(const_cast<uint16_t *>(img4));
    in.width = (size_t)w;
    in.height = (size_t)h;
    in.channels = 4;
    ImageBufferFloat out = view(rgb3, w, h);
    return cpu().apply_white_balance(in, out, wb4, filters, xt) ? 0 : -1;
}

// core/cpu_accelerator.cpp:3034 tone_mapping
int ref_tone_mapping(const float *in3, float *out3, int w, int h, float after_scale) {
    ImageBufferFloat in = view(in3, w, h), out = view(out3, w, h);
    return cpu().tone_mapping(in, out, after_scale) ? 0 : -1;
}

// core/libraw_wrapper.cpp:1398 apply_detail_preserving_tonemap (in place)
int ref_detail_tonemap(float *rgb3, int w, int h) {
    cpu();
    ref_detail_tm_impl(rgb3, w, h);
    return 0;
}

// core/cpu_accelerator.cpp:3939 defringe (in3 == out3 selects the in-place form)
int ref_defringe(const float *in3, float *out3, int w, int h, float radius, float edge_threshold,
                 float chroma_threshold, float strength, int defringe_green,
                 float green_strength_scale) {
    ImageBufferFloat in = view(in3, w, h), out = view(out3, w, h);
    return cpu().defringe(in, out, radius, edge_threshold, chroma_threshold, strength,
                          defringe_green != 0, green_strength_scale)
               ? 0
               : -1;
}

// core/cpu_accelerator.cpp:3496 ca_register_lateral
int ref_lateral_ca(const float *in3, float *out3, int w, int h, int cell_size, int max_iterations,
                   float max_shift, float min_confidence, int pyramid_levels) {
    ImageBufferFloat in = view(in3, w, h), out = view(out3, w, h);
    return cpu().ca_register_lateral(in, out, cell_size, max_iterations, max_shift,
                                     min_confidence, pyramid_levels)
               ? 0
               : -1;
}

// core/cpu_accelerator.cpp:3799 ca_axial_cleanup
int ref_axial_ca(const float *in3, float *out3, int w, int h, int radius, float epsilon,
                 float strength) {
    ImageBufferFloat in = view(in3, w, h), out = view(out3, w, h);
    return cpu().ca_axial_cleanup(in, out, radius, epsilon, strength) ? 0 : -1;
}

}  // extern "C"

extern "C" {

// core/cpu_accelerator.cpp:2874 convert_color_space
int ref_convert_color_space(const float *in3, float *out3, int w, int h, const float *transform12) {
    float t[3][4];
    for (int i = 0; i < 12; i++) t[i / 4][i % 4] = transform12[i];
    ImageBufferFloat in = view(in3, w, h), out = view(out3, w, h);
    return cpu().convert_color_space(in, out, t) ? 0 : -1;
}

// core/cpu_accelerator.cpp:2903 gamma_correct
int ref_gamma_correct(const float *in3, float *out3, int w, int h, float gamma_power,
                      float gamma_slope, int output_color_space) {
    ImageBufferFloat in = view(in3, w, h), out = view(out3, w, h);
    return cpu().gamma_correct(in, out, gamma_power, gamma_slope, output_color_space) ? 0 : -1;
}

}  // extern "C"
