// HIPAccelerator: forwards the reference backend's methods to the C ABI.
#include "hip_accelerator.h"

#include <iostream>

namespace libraw_enhanced {

namespace {
bool same_size(const ImageBufferFloat &a, const ImageBufferFloat &b) {
    return a.is_valid() && b.is_valid() && a.width == b.width && a.height == b.height;
}
const float *fp(const ImageBufferFloat &b) { return reinterpret_cast<const float *>(b.image); }
float *fp(ImageBufferFloat &b) { return reinterpret_cast<float *>(b.image); }
}  // namespace

HIPAccelerator::HIPAccelerator(int device) : device_(device) {}

HIPAccelerator::~HIPAccelerator() {
    if (ctx_) lre_hip_destroy(ctx_);
}

bool HIPAccelerator::initialize() {
    if (ctx_) return true;
    if (!lre_hip_available()) return false;
    if (lre_hip_create(&ctx_, device_) != LRE_OK) {
        ctx_ = nullptr;
        return false;
    }
    device_name_ = "AMD Instinct (HIP, gfx950)";
    return true;
}

bool HIPAccelerator::is_available() const { return ctx_ != nullptr; }
std::string HIPAccelerator::get_device_info() const { return device_name_; }
std::string HIPAccelerator::last_error() const { return lre_hip_last_error(); }

bool HIPAccelerator::demosaic_bayer_amaze(const ImageBufferFloat &raw, ImageBufferFloat &rgb,
                                          uint32_t filters, const float (&cam_mul)[4],
                                          float maximum_value) {
    if (!ctx_ || !same_size(raw, rgb)) return false;
    return lre_hip_amaze(ctx_, fp(raw), fp(rgb), (int)raw.width, (int)raw.height, filters, cam_mul,
                         maximum_value) == LRE_OK;
}

bool HIPAccelerator::demosaic_xtrans_3pass(const ImageBufferFloat &raw, ImageBufferFloat &rgb,
                                           const char (&xtrans)[6][6], const float (&)[3][4],
                                           float maximum_value) {
    if (!ctx_ || !same_size(raw, rgb)) return false;
    signed char xt[36];
    for (int i = 0; i < 36; i++) xt[i] = (signed char)xtrans[i / 6][i % 6];
    return lre_hip_xtrans3(ctx_, fp(raw), fp(rgb), (int)raw.width, (int)raw.height, xt,
                           maximum_value) == LRE_OK;
}

bool HIPAccelerator::apply_white_balance(const ImageBuffer &raw, ImageBufferFloat &rgb,
                                         const float wb[4], uint32_t filters,
                                         const char xtrans[6][6]) {
    if (!ctx_ || !raw.is_valid() || !rgb.is_valid() || raw.width != rgb.width ||
        raw.height != rgb.height || raw.channels != 4)
        return false;
    signed char xt[36] = {0};
    if (xtrans)
        for (int i = 0; i < 36; i++) xt[i] = (signed char)xtrans[i / 6][i % 6];
    return lre_hip_white_balance(ctx_, reinterpret_cast<const uint16_t *>(raw.image), fp(rgb),
                                 (int)raw.width, (int)raw.height, wb, filters, xt) == LRE_OK;
}

bool HIPAccelerator::tone_mapping(const ImageBufferFloat &in, ImageBufferFloat &out,
                                  float after_scale) {
    if (!ctx_ || !same_size(in, out)) return false;
    return lre_hip_tone_mapping(ctx_, fp(in), fp(out), (int)in.width, (int)in.height,
                                after_scale) == LRE_OK;
}

bool HIPAccelerator::apply_detail_preserving_tonemap(const ImageBufferFloat &in,
                                                     ImageBufferFloat &out) {
    if (!ctx_ || !same_size(in, out)) return false;
    return lre_hip_detail_tonemap(ctx_, fp(in), fp(out), (int)in.width, (int)in.height) == LRE_OK;
}

bool HIPAccelerator::enhance_micro_contrast(const ImageBufferFloat &in, ImageBufferFloat &out,
                                            float threshold, float strength, float target_contrast,
                                            const float *mask) {
    if (!ctx_ || !same_size(in, out)) return false;
    return lre_hip_micro_contrast(ctx_, fp(in), fp(out), (int)in.width, (int)in.height, threshold,
                                  strength, target_contrast, mask) == LRE_OK;
}

bool HIPAccelerator::defringe(const ImageBufferFloat &in, ImageBufferFloat &out, float radius,
                              float edge_threshold, float chroma_threshold, float strength,
                              bool defringe_green, float green_strength_scale) {
    if (!ctx_ || !same_size(in, out)) return false;
    // in.image == out.image selects the in-place semantics, as in the reference
    return lre_hip_defringe(ctx_, fp(in), fp(out), (int)in.width, (int)in.height, radius,
                            edge_threshold, chroma_threshold, strength, defringe_green ? 1 : 0,
                            green_strength_scale) == LRE_OK;
}

bool HIPAccelerator::ca_register_lateral(const ImageBufferFloat &in, ImageBufferFloat &out,
                                         int cell_size, int max_iterations, float max_shift,
                                         float min_confidence, int pyramid_levels) {
    if (!ctx_ || !same_size(in, out)) return false;
    return lre_hip_lateral_ca(ctx_, fp(in), fp(out), (int)in.width, (int)in.height, cell_size,
                              max_iterations, max_shift, min_confidence, pyramid_levels) == LRE_OK;
}

bool HIPAccelerator::convert_color_space(const ImageBufferFloat &in, ImageBufferFloat &out,
                                         const float transform[3][4]) {
    if (!ctx_ || !same_size(in, out) || !transform) return false;
    float t[12];
    for (int i = 0; i < 12; i++) t[i] = transform[i / 4][i % 4];
    return lre_hip_convert_color_space(ctx_, fp(in), fp(out), (int)in.width, (int)in.height, t) ==
           LRE_OK;
}

bool HIPAccelerator::gamma_correct(const ImageBufferFloat &in, ImageBufferFloat &out,
                                   float gamma_power, float gamma_slope, int output_color_space) {
    if (!ctx_ || !same_size(in, out)) return false;
    return lre_hip_gamma_correct(ctx_, fp(in), fp(out), (int)in.width, (int)in.height, gamma_power,
                                 gamma_slope, output_color_space) == LRE_OK;
}

bool HIPAccelerator::ca_axial_cleanup(const ImageBufferFloat &in, ImageBufferFloat &out,
                                      int radius, float epsilon, float strength) {
    if (!ctx_ || !same_size(in, out)) return false;
    return lre_hip_axial_ca(ctx_, fp(in), fp(out), (int)in.width, (int)in.height, radius, epsilon,
                            strength) == LRE_OK;
}

}  // namespace libraw_enhanced
