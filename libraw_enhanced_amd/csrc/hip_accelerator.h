// HIPAccelerator — the MI355X backend in the shape of the reference's
// CPUAccelerator / GPUAccelerator (core/cpu_accelerator.h:17-195,
// core/gpu_accelerator.h:18-130).  It sits in core/ next to them and is chosen
// by Accelerator::Impl (core/accelerator.cpp:15-24) when a HIP device is
// present.  Every method forwards to the C ABI in lre_hip.h and returns false
// on any failure, which makes the facade fall back to the CPU backend
// (core/accelerator.cpp:141-157).
//
// Built inside the reference tree: "accelerator.h" is the reference's header
// (ImageBuffer, ImageBufferFloat).
#pragma once

#include <cstdint>
#include <string>

#include "accelerator.h"
#include "lre_hip.h"

namespace libraw_enhanced {

class HIPAccelerator {
public:
    // device: the HIP device index the backend binds to in initialize()
    explicit HIPAccelerator(int device = 0);
    ~HIPAccelerator();
    HIPAccelerator(const HIPAccelerator &) = delete;
    HIPAccelerator &operator=(const HIPAccelerator &) = delete;

    bool initialize();
    bool is_available() const;
    std::string get_device_info() const;

    bool demosaic_bayer_amaze(const ImageBufferFloat &raw_buffer, ImageBufferFloat &rgb_buffer,
                              uint32_t filters, const float (&cam_mul)[4], float maximum_value);
    bool demosaic_xtrans_3pass(const ImageBufferFloat &raw_buffer, ImageBufferFloat &rgb_buffer,
                               const char (&xtrans)[6][6], const float (&color_matrix)[3][4],
                               float maximum_value);
    bool apply_white_balance(const ImageBuffer &raw_buffer, ImageBufferFloat &rgb_buffer,
                             const float wb_multipliers[4], uint32_t filters,
                             const char xtrans[6][6]);
    bool tone_mapping(const ImageBufferFloat &rgb_input, ImageBufferFloat &rgb_output,
                      float after_scale);
    // apply_detail_preserving_tonemap_gpu of the facade (core/accelerator.h:329-330)
    bool apply_detail_preserving_tonemap(const ImageBufferFloat &rgb_input,
                                         ImageBufferFloat &rgb_output);
    bool enhance_micro_contrast(const ImageBufferFloat &rgb_input, ImageBufferFloat &rgb_output,
                                float threshold, float strength, float target_contrast = 0.06f,
                                const float *mask = nullptr);
    bool defringe(const ImageBufferFloat &rgb_input, ImageBufferFloat &rgb_output,
                  float radius = 10.0f, float edge_threshold = 0.1f,
                  float chroma_threshold = 0.15f, float strength = 10.0f,
                  bool defringe_green = false, float green_strength_scale = 0.3f);
    bool ca_register_lateral(const ImageBufferFloat &rgb_input, ImageBufferFloat &rgb_output,
                             int cell_size = 96, int max_iterations = 3, float max_shift = 6.0f,
                             float min_confidence = 0.02f, int pyramid_levels = 3);
    bool convert_color_space(const ImageBufferFloat &rgb_input, ImageBufferFloat &rgb_output,
                             const float transform[3][4]);
    bool gamma_correct(const ImageBufferFloat &rgb_input, ImageBufferFloat &rgb_output,
                       float gamma_power = 0.f, float gamma_slope = 0.f,
                       int output_color_space = 1);
    bool ca_axial_cleanup(const ImageBufferFloat &rgb_input, ImageBufferFloat &rgb_output,
                          int radius = 6, float epsilon = 1e-4f, float strength = 0.3f);

    // the message of the last failed call
    std::string last_error() const;

private:
    int device_ = 0;
    lre_ctx *ctx_ = nullptr;
    std::string device_name_;
};

}  // namespace libraw_enhanced
