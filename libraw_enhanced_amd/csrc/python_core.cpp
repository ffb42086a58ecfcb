// pybind11 module `_core`: the standalone float32 methods of the reference's
// LibRawWrapper (core/python_bindings.cpp:320-351) with their Python
// signatures and defaults, running on the HIP backend through the C ABI
// (include/lre_hip.h).  Each method takes an (H, W, 3) array, converts it to
// C-contiguous float32 as pybind11's array_t<float> does, and returns a new
// array (core/libraw_wrapper.cpp:3063-3214).  A wrong shape raises
// LibRawInvalidArgument before any device work, as the reference's
// std::invalid_argument does (core/libraw_wrapper.cpp:3066-3069, :3100-3104,
// :3137-3141, :3181-3185; python_bindings.cpp:503-505).  There is no CPU path
// in this module: a HIP failure raises LibRawError.  The GIL is released while
// the device works.
//
// Not bound: the LibRaw decode methods (open_file, unpack, process, ...),
// which need LibRaw.  get_threshold() therefore returns the reference's
// fallback for "no raw loaded", 1 (compute_default_threshold,
// core/libraw_wrapper.cpp:3004-3011), and threshold < 0 takes that value.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include <stdexcept>
#include <string>

#include "lre_hip.h"

namespace py = pybind11;

namespace {

using FloatImage = py::array_t<float, py::array::c_style | py::array::forcecast>;

class LibRawWrapper {
  public:
    explicit LibRawWrapper(int device) : device_(device) {}
    ~LibRawWrapper() {
        if (ctx_) lre_hip_destroy(ctx_);
    }
    LibRawWrapper(const LibRawWrapper &) = delete;
    LibRawWrapper &operator=(const LibRawWrapper &) = delete;

    float get_threshold() const { return 1.0f; }

    py::array_t<float> tone_mapping(const FloatImage &image, float after_scale) {
        auto [h, w] = shape_of(image, "tone_mapping_numpy");
        py::array_t<float> out({h, w, (py::ssize_t)3});
        run([&](lre_ctx *c) { return lre_hip_tone_mapping(c, image.data(), out.mutable_data(), (int)w, (int)h, after_scale); });
        return out;
    }

    py::array_t<float> defringe(const FloatImage &image, float radius, float strength, bool defringe_green) {
        auto [h, w] = shape_of(image, "defringe_numpy");
        py::array_t<float> out({h, w, (py::ssize_t)3});
        // distinct input and output buffers: the standalone semantics (:3211)
        run([&](lre_ctx *c) {
            return lre_hip_defringe(c, image.data(), out.mutable_data(), (int)w, (int)h, radius, 0.1f, 0.15f,
                                    strength, defringe_green ? 1 : 0, 0.3f);
        });
        return out;
    }

    py::array_t<float> enhance_micro_contrast(const FloatImage &image, float threshold, float strength,
                                              float target_contrast) {
        auto [h, w] = shape_of(image, "enhance_micro_contrast_numpy");
        if (threshold < 0.f) threshold = get_threshold();
        py::array_t<float> out({h, w, (py::ssize_t)3});
        run([&](lre_ctx *c) {
            return lre_hip_micro_contrast(c, image.data(), out.mutable_data(), (int)w, (int)h, threshold,
                                          strength, target_contrast, nullptr);
        });
        return out;
    }

    py::array_t<float> recover_highlights(const FloatImage &image, float threshold) {
        auto [h, w] = shape_of(image, "recover_highlights_numpy");
        if (threshold < 0.f) threshold = get_threshold();
        py::array_t<float> out({h, w, (py::ssize_t)3});
        run([&](lre_ctx *c) {
            return lre_hip_recover_highlights(c, image.data(), out.mutable_data(), (int)w, (int)h, threshold,
                                              nullptr, nullptr, nullptr);
        });
        return out;
    }

  private:
    static std::pair<py::ssize_t, py::ssize_t> shape_of(const FloatImage &image, const char *what) {
        if (image.ndim() != 3 || image.shape(2) != 3)
            throw std::invalid_argument(std::string(what) + ": image must be shape (H, W, 3) float32");
        if (image.shape(0) < 1 || image.shape(1) < 1)
            throw std::invalid_argument(std::string(what) + ": image must not be empty");
        return {image.shape(0), image.shape(1)};
    }

    lre_ctx *ctx() {
        if (!ctx_) {
            const int rc = lre_hip_create(&ctx_, device_);
            if (rc != LRE_OK) {
                ctx_ = nullptr;
                throw std::runtime_error(std::string("HIP backend unavailable: ") + lre_hip_last_error());
            }
        }
        return ctx_;
    }

    // f(ctx) runs without the GIL; the C ABI does not throw
    template <class F>
    void run(F f) {
        lre_ctx *c = ctx();
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = f(c);
        }
        if (rc == LRE_ERR_ARG) throw std::invalid_argument(lre_hip_last_error());
        if (rc != LRE_OK) throw std::runtime_error(lre_hip_last_error());
    }

    int device_;
    lre_ctx *ctx_ = nullptr;
};

}  // namespace

PYBIND11_MODULE(_core, m) {
    m.doc() = "libraw_enhanced's standalone float32 methods on the MI355X HIP backend";
    py::class_<LibRawWrapper>(m, "LibRawWrapper")
        .def(py::init<int>(), py::arg("device") = 0)
        .def("get_threshold", &LibRawWrapper::get_threshold,
             "Get threshold value = maximum / data_maximum (1 when no RAW is loaded)")
        .def("recover_highlights", &LibRawWrapper::recover_highlights, py::arg("image"),
             py::arg("threshold") = -1.f,
             "Recover highlights on a (H,W,3) float32 numpy array. threshold<0 uses "
             "maximum/data_maximum automatically. Returns a new array.")
        .def("tone_mapping", &LibRawWrapper::tone_mapping, py::arg("image"), py::arg("after_scale") = 1.f,
             "Apply tone mapping on a (H,W,3) float32 numpy array. Returns a new array.")
        .def("enhance_micro_contrast", &LibRawWrapper::enhance_micro_contrast, py::arg("image"),
             py::arg("threshold") = -1.f, py::arg("strength") = 8.f, py::arg("target_contrast") = 0.06f,
             "Enhance micro contrast on a (H,W,3) float32 numpy array. threshold<0 uses "
             "maximum/data_maximum automatically. Returns a new array.")
        .def("defringe", &LibRawWrapper::defringe, py::arg("image"), py::arg("radius") = 10.0f,
             py::arg("strength") = 10.0f, py::arg("defringe_green") = false,
             "Remove chromatic aberration fringe on a linear (H,W,3) float32 numpy array. "
             "Returns a new float32 array.");
    m.def("hip_available", [] { return lre_hip_available() == 1; }, "Whether a HIP device is usable");
    py::register_exception<std::runtime_error>(m, "LibRawError");
    py::register_exception<std::invalid_argument>(m, "LibRawInvalidArgument");
    py::register_exception<std::out_of_range>(m, "LibRawOutOfRange");
}
