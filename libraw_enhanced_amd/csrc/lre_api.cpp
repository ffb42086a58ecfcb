// C ABI of the HIP develop path: context management, host-pointer entry points
// (copy in, run, copy out) and device-resident entry points (enqueue only).
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <atomic>
#include <mutex>
#include <limits>
#include <string>

#include "lre_common.h"
#include "lre_internal.h"

namespace {
thread_local std::string g_last_error;

bool valid_dims(int w, int h) { return w > 0 && h > 0 && (size_t)w * h < ((size_t)1 << 31); }

int set_arg_error(const char *what) {
    lre_set_error("invalid argument: %s", what);
    return LRE_ERR_ARG;
}

int ensure_pinned(lre_ctx *ctx, size_t bytes) {
    if (ctx->pinned_bytes >= bytes) return LRE_OK;
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    ctx->pinned = nullptr;
    ctx->pinned_bytes = 0;
    LRE_HIP_CHECK(hipHostMalloc(&ctx->pinned, bytes, hipHostMallocDefault));
    ctx->pinned_bytes = bytes;
    return LRE_OK;
}

// Device buffers for one host-pointer call (kept in the context's scratch).
struct Carve {
    char *base;
    size_t off = 0;
    explicit Carve(void *b) : base(static_cast<char *>(b)) {}
    template <class T>
    T *take(size_t count) {
        T *p = reinterpret_cast<T *>(base + off);
        off += (count * sizeof(T) + 255) & ~(size_t)255;
        return p;
    }
};

size_t aligned(size_t b) { return (b + 255) & ~(size_t)255; }

int copy_in(lre_ctx *ctx, void *dst, const void *src, size_t bytes) {
    int rc = ensure_pinned(ctx, bytes);
    if (rc) return rc;
    std::memcpy(ctx->pinned, src, bytes);
    LRE_HIP_CHECK(hipMemcpyAsync(dst, ctx->pinned, bytes, hipMemcpyHostToDevice, ctx->stream));
    // the staging buffer is reused by the next copy
    LRE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return LRE_OK;
}

int copy_out(lre_ctx *ctx, void *dst, const void *src, size_t bytes) {
    int rc = ensure_pinned(ctx, bytes);
    if (rc) return rc;
    LRE_HIP_CHECK(hipMemcpyAsync(ctx->pinned, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    LRE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    std::memcpy(dst, ctx->pinned, bytes);
    return LRE_OK;
}
// host-pointer form of an RGB -> RGB stage: copy in, run, copy out
template <class F>
int rgb_stage_host(lre_ctx *ctx, const float *in_rgb, float *out_rgb, int w, int h,
                          size_t stage_scratch, F launch) {
    if (!ctx || !in_rgb || !out_rgb) return set_arg_error("null pointer");
    if (!valid_dims(w, h)) return set_arg_error("image size");
    LRE_HIP_CHECK(hipSetDevice(ctx->device));
    const size_t n = (size_t)w * h;
    int rc = lre_ensure_scratch(ctx, aligned(n * 12) * 2 + stage_scratch, ctx->stream);
    if (rc) return rc;
    lre_scratch_guard scratch_use{ctx, ctx->stream};
    Carve c(ctx->scratch);
    float *d_in = c.take<float>(n * 3), *d_out = c.take<float>(n * 3);
    void *ws = static_cast<char *>(ctx->scratch) + c.off;
    if ((rc = copy_in(ctx, d_in, in_rgb, n * 12))) return rc;
    if ((rc = launch(d_in, d_out, ws))) return rc;
    return copy_out(ctx, out_rgb, d_out, n * 12);
}

// This device's address of the no-alloc AMaZE arenas, looked up once per
// device (the caller has made that device current).
void *noalloc_arenas_for(int device) {
    constexpr int MAX_DEVICES = 64;
    static std::mutex mu;
    static void *addr[MAX_DEVICES] = {};
    if (device < 0 || device >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[device] && lre_amaze_noalloc_setup(&addr[device]) != LRE_OK) addr[device] = nullptr;
    return addr[device];
}

}  // namespace

int lre_current_cus() {
    constexpr int MAXD = 64;
    static std::atomic<int> cus[MAXD] = {};
    int d = 0;
    if (hipGetDevice(&d) != hipSuccess || d < 0 || d >= MAXD) return 256;
    int c = cus[d].load(std::memory_order_relaxed);
    if (c < 1) {
        if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, d) != hipSuccess || c < 1) c = 256;
        cus[d].store(c, std::memory_order_relaxed);
    }
    return c;
}

void lre_set_error(const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

void lre_scratch_release(lre_ctx *ctx, hipStream_t s) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    ctx->scratch_pending = hipStreamIsCapturing(s, &st) == hipSuccess && st == hipStreamCaptureStatusNone &&
                           hipEventRecord(ctx->scratch_done, s) == hipSuccess;
    ctx->scratch_stream = s;
}

int lre_ensure_scratch(lre_ctx *ctx, size_t bytes, hipStream_t s) {
    if (ctx->scratch_stream != s && ctx->scratch_pending) {
        // the previous call ran on another stream and may still use the scratch
        LRE_HIP_CHECK(hipEventSynchronize(ctx->scratch_done));
        ctx->scratch_pending = false;
    }
    if (ctx->scratch_bytes >= bytes) return LRE_OK;
    if (ctx->scratch) {
        // work on any stream may still use the old scratch
        LRE_HIP_CHECK(hipDeviceSynchronize());
        (void)hipFree(ctx->scratch);
    }
    ctx->scratch = nullptr;
    ctx->scratch_bytes = 0;
    hipError_t e = hipMalloc(&ctx->scratch, bytes);
    if (e != hipSuccess) {
        lre_set_error("device scratch allocation of %zu bytes failed: %s", bytes,
                      hipGetErrorString(e));
        return LRE_ERR_NOMEM;
    }
    ctx->scratch_bytes = bytes;
    return LRE_OK;
}

extern "C" {

const char *lre_hip_last_error(void) { return g_last_error.c_str(); }

int lre_hip_available(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n > 0 ? 1 : 0;
}

int lre_hip_create(lre_ctx **out, int device) {
    if (!out) return set_arg_error("ctx");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        lre_set_error("no HIP device");
        return LRE_ERR_NODEV;
    }
    if (device < 0 || device >= n) return set_arg_error("device index");
    LRE_HIP_CHECK(hipSetDevice(device));
    auto *ctx = new lre_ctx();
    ctx->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->num_cus = prop.multiProcessorCount;
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->scratch_done, hipEventDisableTiming);
    if (e != hipSuccess) {
        if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        lre_set_error("hipStreamCreate failed: %s", hipGetErrorString(e));
        return LRE_ERR_HIP;
    }
    // set-up of the no-alloc entry point; the context works without it
    ctx->noalloc_arenas = noalloc_arenas_for(device);
    *out = ctx;
    return LRE_OK;
}

int lre_hip_destroy(lre_ctx *ctx) {
    if (!ctx) return LRE_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    if (ctx->scratch_done) (void)hipEventDestroy(ctx->scratch_done);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return LRE_OK;
}

int lre_hip_synchronize(lre_ctx *ctx) {
    if (!ctx) return set_arg_error("ctx");
    LRE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return LRE_OK;
}

void *lre_hip_stream(lre_ctx *ctx) { return ctx ? ctx->stream : nullptr; }

int lre_hip_dev_malloc(lre_ctx *ctx, void **d_ptr, size_t bytes) {
    if (!ctx || !d_ptr) return set_arg_error("null pointer");
    LRE_HIP_CHECK(hipSetDevice(ctx->device));
    hipError_t e = hipMalloc(d_ptr, bytes ? bytes : 1);
    if (e != hipSuccess) {
        lre_set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
        return LRE_ERR_NOMEM;
    }
    return LRE_OK;
}

int lre_hip_dev_free(lre_ctx *ctx, void *d_ptr) {
    if (!ctx) return set_arg_error("ctx");
    if (!d_ptr) return LRE_OK;
    LRE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    LRE_HIP_CHECK(hipFree(d_ptr));
    return LRE_OK;
}

int lre_hip_memcpy_htod(lre_ctx *ctx, void *d_dst, const void *h_src, size_t bytes) {
    if (!ctx || !d_dst || !h_src) return set_arg_error("null pointer");
    LRE_HIP_CHECK(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream));
    LRE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return LRE_OK;
}

int lre_hip_memcpy_dtoh(lre_ctx *ctx, void *h_dst, const void *d_src, size_t bytes) {
    if (!ctx || !h_dst || !d_src) return set_arg_error("null pointer");
    LRE_HIP_CHECK(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    LRE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return LRE_OK;
}

// ---------------- device-resident entry points ----------------

int lre_hip_dev_amaze(lre_ctx *ctx, const float *d_cfa, float *d_rgb, int w, int h,
                      uint32_t filters, const float cam_mul[4], float maximum_value,
                      void *stream) {
    if (!ctx || !d_cfa || !d_rgb || !cam_mul) return set_arg_error("null pointer");
    if (!valid_dims(w, h)) return set_arg_error("image size");
    if (!(maximum_value > 0.f)) return set_arg_error("maximum_value must be > 0");
    if (filters == 9) return set_arg_error("AMaZE needs a Bayer pattern");
    int rc = lre_ensure_scratch(ctx, lre_amaze_scratch_bytes(ctx, w, h), lre_pick_stream(ctx, stream));
    if (rc) return rc;
    lre_scratch_guard scratch_use{ctx, lre_pick_stream(ctx, stream)};
    return lre_launch_amaze(ctx, d_cfa, d_rgb, w, h, filters, cam_mul, maximum_value,
                            ctx->scratch, lre_pick_stream(ctx, stream));
}

int lre_hip_dev_amaze_noalloc(lre_ctx *ctx, const float *d_cfa, float *d_rgb, int w, int h,
                              uint32_t filters, const float cam_mul[4], float maximum_value,
                              void *stream) {
    if (!ctx || !d_cfa || !d_rgb || !cam_mul) return set_arg_error("null pointer");
    if (!valid_dims(w, h)) return set_arg_error("image size");
    if (!(maximum_value > 0.f)) return set_arg_error("maximum_value must be > 0");
    if (filters == 9) return set_arg_error("AMaZE needs a Bayer pattern");
    if (!ctx->noalloc_arenas) {
        lre_set_error("the no-alloc AMaZE arenas are not available on device %d", ctx->device);
        return LRE_ERR_HIP;
    }
    LRE_HIP_CHECK(hipSetDevice(ctx->device));
    return lre_launch_amaze_noalloc(d_cfa, d_rgb, w, h, filters, cam_mul, maximum_value,
                                    ctx->noalloc_arenas, lre_pick_stream(ctx, stream));
}

int lre_hip_dev_xtrans3(lre_ctx *ctx, const float *d_cfa, float *d_rgb, int w, int h,
                        const signed char xtrans[36], float maximum_value, void *stream) {
    if (!ctx || !d_cfa || !d_rgb || !xtrans) return set_arg_error("null pointer");
    if (!valid_dims(w, h)) return set_arg_error("image size");
    if (!(maximum_value > 0.f)) return set_arg_error("maximum_value must be > 0");
    int rc = lre_ensure_scratch(ctx, lre_xtrans_scratch_bytes(ctx, w, h), lre_pick_stream(ctx, stream));
    if (rc) return rc;
    lre_scratch_guard scratch_use{ctx, lre_pick_stream(ctx, stream)};
    return lre_launch_xtrans3(ctx, d_cfa, d_rgb, w, h, xtrans, maximum_value, ctx->scratch,
                              lre_pick_stream(ctx, stream));
}

int lre_hip_dev_tone_mapping(lre_ctx *ctx, const float *d_in, float *d_out, int w, int h,
                             float after_scale, void *stream) {
    if (!ctx || !d_in || !d_out) return set_arg_error("null pointer");
    if (!valid_dims(w, h)) return set_arg_error("image size");
    return lre_launch_tone_mapping(d_in, d_out, (size_t)w * h * 3, after_scale,
                                   lre_pick_stream(ctx, stream));
}

int lre_hip_dev_detail_tonemap(lre_ctx *ctx, const float *d_in, float *d_out, int w, int h,
                               void *stream) {
    if (!ctx || !d_in || !d_out) return set_arg_error("null pointer");
    if (!valid_dims(w, h)) return set_arg_error("image size");
    // with a separate output the guide is read from the input: no scratch
    if (d_in != d_out) return lre_launch_detail_tonemap(d_in, d_out, w, h, nullptr, lre_pick_stream(ctx, stream));
    int rc = lre_ensure_scratch(ctx, lre_detail_tm_scratch_bytes(w, h), lre_pick_stream(ctx, stream));
    if (rc) return rc;
    lre_scratch_guard scratch_use{ctx, lre_pick_stream(ctx, stream)};
    return lre_launch_detail_tonemap(d_in, d_out, w, h, ctx->scratch, lre_pick_stream(ctx, stream));
}

int lre_hip_dev_micro_contrast(lre_ctx *ctx, const float *d_in, float *d_out, int w, int h,
                               float threshold, float strength, float target_contrast,
                               const float *d_mask, void *stream) {
    if (!ctx || !d_in || !d_out) return set_arg_error("null pointer");
    if (!valid_dims(w, h)) return set_arg_error("image size");
    int rc = lre_ensure_scratch(ctx, lre_micro_contrast_scratch_bytes(w, h), lre_pick_stream(ctx, stream));
    if (rc) return rc;
    lre_scratch_guard scratch_use{ctx, lre_pick_stream(ctx, stream)};
    return lre_launch_micro_contrast(d_in, d_out, w, h, threshold, strength, target_contrast, d_mask,
                                     ctx->scratch, lre_pick_stream(ctx, stream));
}

int lre_hip_dev_recover_highlights(lre_ctx *ctx, const float *d_in, float *d_out, int w, int h,
                                   float saturation_threshold, float *d_mask, const float *wb_rgb,
                                   float *d_max_val, void *stream) {
    if (!ctx || !d_in || !d_out) return set_arg_error("null pointer");
    if (!valid_dims(w, h)) return set_arg_error("image size");
    const size_t n = (size_t)w * h;
    const bool inplace = d_in == d_out;
    int rc = lre_ensure_scratch(ctx, (inplace ? aligned(n * 12) : 0) + lre_highlight_scratch_bytes(w, h), lre_pick_stream(ctx, stream));
    if (rc) return rc;
    lre_scratch_guard scratch_use{ctx, lre_pick_stream(ctx, stream)};
    hipStream_t s = lre_pick_stream(ctx, stream);
    Carve c(ctx->scratch);
    const float *src = d_in;
    if (inplace) {
        float *copy = c.take<float>(n * 3);
        LRE_HIP_CHECK(hipMemcpyAsync(copy, d_in, n * 12, hipMemcpyDeviceToDevice, s));
        src = copy;
    }
    return lre_launch_recover_highlights(src, d_out, w, h, saturation_threshold, d_mask, wb_rgb, d_max_val,
                                         static_cast<char *>(ctx->scratch) + c.off, s);
}

int lre_hip_dev_defringe(lre_ctx *ctx, const float *d_in, float *d_out, int w, int h,
                         float radius, float edge_threshold, float chroma_threshold,
                         float strength, int defringe_green, float green_strength_scale,
                         void *stream) {
    if (!ctx || !d_in || !d_out) return set_arg_error("null pointer");
    if (!valid_dims(w, h)) return set_arg_error("image size");
    int rc = lre_ensure_scratch(ctx, lre_defringe_scratch_bytes(w, h), lre_pick_stream(ctx, stream));
    if (rc) return rc;
    lre_scratch_guard scratch_use{ctx, lre_pick_stream(ctx, stream)};
    return lre_launch_defringe(d_in, d_out, w, h, radius, edge_threshold, chroma_threshold,
                               strength, defringe_green, green_strength_scale, ctx->scratch,
                               lre_pick_stream(ctx, stream));
}

int lre_hip_dev_axial_ca(lre_ctx *ctx, const float *d_in, float *d_out, int w, int h, int radius,
                         float epsilon, float strength, void *stream) {
    if (!ctx || !d_in || !d_out) return set_arg_error("null pointer");
    if (!valid_dims(w, h)) return set_arg_error("image size");
    int rc = lre_ensure_scratch(ctx, lre_axial_scratch_bytes(w, h), lre_pick_stream(ctx, stream));
    if (rc) return rc;
    lre_scratch_guard scratch_use{ctx, lre_pick_stream(ctx, stream)};
    return lre_launch_axial_ca(d_in, d_out, w, h, radius, epsilon, strength, ctx->scratch,
                               lre_pick_stream(ctx, stream));
}

int lre_hip_dev_lateral_ca(lre_ctx *ctx, const float *d_in, float *d_out, int w, int h,
                           int cell_size, int max_iterations, float max_shift,
                           float min_confidence, int pyramid_levels, void *stream) {
    if (!ctx || !d_in || !d_out) return set_arg_error("null pointer");
    if (!valid_dims(w, h)) return set_arg_error("image size");
    int rc = lre_ensure_scratch(ctx, lre_lateral_scratch_bytes(w, h, cell_size), lre_pick_stream(ctx, stream));
    if (rc) return rc;
    lre_scratch_guard scratch_use{ctx, lre_pick_stream(ctx, stream)};
    return lre_launch_lateral_ca(d_in, d_out, w, h, cell_size, max_iterations, max_shift,
                                 min_confidence, pyramid_levels, ctx->scratch,
                                 lre_pick_stream(ctx, stream));
}

int lre_hip_dev_ingest_u16(lre_ctx *ctx, const uint16_t *d_raw, float *d_cfa, int w, int h,
                           uint32_t filters, const signed char xtrans[36], const int cblack[4],
                           int black, int maximum, const float wb[4], float *d_data_max,
                           void *stream) {
    if (!ctx || !d_raw || !d_cfa || !wb) return set_arg_error("null pointer");
    if (!valid_dims(w, h)) return set_arg_error("image size");
    if (filters == 9 && !xtrans) return set_arg_error("X-Trans table");
    return lre_launch_ingest_u16(d_raw, d_cfa, w, h, filters, xtrans, cblack, black, maximum, wb,
                                 d_data_max, lre_pick_stream(ctx, stream));
}

int lre_hip_dev_rgb3_to_cfa(lre_ctx *ctx, const float *d_raw_rgb, float *d_cfa, int w, int h,
                            uint32_t filters, const signed char xtrans[36], void *stream) {
    if (!ctx || !d_raw_rgb || !d_cfa) return set_arg_error("null pointer");
    if (!valid_dims(w, h)) return set_arg_error("image size");
    if (filters == 9 && !xtrans) return set_arg_error("X-Trans table");
    return lre_launch_rgb3_to_cfa(d_raw_rgb, d_cfa, w, h, filters, xtrans,
                                  lre_pick_stream(ctx, stream));
}

// ---------------- host-pointer entry points ----------------
// Device buffers live at the front of the context scratch; the stage's own
// scratch follows them.

int lre_hip_amaze(lre_ctx *ctx, const float *raw_rgb, float *rgb, int w, int h,
                  uint32_t filters, const float cam_mul[4], float maximum_value) {
    if (!ctx || !raw_rgb || !rgb || !cam_mul) return set_arg_error("null pointer");
    if (!valid_dims(w, h)) return set_arg_error("image size");
    if (!(maximum_value > 0.f)) return set_arg_error("maximum_value must be > 0");
    if (filters == 9) return set_arg_error("AMaZE needs a Bayer pattern");
    LRE_HIP_CHECK(hipSetDevice(ctx->device));
    const size_t n = (size_t)w * h;
    const size_t io = aligned(n * 3 * 4) * 2 + aligned(n * 4);
    int rc = lre_ensure_scratch(ctx, io + lre_amaze_scratch_bytes(ctx, w, h), ctx->stream);
    if (rc) return rc;
    lre_scratch_guard scratch_use{ctx, ctx->stream};
    Carve c(ctx->scratch);
    float *d_in = c.take<float>(n * 3), *d_out = c.take<float>(n * 3), *d_cfa = c.take<float>(n);
    void *d_ws = static_cast<char *>(ctx->scratch) + c.off;
    if ((rc = copy_in(ctx, d_in, raw_rgb, n * 12))) return rc;
    if ((rc = lre_launch_rgb3_to_cfa(d_in, d_cfa, w, h, filters, nullptr, ctx->stream))) return rc;
    // the reference writes every output pixel; start from a copy of the input so
    // that nothing is left undefined
    LRE_HIP_CHECK(hipMemcpyAsync(d_out, d_in, n * 12, hipMemcpyDeviceToDevice, ctx->stream));
    if ((rc = lre_launch_amaze(ctx, d_cfa, d_out, w, h, filters, cam_mul, maximum_value, d_ws,
                               ctx->stream)))
        return rc;
    return copy_out(ctx, rgb, d_out, n * 12);
}

int lre_hip_xtrans3(lre_ctx *ctx, const float *raw_rgb, float *rgb, int w, int h,
                    const signed char xtrans[36], float maximum_value) {
    if (!ctx || !raw_rgb || !rgb || !xtrans) return set_arg_error("null pointer");
    if (!valid_dims(w, h)) return set_arg_error("image size");
    if (!(maximum_value > 0.f)) return set_arg_error("maximum_value must be > 0");
    LRE_HIP_CHECK(hipSetDevice(ctx->device));
    const size_t n = (size_t)w * h;
    const size_t io = aligned(n * 3 * 4) * 2 + aligned(n * 4);
    int rc = lre_ensure_scratch(ctx, io + lre_xtrans_scratch_bytes(ctx, w, h), ctx->stream);
    if (rc) return rc;
    lre_scratch_guard scratch_use{ctx, ctx->stream};
    Carve c(ctx->scratch);
    float *d_in = c.take<float>(n * 3), *d_out = c.take<float>(n * 3), *d_cfa = c.take<float>(n);
    void *d_ws = static_cast<char *>(ctx->scratch) + c.off;
    if ((rc = copy_in(ctx, d_in, raw_rgb, n * 12))) return rc;
    if ((rc = lre_launch_rgb3_to_cfa(d_in, d_cfa, w, h, 9, xtrans, ctx->stream))) return rc;
    // The reference writes the 8-px frame and whatever its tiles cover; with no
    // tile (a side of 22 px or less) the pixels inside the frame keep the
    // caller's values, so those go to the device first.
    if (lre_xtrans_tiles(w, h) == 0 && (rc = copy_in(ctx, d_out, rgb, n * 12))) return rc;
    if ((rc = lre_launch_xtrans3(ctx, d_cfa, d_out, w, h, xtrans, maximum_value, d_ws, ctx->stream)))
        return rc;
    return copy_out(ctx, rgb, d_out, n * 12);
}

int lre_hip_tone_mapping(lre_ctx *ctx, const float *in_rgb, float *out_rgb, int w, int h,
                         float after_scale) {
    if (!ctx || !in_rgb || !out_rgb) return set_arg_error("null pointer");
    if (!valid_dims(w, h)) return set_arg_error("image size");
    LRE_HIP_CHECK(hipSetDevice(ctx->device));
    const size_t n = (size_t)w * h * 3;
    int rc = lre_ensure_scratch(ctx, aligned(n * 4), ctx->stream);
    if (rc) return rc;
    lre_scratch_guard scratch_use{ctx, ctx->stream};
    float *d = static_cast<float *>(ctx->scratch);
    if ((rc = copy_in(ctx, d, in_rgb, n * 4))) return rc;
    if ((rc = lre_launch_tone_mapping(d, d, n, after_scale, ctx->stream))) return rc;
    return copy_out(ctx, out_rgb, d, n * 4);
}

int lre_hip_detail_tonemap(lre_ctx *ctx, const float *in_rgb, float *out_rgb, int w, int h) {
    if (!ctx || !in_rgb || !out_rgb) return set_arg_error("null pointer");
    if (!valid_dims(w, h)) return set_arg_error("image size");
    LRE_HIP_CHECK(hipSetDevice(ctx->device));
    const size_t n = (size_t)w * h * 3;
    int rc = lre_ensure_scratch(ctx, aligned(n * 4) + lre_detail_tm_scratch_bytes(w, h), ctx->stream);
    if (rc) return rc;
    lre_scratch_guard scratch_use{ctx, ctx->stream};
    float *d = static_cast<float *>(ctx->scratch);
    void *ws = static_cast<char *>(ctx->scratch) + aligned(n * 4);
    if ((rc = copy_in(ctx, d, in_rgb, n * 4))) return rc;
    if ((rc = lre_launch_detail_tonemap(d, d, w, h, ws, ctx->stream))) return rc;
    return copy_out(ctx, out_rgb, d, n * 4);
}

int lre_hip_micro_contrast(lre_ctx *ctx, const float *in_rgb, float *out_rgb, int w, int h,
                           float threshold, float strength, float target_contrast,
                           const float *mask) {
    if (!ctx || !in_rgb || !out_rgb) return set_arg_error("null pointer");
    if (!valid_dims(w, h)) return set_arg_error("image size");
    LRE_HIP_CHECK(hipSetDevice(ctx->device));
    const size_t n = (size_t)w * h;
    int rc = lre_ensure_scratch(ctx, aligned(n * 12) + (mask ? aligned(n * 4) : 0) +
                                         lre_micro_contrast_scratch_bytes(w, h), ctx->stream);
    if (rc) return rc;
    lre_scratch_guard scratch_use{ctx, ctx->stream};
    Carve c(ctx->scratch);
    float *d = c.take<float>(n * 3);
    float *d_mask = mask ? c.take<float>(n) : nullptr;
    void *ws = static_cast<char *>(ctx->scratch) + c.off;
    if ((rc = copy_in(ctx, d, in_rgb, n * 12))) return rc;
    if (mask && (rc = copy_in(ctx, d_mask, mask, n * 4))) return rc;
    if ((rc = lre_launch_micro_contrast(d, d, w, h, threshold, strength, target_contrast, d_mask, ws,
                                        ctx->stream)))
        return rc;
    return copy_out(ctx, out_rgb, d, n * 12);
}

int lre_hip_recover_highlights(lre_ctx *ctx, const float *in_rgb, float *out_rgb, int w, int h,
                               float saturation_threshold, float *mask_out, const float *wb_rgb,
                               float *max_val_out) {
    if (!ctx || !in_rgb || !out_rgb) return set_arg_error("null pointer");
    if (!valid_dims(w, h)) return set_arg_error("image size");
    LRE_HIP_CHECK(hipSetDevice(ctx->device));
    const size_t n = (size_t)w * h;
    int rc = lre_ensure_scratch(ctx, 2 * aligned(n * 12) + (mask_out ? aligned(n * 4) : 0) + 256 +
                                         lre_highlight_scratch_bytes(w, h), ctx->stream);
    if (rc) return rc;
    lre_scratch_guard scratch_use{ctx, ctx->stream};
    Carve c(ctx->scratch);
    float *d_in = c.take<float>(n * 3);
    float *d_out = c.take<float>(n * 3);
    float *d_mask = mask_out ? c.take<float>(n) : nullptr;
    float *d_max = c.take<float>(1);
    void *ws = static_cast<char *>(ctx->scratch) + c.off;
    if ((rc = copy_in(ctx, d_in, in_rgb, n * 12))) return rc;
    if ((rc = lre_launch_recover_highlights(d_in, d_out, w, h, saturation_threshold, d_mask, wb_rgb, d_max, ws,
                                            ctx->stream)))
        return rc;
    if (mask_out && (rc = copy_out(ctx, mask_out, d_mask, n * 4))) return rc;
    if (max_val_out && (rc = copy_out(ctx, max_val_out, d_max, 4))) return rc;
    return copy_out(ctx, out_rgb, d_out, n * 12);
}

int lre_hip_defringe(lre_ctx *ctx, const float *in_rgb, float *out_rgb, int w, int h,
                     float radius, float edge_threshold, float chroma_threshold, float strength,
                     int defringe_green, float green_strength_scale) {
    if (!ctx || !in_rgb || !out_rgb) return set_arg_error("null pointer");
    if (!valid_dims(w, h)) return set_arg_error("image size");
    LRE_HIP_CHECK(hipSetDevice(ctx->device));
    const size_t n = (size_t)w * h;
    const bool inplace = in_rgb == out_rgb;
    int rc = lre_ensure_scratch(ctx, aligned(n * 12) * 2 + lre_defringe_scratch_bytes(w, h), ctx->stream);
    if (rc) return rc;
    lre_scratch_guard scratch_use{ctx, ctx->stream};
    Carve c(ctx->scratch);
    float *d_in = c.take<float>(n * 3), *d_out = c.take<float>(n * 3);
    void *ws = static_cast<char *>(ctx->scratch) + c.off;
    if ((rc = copy_in(ctx, d_in, in_rgb, n * 12))) return rc;
    if ((rc = lre_launch_defringe(d_in, inplace ? d_in : d_out, w, h, radius, edge_threshold,
                                  chroma_threshold, strength, defringe_green,
                                  green_strength_scale, ws, ctx->stream)))
        return rc;
    return copy_out(ctx, out_rgb, inplace ? d_in : d_out, n * 12);
}

int lre_hip_axial_ca(lre_ctx *ctx, const float *in_rgb, float *out_rgb, int w, int h, int radius,
                     float epsilon, float strength) {
    return rgb_stage_host(ctx, in_rgb, out_rgb, w, h, lre_axial_scratch_bytes(w, h),
                          [&](const float *di, float *dout, void *ws) {
                              return lre_launch_axial_ca(di, dout, w, h, radius, epsilon, strength,
                                                         ws, ctx->stream);
                          });
}

int lre_hip_lateral_ca(lre_ctx *ctx, const float *in_rgb, float *out_rgb, int w, int h,
                       int cell_size, int max_iterations, float max_shift, float min_confidence,
                       int pyramid_levels) {
    return rgb_stage_host(ctx, in_rgb, out_rgb, w, h, lre_lateral_scratch_bytes(w, h, cell_size),
                          [&](const float *di, float *dout, void *ws) {
                              return lre_launch_lateral_ca(di, dout, w, h, cell_size,
                                                           max_iterations, max_shift,
                                                           min_confidence, pyramid_levels, ws,
                                                           ctx->stream);
                          });
}

int lre_hip_convert_color_space(lre_ctx *ctx, const float *in_rgb, float *out_rgb, int w, int h,
                                const float transform[12]) {
    if (!transform) return set_arg_error("null transform");
    return rgb_stage_host(ctx, in_rgb, out_rgb, w, h, 0, [&](const float *di, float *dout, void *) {
        return lre_launch_epilogue(di, dout, (size_t)w * h, transform, 0.f, 0.f, -1, 32, ctx->stream);
    });
}

int lre_hip_gamma_correct(lre_ctx *ctx, const float *in_rgb, float *out_rgb, int w, int h,
                          float gamma_power, float gamma_slope, int output_color_space) {
    if (output_color_space < 0) return set_arg_error("output colour space");
    return rgb_stage_host(ctx, in_rgb, out_rgb, w, h, 0, [&](const float *di, float *dout, void *) {
        return lre_launch_epilogue(di, dout, (size_t)w * h, nullptr, gamma_power, gamma_slope,
                                   output_color_space, 32, ctx->stream);
    });
}

int lre_hip_quantize(lre_ctx *ctx, const float *in, void *out, size_t count, int bits) {
    if (!ctx || !in || !out) return set_arg_error("null pointer");
    if (bits != 8 && bits != 16) return set_arg_error("bits must be 8 or 16");
    if (count % 3) return set_arg_error("count must be a multiple of 3 (RGB samples)");
    LRE_HIP_CHECK(hipSetDevice(ctx->device));
    const size_t ob = count * (size_t)(bits / 8);
    int rc = lre_ensure_scratch(ctx, aligned(count * 4) + aligned(ob), ctx->stream);
    if (rc) return rc;
    lre_scratch_guard scratch_use{ctx, ctx->stream};
    Carve c(ctx->scratch);
    float *d_in = c.take<float>(count);
    void *d_out = c.take<char>(ob);
    if ((rc = copy_in(ctx, d_in, in, count * 4))) return rc;
    if ((rc = lre_launch_epilogue(d_in, d_out, count / 3, nullptr, 0.f, 0.f, -1, bits, ctx->stream)))
        return rc;
    return copy_out(ctx, out, d_out, ob);
}

int lre_hip_dev_convert_color_space(lre_ctx *ctx, const float *d_in, float *d_out, int w, int h,
                                    const float transform[12], void *stream) {
    if (!ctx || !d_in || !d_out || !transform) return set_arg_error("null pointer");
    if (!valid_dims(w, h)) return set_arg_error("image size");
    LRE_HIP_CHECK(hipSetDevice(ctx->device));
    return lre_launch_epilogue(d_in, d_out, (size_t)w * h, transform, 0.f, 0.f, -1, 32,
                               lre_pick_stream(ctx, stream));
}

int lre_hip_dev_gamma_correct(lre_ctx *ctx, const float *d_in, float *d_out, int w, int h,
                              float gamma_power, float gamma_slope, int output_color_space,
                              void *stream) {
    if (!ctx || !d_in || !d_out) return set_arg_error("null pointer");
    if (!valid_dims(w, h)) return set_arg_error("image size");
    if (output_color_space < 0) return set_arg_error("output colour space");
    LRE_HIP_CHECK(hipSetDevice(ctx->device));
    return lre_launch_epilogue(d_in, d_out, (size_t)w * h, nullptr, gamma_power, gamma_slope,
                               output_color_space, 32, lre_pick_stream(ctx, stream));
}

int lre_hip_dev_epilogue(lre_ctx *ctx, const float *d_rgb, void *d_out, int w, int h,
                         const float transform[12], float gamma_power, float gamma_slope,
                         int output_color_space, int bits, void *stream) {
    if (!ctx || !d_rgb || !d_out) return set_arg_error("null pointer");
    if (!valid_dims(w, h)) return set_arg_error("image size");
    if (bits != 32 && d_out == static_cast<const void *>(d_rgb))
        return set_arg_error("a quantized output cannot overwrite its input");
    LRE_HIP_CHECK(hipSetDevice(ctx->device));
    return lre_launch_epilogue(d_rgb, d_out, (size_t)w * h, transform, gamma_power, gamma_slope,
                               output_color_space, bits, lre_pick_stream(ctx, stream));
}

int lre_hip_green_matching(lre_ctx *ctx, uint16_t *img4, int w, int h, uint32_t filters,
                           int maximum) {
    if (!ctx || !img4) return set_arg_error("null pointer");
    if (!valid_dims(w, h)) return set_arg_error("image size");
    if (filters == 9) return LRE_OK;
    LRE_HIP_CHECK(hipSetDevice(ctx->device));
    const size_t n = (size_t)w * h;
    int rc = lre_ensure_scratch(ctx, aligned(n * 8) * 2, ctx->stream);
    if (rc) return rc;
    lre_scratch_guard scratch_use{ctx, ctx->stream};
    Carve c(ctx->scratch);
    uint16_t *d_src = c.take<uint16_t>(n * 4), *d_dst = c.take<uint16_t>(n * 4);
    if ((rc = copy_in(ctx, d_src, img4, n * 8))) return rc;
    if ((rc = lre_launch_green_matching(d_src, d_dst, w, h, filters, maximum, ctx->stream))) return rc;
    return copy_out(ctx, img4, d_dst, n * 8);
}

int lre_hip_dev_green_matching(lre_ctx *ctx, const uint16_t *d_src, uint16_t *d_dst, int w, int h,
                               uint32_t filters, int maximum, void *stream) {
    if (!ctx || !d_src || !d_dst) return set_arg_error("null pointer");
    if (!valid_dims(w, h)) return set_arg_error("image size");
    if (d_src == d_dst && filters != 9) return set_arg_error("source and destination must differ");
    LRE_HIP_CHECK(hipSetDevice(ctx->device));
    return lre_launch_green_matching(d_src, d_dst, w, h, filters, maximum, lre_pick_stream(ctx, stream));
}

int lre_hip_remove_zeroes(lre_ctx *ctx, uint16_t *img4, int w, int h, uint32_t filters) {
    if (!ctx || !img4) return set_arg_error("null pointer");
    if (!valid_dims(w, h)) return set_arg_error("image size");
    LRE_HIP_CHECK(hipSetDevice(ctx->device));
    const size_t n = (size_t)w * h;
    int rc = lre_ensure_scratch(ctx, aligned(n * 8) + lre_remove_zeroes_scratch_bytes(w, h), ctx->stream);
    if (rc) return rc;
    lre_scratch_guard scratch_use{ctx, ctx->stream};
    Carve c(ctx->scratch);
    uint16_t *d = c.take<uint16_t>(n * 4);
    void *ws = static_cast<char *>(ctx->scratch) + c.off;
    if ((rc = copy_in(ctx, d, img4, n * 8))) return rc;
    if ((rc = lre_launch_remove_zeroes(d, w, h, filters, ws, ctx->stream))) return rc;
    return copy_out(ctx, img4, d, n * 8);
}

int lre_hip_dev_remove_zeroes(lre_ctx *ctx, uint16_t *d_img4, int w, int h, uint32_t filters,
                              void *stream) {
    if (!ctx || !d_img4) return set_arg_error("null pointer");
    if (!valid_dims(w, h)) return set_arg_error("image size");
    LRE_HIP_CHECK(hipSetDevice(ctx->device));
    int rc = lre_ensure_scratch(ctx, lre_remove_zeroes_scratch_bytes(w, h), lre_pick_stream(ctx, stream));
    if (rc) return rc;
    lre_scratch_guard scratch_use{ctx, lre_pick_stream(ctx, stream)};
    return lre_launch_remove_zeroes(d_img4, w, h, filters, ctx->scratch, lre_pick_stream(ctx, stream));
}

int lre_hip_auto_wb(lre_ctx *ctx, const uint16_t *img4, int w, int h, uint32_t filters,
                    const signed char xtrans[36], float pre_mul[4]) {
    if (!ctx || !img4 || !pre_mul || (filters == 9 && !xtrans)) return set_arg_error("null pointer");
    if (!valid_dims(w, h)) return set_arg_error("image size");
    LRE_HIP_CHECK(hipSetDevice(ctx->device));
    const size_t n = (size_t)w * h;
    int rc = lre_ensure_scratch(ctx, aligned(n * 8) + 256 + 256, ctx->stream);
    if (rc) return rc;
    lre_scratch_guard scratch_use{ctx, ctx->stream};
    Carve c(ctx->scratch);
    uint16_t *d = c.take<uint16_t>(n * 4);
    float *d_pm = c.take<float>(4);
    void *ws = c.take<char>(64);
    if ((rc = copy_in(ctx, d, img4, n * 8))) return rc;
    if ((rc = lre_launch_auto_wb(d, w, h, filters, xtrans, d_pm, ws, ctx->stream))) return rc;
    return copy_out(ctx, pre_mul, d_pm, 16);
}

int lre_hip_dev_auto_wb(lre_ctx *ctx, const uint16_t *d_img4, int w, int h, uint32_t filters,
                        const signed char xtrans[36], float *d_pre_mul, void *stream) {
    if (!ctx || !d_img4 || !d_pre_mul || (filters == 9 && !xtrans)) return set_arg_error("null pointer");
    if (!valid_dims(w, h)) return set_arg_error("image size");
    LRE_HIP_CHECK(hipSetDevice(ctx->device));
    int rc = lre_ensure_scratch(ctx, 256, lre_pick_stream(ctx, stream));
    if (rc) return rc;
    lre_scratch_guard scratch_use{ctx, lre_pick_stream(ctx, stream)};
    return lre_launch_auto_wb(d_img4, w, h, filters, xtrans, d_pre_mul, ctx->scratch,
                              lre_pick_stream(ctx, stream));
}

int lre_hip_scale_aber(lre_ctx *ctx, uint16_t *img4, int w, int h, const float aber[4]) {
    if (!ctx || !img4 || !aber) return set_arg_error("null pointer");
    if (!valid_dims(w, h)) return set_arg_error("image size");
    LRE_HIP_CHECK(hipSetDevice(ctx->device));
    const size_t n = (size_t)w * h;
    int rc = lre_ensure_scratch(ctx, aligned(n * 8) + lre_scale_aber_scratch_bytes(w, h), ctx->stream);
    if (rc) return rc;
    lre_scratch_guard scratch_use{ctx, ctx->stream};
    Carve c(ctx->scratch);
    uint16_t *d = c.take<uint16_t>(n * 4);
    void *ws = static_cast<char *>(ctx->scratch) + c.off;
    if ((rc = copy_in(ctx, d, img4, n * 8))) return rc;
    if ((rc = lre_launch_scale_aber(d, w, h, aber, ws, ctx->stream))) return rc;
    return copy_out(ctx, img4, d, n * 8);
}

int lre_hip_dev_scale_aber(lre_ctx *ctx, uint16_t *d_img4, int w, int h, const float aber[4],
                           void *stream) {
    if (!ctx || !d_img4 || !aber) return set_arg_error("null pointer");
    if (!valid_dims(w, h)) return set_arg_error("image size");
    LRE_HIP_CHECK(hipSetDevice(ctx->device));
    int rc = lre_ensure_scratch(ctx, lre_scale_aber_scratch_bytes(w, h), lre_pick_stream(ctx, stream));
    if (rc) return rc;
    lre_scratch_guard scratch_use{ctx, lre_pick_stream(ctx, stream)};
    return lre_launch_scale_aber(d_img4, w, h, aber, ctx->scratch, lre_pick_stream(ctx, stream));
}

int lre_hip_black_level(lre_ctx *ctx, uint16_t *img4, int w, int h, const int cblack[4],
                        const int *pos_table, int pos_rows, int pos_cols, int black, int maximum) {
    if (!ctx || !img4 || !cblack) return set_arg_error("null pointer");
    if (!valid_dims(w, h)) return set_arg_error("image size");
    const bool positional = pos_table && pos_rows > 0 && pos_cols > 0;
    if (positional && (size_t)pos_rows * pos_cols > 4096) return set_arg_error("positional table size");
    LRE_HIP_CHECK(hipSetDevice(ctx->device));
    const size_t n = (size_t)w * h;
    int rc = lre_ensure_scratch(ctx, aligned(n * 8) + aligned(4096 * 4), ctx->stream);
    if (rc) return rc;
    lre_scratch_guard scratch_use{ctx, ctx->stream};
    Carve c(ctx->scratch);
    uint16_t *d_img = c.take<uint16_t>(n * 4);
    int *d_pos = c.take<int>(4096);
    if ((rc = copy_in(ctx, d_img, img4, n * 8))) return rc;
    if (positional && (rc = copy_in(ctx, d_pos, pos_table, sizeof(int) * pos_rows * pos_cols))) return rc;
    if ((rc = lre_launch_black_level(d_img, w, h, cblack, d_pos, positional ? pos_rows : 0,
                                     positional ? pos_cols : 0, black, maximum, ctx->stream)))
        return rc;
    return copy_out(ctx, img4, d_img, n * 8);
}

int lre_hip_white_balance(lre_ctx *ctx, const uint16_t *img4, float *rgb, int w, int h,
                          const float wb[4], uint32_t filters, const signed char xtrans[36]) {
    if (!ctx || !img4 || !rgb || !wb) return set_arg_error("null pointer");
    if (!valid_dims(w, h)) return set_arg_error("image size");
    if (filters == 9 && !xtrans) return set_arg_error("X-Trans table");
    LRE_HIP_CHECK(hipSetDevice(ctx->device));
    const size_t n = (size_t)w * h;
    int rc = lre_ensure_scratch(ctx, aligned(n * 8) + aligned(n * 12), ctx->stream);
    if (rc) return rc;
    lre_scratch_guard scratch_use{ctx, ctx->stream};
    Carve c(ctx->scratch);
    uint16_t *d_in = c.take<uint16_t>(n * 4);
    float *d_out = c.take<float>(n * 3);
    if ((rc = copy_in(ctx, d_in, img4, n * 8))) return rc;
    // channels other than the native one keep the caller's values
    if ((rc = copy_in(ctx, d_out, rgb, n * 12))) return rc;
    if ((rc = lre_launch_white_balance(d_in, d_out, w, h, wb, filters, xtrans, ctx->stream)))
        return rc;
    return copy_out(ctx, rgb, d_out, n * 12);
}

}  // extern "C"

// ---- scale_colors multiplier logic (core/libraw_wrapper.cpp:612-862) --------
int lre_wb_needs_auto(const lre_wb_meta *m) {
    if (!m) return 0;
    return m->use_auto_wb ||
           (m->use_camera_wb &&
            (m->cam_mul[0] < -0.5 || (m->cam_mul[0] <= 0.00001f && !m->camerawb_daylight_fallback)));
}

int lre_scale_colors_meta(lre_wb_meta *m, const float auto_pre_mul[4], float scale_mul[4]) {
    if (!m || !scale_mul) return set_arg_error("null pointer");
    const bool use_auto = lre_wb_needs_auto(m) != 0;
    if (use_auto && !auto_pre_mul) return set_arg_error("automatic WB needs the image-wide multipliers");
    // 1. user multipliers (:612-615)
    if (m->user_mul[0] != 0.f) std::memcpy(m->pre_mul, m->user_mul, sizeof(m->pre_mul));
    // 2. automatic multipliers (:620-663), from the image-wide sums
    if (use_auto) std::memcpy(m->pre_mul, auto_pre_mul, sizeof(m->pre_mul));
    // 3. camera white balance from the white samples (:743-779)
    if (m->use_camera_wb && m->cam_mul[0] > 0.00001f) {
        unsigned sum[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (unsigned row = 0; row < 8; row++)
            for (unsigned col = 0; col < 8; col++) {
                const unsigned c = lre::bayer_native(m->filters, row, col);
                // unsigned difference stored into an int, as in the reference
                const int val = (int)((unsigned)m->white[row][col] - m->cblack[c]);
                if (val > 0) sum[c] += (unsigned)val;
                sum[c + 4]++;
            }
        if (m->as_shot_wb_applied) {
            m->pre_mul[0] = m->pre_mul[1] = m->pre_mul[2] = m->pre_mul[3] = 1.0f;
        } else if (sum[0] && sum[1] && sum[2] && sum[3]) {
            for (int c = 0; c < 4; c++) m->pre_mul[c] = (float)sum[c + 4] / sum[c];
        } else if (m->cam_mul[0] > 0.00001f && m->cam_mul[2] > 0.00001f) {
            std::memcpy(m->pre_mul, m->cam_mul, sizeof(m->pre_mul));
        } else {
            m->warnings |= LRE_WARN_BAD_CAMERA_WB;
        }
    }
    // 4. Nikon sRAW daylight (:784-795)
    if (m->as_shot_wb_applied && !m->use_camera_wb && !m->use_auto_wb && m->cam_mul[0] > 0.00001f &&
        m->cam_mul[1] > 0.00001f && m->cam_mul[2] > 0.00001f)
        for (int c = 0; c < 3; c++) m->pre_mul[c] /= m->cam_mul[c];
    // 5. normalisation (:800-806)
    if (m->pre_mul[1] == 0) m->pre_mul[1] = 1;
    if (m->pre_mul[3] == 0) m->pre_mul[3] = m->colors < 4 ? m->pre_mul[1] : 1;
    // 7. scaling multipliers (:819-845); maximum is unsigned, as LibRaw's
    m->maximum -= m->black;
    if (!use_auto) {
        double dmin = std::numeric_limits<double>::max();
        for (int c = 0; c < 4; c++)
            if (dmin > m->pre_mul[c]) dmin = m->pre_mul[c];
        if (dmin > 0.00001 && m->maximum > 0)
            for (int c = 0; c < 4; c++) scale_mul[c] = (float)(m->pre_mul[c] / dmin);
        else
            for (int c = 0; c < 4; c++) scale_mul[c] = 1.0f;
    } else {
        for (int c = 0; c < 4; c++) scale_mul[c] = m->pre_mul[c];
    }
    // 8. a 1x1 or 2x2 cblack pattern folded into the per-channel blacks (:850-862)
    unsigned *cb = m->cblack;
    if (m->filters > 1000 && (cb[4] + 1) / 2 == 1 && (cb[5] + 1) / 2 == 1) {
        for (unsigned c = 0; c < 4; c++)
            cb[lre::bayer_native(m->filters, c / 2, c % 2)] += cb[6 + c / 2 % cb[4] * cb[5] + c % 2 % cb[5]];
        cb[4] = cb[5] = 0;
    }
    return LRE_OK;
}
