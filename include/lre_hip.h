/*
 * lre_hip.h — C ABI of the MI355X (gfx950) develop path of libraw_enhanced.
 *
 * A HIP backend that sits where the reference's Accelerator dispatches to its
 * CPU or Metal backends (core/accelerator.h:247-375, core/accelerator.cpp:15-24).
 * Every entry point returns 0 on success and a negative status on failure
 * (the caller falls back to the CPU path, as the reference's facade does on a
 * `false` return, core/accelerator.cpp:141-157); lre_hip_last_error() gives the
 * message.  Nothing here throws.
 *
 * Image layout: row-major H x W.  "rgb" buffers are interleaved float RGB
 * (the reference's ImageBufferFloat, core/accelerator.h:71-80); "cfa" planes are
 * one float per site; "img4" is the reference's uint16 x 4 ImageBuffer
 * (core/accelerator.h:59-68).
 *
 * Two families of entry points:
 *   lre_hip_<stage>(ctx, host pointers...)  copy in, run, copy out, synchronise —
 *       the drop-in for the corresponding Accelerator method;
 *   lre_hip_dev_<stage>(ctx, device pointers..., stream)  enqueue only, on the
 *       given HIP stream (NULL = the context's stream); inputs stay in HBM.
 */
#ifndef LRE_HIP_H
#define LRE_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    LRE_OK = 0,
    LRE_ERR_ARG = -1,      /* invalid argument (null pointer, bad size, mismatch) */
    LRE_ERR_HIP = -2,      /* HIP runtime error */
    LRE_ERR_NODEV = -3,    /* no usable gfx950 device */
    LRE_ERR_NOMEM = -4,    /* device allocation failed */
};

typedef struct lre_ctx lre_ctx;

/* 1 when a HIP device is present and usable. */
int lre_hip_available(void);
/* Context bound to one device: owns a stream and grow-only device scratch. */
int lre_hip_create(lre_ctx **ctx, int device);
int lre_hip_destroy(lre_ctx *ctx);
/* Message of the last failure on the calling thread ("" if none). */
const char *lre_hip_last_error(void);
/* Wait for all work enqueued on the context's stream. */
int lre_hip_synchronize(lre_ctx *ctx);
/* The context's HIP stream (hipStream_t), for callers that enqueue their own work. */
void *lre_hip_stream(lre_ctx *ctx);

/* Device memory plumbing for callers without their own allocator. */
int lre_hip_dev_malloc(lre_ctx *ctx, void **d_ptr, size_t bytes);
int lre_hip_dev_free(lre_ctx *ctx, void *d_ptr);
int lre_hip_memcpy_htod(lre_ctx *ctx, void *d_dst, const void *h_src, size_t bytes);
int lre_hip_memcpy_dtoh(lre_ctx *ctx, void *h_dst, const void *d_src, size_t bytes);

/* ---- host-pointer entry points (drop-in for the Accelerator methods) ---- */

/* Accelerator::demosaic_bayer_amaze (core/accelerator.h:279-281,
 * core/cpu_accelerator.cpp:1635-1653): raw_rgb holds the CFA sample at its
 * native channel (other channels ignored); the output is clipped at 0 and the
 * 4-px frame is rebuilt by border interpolation. */
int lre_hip_amaze(lre_ctx *ctx, const float *raw_rgb, float *rgb, int width, int height,
                  uint32_t filters, const float cam_mul[4], float maximum_value);

/* Accelerator::demosaic_xtrans_3pass (core/accelerator.h:291-295,
 * core/cpu_accelerator.cpp:1870-2676 and :2776-2783): Markesteijn 3-pass
 * with YPbPr homogeneity, then the weighted border interpolation of the 8-px
 * frame.  xtrans is the 6x6 colour table (0 R, 1 G, 2 B), row-major.  The
 * reference's colour matrix argument only feeds its disabled CIELab branch and
 * is not needed.  Width at most 16382 (16-bit hexagon offsets, as the reference). */
int lre_hip_xtrans3(lre_ctx *ctx, const float *raw_rgb, float *rgb, int width, int height,
                    const signed char xtrans[36], float maximum_value);

/* Accelerator::tone_mapping (core/accelerator.h:322-323,
 * core/cpu_accelerator.cpp:3034-3068): ACES per channel times after_scale.
 * in == out is allowed. */
int lre_hip_tone_mapping(lre_ctx *ctx, const float *in_rgb, float *out_rgb, int width,
                         int height, float after_scale);

/* Accelerator::apply_detail_preserving_tonemap_gpu (core/accelerator.h:329-330);
 * the CPU version is LibRawWrapper's apply_detail_preserving_tonemap
 * (core/libraw_wrapper.cpp:1398-1471).  in == out is allowed. */
int lre_hip_detail_tonemap(lre_ctx *ctx, const float *in_rgb, float *out_rgb, int width,
                           int height);

/* Accelerator::defringe (core/accelerator.h:344-351, core/cpu_accelerator.cpp:3939-5064),
 * defaults radius 10, edge_threshold 0.1, chroma_threshold 0.15, strength 10,
 * defringe_green 0, green_strength_scale 0.3.  in_rgb == out_rgb selects the
 * reference's in-place semantics (the solo cleanup then sees the pixel's
 * corrected values, core/libraw_wrapper.cpp:1810); distinct buffers select the
 * standalone semantics (core/libraw_wrapper.cpp:3211). */
int lre_hip_defringe(lre_ctx *ctx, const float *in_rgb, float *out_rgb, int width, int height,
                     float radius, float edge_threshold, float chroma_threshold, float strength,
                     int defringe_green, float green_strength_scale);

/* Accelerator::ca_axial_cleanup (core/accelerator.h:363-367,
 * core/cpu_accelerator.cpp:3799-3928): guided filter of R and B with G as the
 * guide; defaults radius 6 (at most 16 here), epsilon 1e-4, strength 0.3.
 * in == out is allowed. */
int lre_hip_axial_ca(lre_ctx *ctx, const float *in_rgb, float *out_rgb, int width, int height,
                     int radius, float epsilon, float strength);

/* Accelerator::ca_register_lateral (core/accelerator.h:354-360,
 * core/cpu_accelerator.cpp:3496-3673): pyramidal Lucas-Kanade registration of R
 * and B to G per cell, then a bilinear warp; defaults cell 96, 3 iterations,
 * max shift 6, min confidence 0.02, 3 levels.  in == out is allowed. */
int lre_hip_lateral_ca(lre_ctx *ctx, const float *in_rgb, float *out_rgb, int width, int height,
                       int cell_size, int max_iterations, float max_shift, float min_confidence,
                       int pyramid_levels);

/* LibRawWrapper's apply_black_level_correction (core/libraw_wrapper.cpp:401-491)
 * on the uint16 x 4 image, in place: every channel of every pixel becomes
 * clamp(v - pos_black - cblack[c], 0, maximum) when a positional table
 * (pos_rows x pos_cols, LibRaw's cblack[6..]) is given, else
 * clamp(v - cblack[c], 0, maximum) when any cblack[c] is non-zero, else
 * clamp(v - black, 0, maximum). */
int lre_hip_black_level(lre_ctx *ctx, uint16_t *img4, int width, int height, const int cblack[4],
                        const int *pos_table, int pos_rows, int pos_cols, int black, int maximum);

/* Accelerator::apply_white_balance (core/accelerator.h:304-307,
 * core/cpu_accelerator.cpp:2789-2869): writes float(sample) * wb[c] to the
 * native channel of rgb and leaves the other two channels untouched.
 * filters == 9 selects the X-Trans table xtrans (36 entries, row-major). */
int lre_hip_white_balance(lre_ctx *ctx, const uint16_t *img4, float *rgb, int width,
                          int height, const float wb[4], uint32_t filters,
                          const signed char xtrans[36]);

/* ---- device-resident entry points (enqueue on `stream`, no host sync) ---- */

/* AMaZE from a float CFA plane (one sample per site) to interleaved RGB. */
int lre_hip_dev_amaze(lre_ctx *ctx, const float *d_cfa, float *d_rgb, int width, int height,
                      uint32_t filters, const float cam_mul[4], float maximum_value,
                      void *stream);
/* X-Trans 3-pass from a float CFA plane to interleaved RGB. */
int lre_hip_dev_xtrans3(lre_ctx *ctx, const float *d_cfa, float *d_rgb, int width, int height,
                        const signed char xtrans[36], float maximum_value, void *stream);
int lre_hip_dev_tone_mapping(lre_ctx *ctx, const float *d_in, float *d_out, int width,
                             int height, float after_scale, void *stream);
int lre_hip_dev_detail_tonemap(lre_ctx *ctx, const float *d_in, float *d_out, int width,
                               int height, void *stream);
/* Defringe on resident buffers; d_in == d_out selects the in-place semantics. */
int lre_hip_dev_defringe(lre_ctx *ctx, const float *d_in, float *d_out, int width, int height,
                         float radius, float edge_threshold, float chroma_threshold,
                         float strength, int defringe_green, float green_strength_scale,
                         void *stream);
int lre_hip_dev_axial_ca(lre_ctx *ctx, const float *d_in, float *d_out, int width, int height,
                         int radius, float epsilon, float strength, void *stream);
int lre_hip_dev_lateral_ca(lre_ctx *ctx, const float *d_in, float *d_out, int width, int height,
                           int cell_size, int max_iterations, float max_shift,
                           float min_confidence, int pyramid_levels, void *stream);
/* Fused ingest of a uint16 CFA plane (one sample per site): black subtraction
 * and clip as lre_hip_black_level (per-channel or global black, no positional
 * table), white balance float(v) * wb[c] at the native channel (as
 * apply_white_balance), into a float CFA plane.  When d_data_max is not NULL it
 * receives the data maximum that adjust_maximum tests (the maximum of each
 * pixel's fcol_bayer channel after white balance, core/libraw_wrapper.cpp:1037-1062). */
int lre_hip_dev_ingest_u16(lre_ctx *ctx, const uint16_t *d_raw, float *d_cfa, int width,
                           int height, uint32_t filters, const signed char xtrans[36],
                           const int cblack[4], int black, int maximum, const float wb[4],
                           float *d_data_max, void *stream);
/* Demosaic input planes: extract the native channel of the reference's 3-channel
 * layout into a CFA plane. */
int lre_hip_dev_rgb3_to_cfa(lre_ctx *ctx, const float *d_raw_rgb, float *d_cfa, int width,
                            int height, uint32_t filters, const signed char xtrans[36],
                            void *stream);

#ifdef __cplusplus
}
#endif
#endif /* LRE_HIP_H */
