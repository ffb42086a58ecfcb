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
 *
 * The calls on one context share its device scratch.  A call on a different
 * stream from the previous call waits on the host for that call's work first,
 * so calls on several streams are correct but do not overlap; use one context
 * per stream for overlap.  Work captured into a HIP graph uses the scratch
 * whenever the graph is launched: order such launches against other calls on
 * the context yourself.
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

/* Accelerator::enhance_micro_contrast (core/accelerator.h:338-341,
 * core/cpu_accelerator.cpp:3135-3319); defaults strength 8, target_contrast
 * 0.06 (the pipeline passes 0.04 and its highlight mask).  mask: w x h floats
 * in [0, 1] blending input and enhanced result, or NULL for the gate on green
 * (green >= threshold enhances, else the pixel passes through).  in == out is
 * allowed. */
int lre_hip_micro_contrast(lre_ctx *ctx, const float *in_rgb, float *out_rgb, int width,
                           int height, float threshold, float strength, float target_contrast,
                           const float *mask);

/* LibRawWrapper::Impl::recover_highlights (core/libraw_wrapper.cpp:1095-1383),
 * the float highlight recovery of postprocess for highlight_mode > 2
 * (:1659-1683) and of recover_highlights_numpy (:3063-3095).
 * saturation_threshold: the pipeline's maximum / data_maximum.  mask_out: w x h
 * floats receiving the soft highlight mask of the input, or NULL (the
 * reference computes it only when asked).  wb_rgb: the effective R, G, B white
 * balance (it only changes the mask), or NULL.  max_val_out: the largest
 * channel of the highlight pixels before the rewrite (the value whose
 * non-zeroness the reference returns), or NULL.  in == out is allowed. */
int lre_hip_recover_highlights(lre_ctx *ctx, const float *in_rgb, float *out_rgb, int width,
                               int height, float saturation_threshold, float *mask_out,
                               const float *wb_rgb, float *max_val_out);

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
/* No counterpart in the reference.  lre_hip_dev_amaze for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): same arguments, same output bits, and no
 * allocation of any kind.  Its tile arenas and tile-queue counter live in a
 * fixed device array of the library, 48 MiB per GPU (32 tile arenas), set up
 * by lre_hip_create and held until the process ends.  With 32 arenas at most
 * 32 workgroups run, so it is slower than lre_hip_dev_amaze on large images.
 * The arena pool does not depend on the image: every size lre_hip_dev_amaze
 * accepts is accepted (width * height below 2^31).  All calls on one GPU share
 * that GPU's array, whatever the context or stream: calls that overlap in time
 * on one GPU race and must be serialised by the caller. */
int lre_hip_dev_amaze_noalloc(lre_ctx *ctx, const float *d_cfa, float *d_rgb, int width,
                              int height, uint32_t filters, const float cam_mul[4],
                              float maximum_value, void *stream);
/* X-Trans 3-pass from a float CFA plane to interleaved RGB. */
int lre_hip_dev_xtrans3(lre_ctx *ctx, const float *d_cfa, float *d_rgb, int width, int height,
                        const signed char xtrans[36], float maximum_value, void *stream);
int lre_hip_dev_tone_mapping(lre_ctx *ctx, const float *d_in, float *d_out, int width,
                             int height, float after_scale, void *stream);
/* Detail-preserving tone map on resident buffers.  In place (d_in == d_out)
 * it first writes a guide plane into the context scratch; with a separate
 * output it reads the guide from d_in directly and needs no scratch (faster,
 * same result).  d_in and d_out must not otherwise overlap. */
int lre_hip_dev_detail_tonemap(lre_ctx *ctx, const float *d_in, float *d_out, int width,
                               int height, void *stream);
/* Micro-contrast on resident buffers; d_mask is a device plane or NULL. */
int lre_hip_dev_micro_contrast(lre_ctx *ctx, const float *d_in, float *d_out, int width,
                               int height, float threshold, float strength,
                               float target_contrast, const float *d_mask, void *stream);
/* Highlight recovery on resident buffers; d_mask and d_max_val are device
 * pointers or NULL, wb_rgb a host array of 3 or NULL; d_in == d_out is allowed. */
int lre_hip_dev_recover_highlights(lre_ctx *ctx, const float *d_in, float *d_out, int width,
                                   int height, float saturation_threshold, float *d_mask,
                                   const float *wb_rgb, float *d_max_val, void *stream);
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
/* apply_green_matching (core/libraw_wrapper.cpp:496-585) on the uint16 x 4 image
 * after black subtraction: flat, unsaturated channel-3 (G2) sites are scaled by
 * the ratio of their channel-1 and channel-3 neighbourhood means.  No-op for
 * X-Trans (filters == 9).  The host form works in place. */
int lre_hip_green_matching(lre_ctx *ctx, uint16_t *img4, int width, int height, uint32_t filters,
                           int maximum);
/* Device form: reads d_src, writes d_dst (which must differ from d_src). */
int lre_hip_dev_green_matching(lre_ctx *ctx, const uint16_t *d_src, uint16_t *d_dst, int width,
                               int height, uint32_t filters, int maximum, void *stream);
/* remove_zeroes (core/libraw_wrapper.cpp:150-212), in place: in raster order, a
 * zero at its native channel becomes the integer mean of the non-zero samples of
 * that channel in its 5x5 window (earlier replacements included). */
int lre_hip_remove_zeroes(lre_ctx *ctx, uint16_t *img4, int width, int height, uint32_t filters);
int lre_hip_dev_remove_zeroes(lre_ctx *ctx, uint16_t *d_img4, int width, int height,
                              uint32_t filters, void *stream);
/* The LibRaw fields that scale_colors (core/libraw_wrapper.cpp:602-862) reads
 * and updates, outside the image itself. */
typedef struct lre_wb_meta {
    float user_mul[4];             /* params.user_mul; user_mul[0] == 0: not set */
    int use_auto_wb;               /* params.use_auto_wb */
    int use_camera_wb;             /* params.use_camera_wb */
    int camerawb_daylight_fallback; /* rawparams.options has CAMERAWB_FALLBACK_TO_DAYLIGHT */
    int as_shot_wb_applied;        /* color.as_shot_wb_applied (Nikon sRAW) */
    int colors;                    /* idata.colors */
    uint32_t filters;              /* idata.filters */
    float cam_mul[4];              /* color.cam_mul */
    float pre_mul[4];              /* color.pre_mul, updated */
    uint16_t white[8][8];          /* color.white */
    unsigned cblack[10];           /* color.cblack[0..9]; updated (pattern folded) */
    unsigned black;                /* color.black */
    unsigned maximum;              /* color.maximum; updated (black subtracted) */
    unsigned warnings;             /* LRE_WARN_BAD_CAMERA_WB when camera WB is unusable */
} lre_wb_meta;
#define LRE_WARN_BAD_CAMERA_WB 0x4u /* LibRaw's LIBRAW_WARN_BAD_CAMERA_WB */

/* Whether scale_colors takes the image-wide automatic multipliers
 * (core/libraw_wrapper.cpp:620-629): then the caller passes the result of
 * lre_hip_auto_wb to lre_scale_colors_meta.  Host logic only. */
int lre_wb_needs_auto(const lre_wb_meta *meta);
/* The multiplier part of scale_colors (core/libraw_wrapper.cpp:612-862 without
 * the image-wide sums and the radial aberration): user, automatic and camera
 * multipliers, the Nikon sRAW daylight case, pre_mul normalisation, maximum
 * less black, scale_mul = pre_mul / min(pre_mul) (pre_mul itself on the
 * automatic path), and the folding of a 1x1 / 2x2 cblack pattern into
 * cblack[0..3].  auto_pre_mul: required when lre_wb_needs_auto, else ignored.
 * Host logic only (no device work). */
int lre_scale_colors_meta(lre_wb_meta *meta, const float auto_pre_mul[4], float scale_mul[4]);

/* The automatic white-balance multipliers of scale_colors
 * (core/libraw_wrapper.cpp:621-663): the mean of every native channel over the
 * image (channel 3 folded into 1), pre_mul = (G/R, 1, G/B, 1). */
int lre_hip_auto_wb(lre_ctx *ctx, const uint16_t *img4, int width, int height, uint32_t filters,
                    const signed char xtrans[36], float pre_mul[4]);
/* Device form: d_pre_mul receives the 4 multipliers. */
int lre_hip_dev_auto_wb(lre_ctx *ctx, const uint16_t *d_img4, int width, int height,
                        uint32_t filters, const signed char xtrans[36], float *d_pre_mul,
                        void *stream);
/* The legacy radial chromatic-aberration correction of scale_colors
 * (core/libraw_wrapper.cpp:833-880), in place: channels 0 and 2 resampled
 * bilinearly about the image centre by aber[0] and aber[2] (1 = unchanged). */
int lre_hip_scale_aber(lre_ctx *ctx, uint16_t *img4, int width, int height, const float aber[4]);
int lre_hip_dev_scale_aber(lre_ctx *ctx, uint16_t *d_img4, int width, int height,
                           const float aber[4], void *stream);
/* Demosaic input planes: extract the native channel of the reference's 3-channel
 * layout into a CFA plane. */
int lre_hip_dev_rgb3_to_cfa(lre_ctx *ctx, const float *d_raw_rgb, float *d_cfa, int width,
                            int height, uint32_t filters, const signed char xtrans[36],
                            void *stream);

/* ---- output epilogue ---------------------------------------------------------- */

/* Accelerator::convert_color_space (core/accelerator.h:310-313,
 * core/cpu_accelerator.cpp:2874-2897): out = T[:, :3] * rgb + T[:, 3], with T the
 * row-major 3x4 transform (12 floats). */
int lre_hip_convert_color_space(lre_ctx *ctx, const float *in_rgb, float *out_rgb, int width,
                                int height, const float transform[12]);
/* Accelerator::gamma_correct (core/accelerator.h:315-319,
 * core/cpu_accelerator.cpp:2903-2976): per channel, by output colour space
 * (0 raw, 1 sRGB, 2 Adobe, 3 wide gamut, 4 ProPhoto, 5 XYZ, 6 ACEScg, 7 P3-D65,
 * 8 Rec.2020; core/metal/constants.h:28-38). */
int lre_hip_gamma_correct(lre_ctx *ctx, const float *in_rgb, float *out_rgb, int width,
                          int height, float gamma_power, float gamma_slope,
                          int output_color_space);
/* The 8/16-bit output of the processed image (core/python_bindings.cpp:235-257):
 * clamp(v, 0, 1) * 255 (65535) + 0.5, truncated.  out holds count uint8 or uint16. */
int lre_hip_quantize(lre_ctx *ctx, const float *in, void *out, size_t count, int bits);
int lre_hip_dev_convert_color_space(lre_ctx *ctx, const float *d_in, float *d_out, int width,
                                    int height, const float transform[12], void *stream);
int lre_hip_dev_gamma_correct(lre_ctx *ctx, const float *d_in, float *d_out, int width,
                              int height, float gamma_power, float gamma_slope,
                              int output_color_space, void *stream);
/* All three in one pass over resident RGB (the order of the develop pipeline,
 * core/libraw_wrapper.cpp:1829-1842, then the output conversion): transform may be
 * NULL, output_color_space < 0 skips the gamma, bits is 8, 16 or 32 (float out).
 * d_out may equal d_rgb when bits is 32. */
int lre_hip_dev_epilogue(lre_ctx *ctx, const float *d_rgb, void *d_out, int width, int height,
                         const float transform[12], float gamma_power, float gamma_slope,
                         int output_color_space, int bits, void *stream);

/* ---- many images: batched develop and the request loop --------------------
 * The reference develops one image per postprocess() call
 * (core/libraw_wrapper.cpp:1477-1656) and has no batch or serving layer; these
 * entry points run the same stage functions (the ones above, same parameters
 * and defaults) over many device-resident images. */

/* Stage codes of an ordered stage list.  LRE_STAGE_DEMOSAIC, when present, must
 * come first (AMaZE for Bayer images, X-Trans 3-pass for filters == 9), after
 * the raw-preprocessing stages when those are listed: LRE_STAGE_REMOVE_ZEROES
 * and LRE_STAGE_GREEN_MATCHING, in that order, each at most once, followed by
 * LRE_STAGE_DEMOSAIC.  They act on Bayer images with uint16 input in the
 * reference's order (core/libraw_wrapper.cpp:1496-1547): zero removal on the
 * raw samples, black subtraction and clip, green matching with the white level
 * as maximum, then white balance; float-CFA and X-Trans images skip them. */
enum {
    LRE_STAGE_DEMOSAIC = 1,
    LRE_STAGE_LATERAL_CA = 2,
    LRE_STAGE_AXIAL_CA = 3,
    LRE_STAGE_DEFRINGE = 4,
    LRE_STAGE_DETAIL_TM = 5,
    LRE_STAGE_TONE_MAPPING = 6,
    LRE_STAGE_COLOR_SPACE = 7,
    LRE_STAGE_GAMMA = 8,
    /* recover_highlights (core/libraw_wrapper.cpp:1095-1383) with the image's
     * soft highlight mask kept for a later micro-contrast stage; WB for the
     * mask: the image's wb when it is ingested from uint16, else none */
    LRE_STAGE_HIGHLIGHT = 9,
    /* enhance_micro_contrast as postprocess calls it (:1698-1705): threshold
     * highlight_threshold + 0.3, micro_strength, micro_target_contrast, and the
     * highlight mask when a highlight stage came earlier in the list */
    LRE_STAGE_MICRO_CONTRAST = 10,
    /* remove_zeroes (core/libraw_wrapper.cpp:150-212), as lre_hip_remove_zeroes */
    LRE_STAGE_REMOVE_ZEROES = 11,
    /* apply_green_matching (:496-585), as lre_hip_green_matching */
    LRE_STAGE_GREEN_MATCHING = 12
};

/* Parameters of the post-demosaic stages. */
typedef struct lre_chain_params {
    float defringe_radius, defringe_edge_threshold, defringe_chroma_threshold, defringe_strength;
    int defringe_green;
    float defringe_green_scale;
    int axial_radius;
    float axial_epsilon, axial_strength;
    int lateral_cell, lateral_iterations;
    float lateral_max_shift, lateral_min_confidence;
    int lateral_levels;
    float tone_after_scale;
    float color_transform[12];   /* row-major 3x4 (identity by default) */
    float gamma_power, gamma_slope;
    int output_color_space;
    float highlight_threshold;   /* maximum / data_maximum of the raw (1 when unknown) */
    float micro_strength, micro_target_contrast;
} lre_chain_params;

/* The reference's defaults (core/accelerator.h:344-367, :322-323, :315-319). */
void lre_chain_params_default(lre_chain_params *p);

/* One image.  Input is either a float CFA plane (d_cfa, with maximum_value as
 * for lre_hip_dev_amaze) or, when d_cfa is NULL, a uint16 CFA plane d_raw that
 * is preprocessed on the device exactly as lre_hip_dev_ingest_u16 (black,
 * cblack, white = clip level, wb) followed by the adjust_maximum rule
 * (core/libraw_wrapper.cpp:1070-1077) with maximum = white - black and
 * adjust_threshold.  All stages run in place on d_rgb (H x W x 3 float). */
typedef struct lre_image {
    int width, height;
    uint32_t filters;          /* Bayer filters word, or 9 for X-Trans */
    signed char xtrans[36];    /* X-Trans pattern when filters == 9 */
    float cam_mul[4];          /* AMaZE clip point, as lre_hip_amaze */
    const float *d_cfa;
    float maximum_value;
    const uint16_t *d_raw;
    int black, white;
    int cblack[4];
    float wb[4];
    float adjust_threshold;
    float *d_rgb;
} lre_image;

/* Batched develop over resident images.  One multi-image tile queue per
 * demosaic kind (all Bayer images in one AMaZE launch, all X-Trans images in
 * one X-Trans launch).  The remaining stages run in group mode when it is on
 * (see lre_hip_batch_set_grouping) and every one of them is axial CA, defringe
 * (when its blur kernel has at most 16 taps and every gate's ramp width lies
 * in [2^-20, 2^100], as at the reference's defaults), detail tone map, tone
 * mapping, colour matrix or gamma: the images are spread over `lanes` HIP
 * streams, each stream's images form groups whose stage scratch fits its share
 * of the post-stage budget, and each stage is one launch over a whole group.
 * Otherwise each image runs its stages on one of the `lanes` streams.  Both
 * give the same bits as the per-image entry points.  lre_hip_batch_run only
 * enqueues; lre_hip_batch_synchronize waits.  A batch object is used by one
 * host thread at a time. */
typedef struct lre_batch lre_batch;
int lre_hip_batch_create(lre_batch **out, int device, int lanes);
int lre_hip_batch_destroy(lre_batch *b);
int lre_hip_batch_run(lre_batch *b, const lre_image *images, int n, const int *stages,
                      int nstages, const lre_chain_params *params);
int lre_hip_batch_synchronize(lre_batch *b);
/* Record the batch (same arguments as lre_hip_batch_run) into a HIP graph after
 * one eager run that sizes every buffer; lre_hip_batch_replay launches it.  The
 * graph holds kernel nodes only and refers to the images' device buffers.  A
 * later lre_hip_batch_run that needs larger arenas or scratch reallocates them
 * and drops the graph: replay then fails until the batch is captured again. */
int lre_hip_batch_capture(lre_batch *b, const lre_image *images, int n, const int *stages,
                          int nstages, const lre_chain_params *params);
int lre_hip_batch_replay(lre_batch *b);
/* Launch shape: lanes (1..8) and persistent workgroups per CU of the AMaZE and
 * X-Trans tile queues (1..16). */
int lre_hip_batch_set_config(lre_batch *b, int lanes, int amaze_slots_per_cu,
                             int xtrans_slots_per_cu);
int lre_hip_batch_get_config(lre_batch *b, int *lanes, int *amaze_slots_per_cu,
                             int *xtrans_slots_per_cu);
/* Group mode off (0), on (1) or by image size (2, the default: on when the
 * batch's mean image is below 2 MP, where one launch per stage is faster than
 * per-image chains), and the device bytes the groups' stage scratch may take
 * (0: 8 GiB, or a quarter of the device's memory if less; an image that alone
 * exceeds its lane's share forms a group of its own). */
int lre_hip_batch_set_grouping(lre_batch *b, int group_mode, size_t post_budget_bytes);
int lre_hip_batch_get_grouping(lre_batch *b, int *group_mode, size_t *post_budget_bytes);
/* Tile height of the defringe blur: 48 or 96 rows, or 0 (the default) to pick
 * it by image size (48 while 64 x 96 tiles would give fewer than four per CU).
 * Both heights give the same bits. */
int lre_hip_batch_set_tiles(lre_batch *b, int defringe_blur_rows);
int lre_hip_batch_get_tiles(lre_batch *b, int *defringe_blur_rows);
/* Time `reps` runs of each candidate launch shape (group mode on and off,
 * lanes, slots per CU) on this batch and keep the fastest, then the defringe
 * blur tile height (48, 96) on that shape; *best_ms receives the time per batch
 * of what is kept. */
int lre_hip_batch_autotune(lre_batch *b, const lre_image *images, int n, const int *stages,
                           int nstages, const lre_chain_params *params, int reps, float *best_ms);

/* Request loop.  Requests carry host pointers (uint16 CFA in, float RGB out);
 * lre_hip_host_alloc gives pinned buffers for the fastest copies.  A request is
 * admitted when one of `lanes` worker streams is free and its device footprint
 * (input, float CFA, RGB and stage scratch) fits in hbm_budget_bytes beside the
 * requests already admitted; admitted requests run concurrently on the device.
 * The stage list must start with LRE_STAGE_DEMOSAIC (a request carries a raw
 * plane); the parameters of its stages are checked at creation.  A request
 * that fails posts its error status and leaves the others unaffected: its
 * completion is posted only after all of its device work has finished. */
typedef struct lre_server lre_server;
typedef struct lre_request {
    uint64_t id;
    lre_image image;           /* geometry and preprocessing; device pointers ignored */
    const uint16_t *h_raw;     /* width * height uint16 CFA samples */
    float *h_rgb;              /* width * height * 3 floats */
} lre_request;
typedef struct lre_completion {
    uint64_t id;
    int status;                /* LRE_OK or an error code */
    double t_submit, t_admit, t_done;  /* seconds on the server's monotonic clock */
} lre_completion;
int lre_hip_serve_create(lre_server **out, int device, int lanes, size_t hbm_budget_bytes,
                         const int *stages, int nstages, const lre_chain_params *params);
/* Queue a request (FIFO).  The host buffers must stay valid until its completion. */
int lre_hip_serve_submit(lre_server *s, const lre_request *req);
/* Up to max completions; waits up to timeout_ms for the first one.  Returns the count. */
int lre_hip_serve_poll(lre_server *s, lre_completion *out, int max, int timeout_ms);
/* The server's monotonic clock (seconds), for arrival schedules. */
double lre_hip_serve_now(lre_server *s);
/* Device bytes the server holds for admitted requests, and the peak so far. */
int lre_hip_serve_footprint(lre_server *s, size_t *current, size_t *peak);
int lre_hip_serve_destroy(lre_server *s);
void *lre_hip_host_alloc(size_t bytes);
int lre_hip_host_free(void *p);

#ifdef __cplusplus
}
#endif
#endif /* LRE_HIP_H */
