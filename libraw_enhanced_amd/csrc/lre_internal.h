// Internal host-side declarations shared by the C-ABI and the kernel files.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>

#include "lre_epilogue.h"
#include "lre_hip.h"

struct lre_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // Grow-only device scratch, carved by each stage.  A context is used by one
    // host thread at a time (the same contract as the reference's Accelerator).
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    // The stream of the last call that used the scratch, and an event recorded
    // there after its work (scratch_pending: the event is set and not yet waited
    // for).  A call on another stream first waits for that event.
    hipStream_t scratch_stream = nullptr;
    // this device's address of the no-alloc AMaZE arenas (null if unavailable)
    void *noalloc_arenas = nullptr;
    hipEvent_t scratch_done = nullptr;
    bool scratch_pending = false;
    // Pinned staging for the host-pointer entry points.
    void *pinned = nullptr;
    size_t pinned_bytes = 0;
    int num_cus = 0;
};

// Scratch of a call that enqueues on stream s: waits (on the host) for the work
// of an earlier call on another stream, then grows the scratch to `bytes`.
int lre_ensure_scratch(lre_ctx *ctx, size_t bytes, hipStream_t s);
// Marks the end of a call's use of the scratch on stream s: records the
// context's event there (not while s is being captured into a graph: a graph
// uses the scratch when it is launched, and its launches are the caller's to
// order against other calls on the context).
void lre_scratch_release(lre_ctx *ctx, hipStream_t s);
struct lre_scratch_guard {
    lre_ctx *ctx;
    hipStream_t s;
    ~lre_scratch_guard() { lre_scratch_release(ctx, s); }
};
inline hipStream_t lre_pick_stream(lre_ctx *ctx, void *stream) {
    return stream ? static_cast<hipStream_t>(stream) : ctx->stream;
}

// Bytes of device scratch each stage needs for an image of w x h.
size_t lre_amaze_scratch_bytes(lre_ctx *ctx, int w, int h);
size_t lre_detail_tm_scratch_bytes(int w, int h);
size_t lre_micro_contrast_scratch_bytes(int w, int h);
size_t lre_highlight_scratch_bytes(int w, int h);
size_t lre_defringe_scratch_bytes(int w, int h);
size_t lre_axial_scratch_bytes(int w, int h);
size_t lre_xtrans_scratch_bytes(lre_ctx *ctx, int w, int h);
size_t lre_lateral_scratch_bytes(int w, int h, int cell_size);

// One image of a multi-image demosaic launch.  d_maxv, when set, points at the
// demosaic maximum computed on the device (it then replaces maximum_value).
struct lre_demosaic_job {
    const float *d_cfa;
    float *d_rgb;
    int w, h;
    uint32_t filters;
    signed char xtrans[36];
    float cam_mul[4];
    float maximum_value;
    const float *d_maxv;
};

// One image of a multi-image post-demosaic pass: its RGB in and out (the
// same buffer for in place) and its own scratch for the stage.
struct lre_post_job {
    const float *d_in;
    float *d_out;
    int w, h;
    void *d_scratch;
};
// Multi-image passes: *_multi_table_bytes sizes the host and device table,
// *_multi_fill writes the host table (grid sizes out), the caller copies it
// to the device, *_launch_*_multi enqueues one launch per kernel over all
// images.  Parameters are common to the batch.
size_t lre_detail_tm_multi_table_bytes(int n);
int lre_detail_tm_multi_fill(const lre_post_job *jobs, int n, void *h_table, int grids[2]);
int lre_launch_detail_tm_multi(const void *d_table, int n, const int grids[2], hipStream_t s,
                               const lre::Epilogue *epi = nullptr);

size_t lre_axial_multi_table_bytes(int n);
int lre_axial_multi_fill(const lre_post_job *jobs, int n, int radius, float epsilon, float strength,
                         void *h_table, int grids[3]);
int lre_launch_axial_multi(const void *d_table, int n, int radius, const int grids[3], hipStream_t s);

size_t lre_defringe_multi_table_bytes(int n);
size_t lre_defringe_multi_scratch_bytes(int w, int h);
// 1 when the parameters have a multi-image pass (fused blur radius and the
// quotient smoothsteps), else the images take the per-image launcher
int lre_defringe_multi_supported(float radius, float edge_threshold, float chroma_threshold, float strength,
                                 int defringe_green, float green_strength_scale);
// rows: the blur tile height, as lre_defringe_blur_rows returns it (the same
// value for the fill and the launch)
int lre_defringe_multi_fill(const lre_post_job *jobs, int n, void *h_table, int grids[6], int rows);
int lre_launch_defringe_multi(const void *d_table, int n, const int grids[6], float radius, float edge_threshold,
                              float chroma_threshold, float strength, int defringe_green,
                              float green_strength_scale, int rows, hipStream_t s);
// Blur tile height of defringe: blur_rows if it is one of the two heights
// (48, 96), else chosen by the number of 64 x 96 tiles to cover (short tiles
// below four per CU).  Both heights give the same bits.
int lre_defringe_blur_rows(int blur_rows, size_t tall_tiles);
size_t lre_defringe_tall_tiles(int w, int h);
// compute units of the current device (cached per device)
int lre_current_cus();

// The parameter checks the launchers make (LRE_OK or LRE_ERR_ARG with the
// message set), for callers that validate a stage list before enqueueing.
int lre_check_axial_params(int radius);
int lre_check_defringe_params(float radius);

// Kernel launchers (device pointers; scratch must already be large enough).
int lre_launch_amaze(lre_ctx *ctx, const float *d_cfa, float *d_rgb, int w, int h,
                     uint32_t filters, const float cam_mul[4], float maximum_value,
                     void *d_scratch, hipStream_t s, const float *d_maxv = nullptr);
// AMaZE without allocation: a fixed pool of LRE_AMAZE_NOALLOC_SLOTS tile
// arenas in a device array of the library.  _setup gives the current device's
// address of that array (once per device, that device current).
constexpr int LRE_AMAZE_NOALLOC_SLOTS = 32;
constexpr size_t LRE_AMAZE_NOALLOC_BYTES = (size_t)48 << 20;
int lre_amaze_noalloc_setup(void **d_arenas);
int lre_launch_amaze_noalloc(const float *d_cfa, float *d_rgb, int w, int h, uint32_t filters,
                             const float cam_mul[4], float maximum_value, void *d_arenas, hipStream_t s);
// Multi-image tile queues.  *_fill writes the per-image table on the host and
// returns the total tile count (<0 on error); the caller copies it to d_table
// before the launch.  Table and arena sizes come from the *_bytes helpers.
size_t lre_amaze_multi_table_bytes(int n);
size_t lre_amaze_multi_scratch_bytes(int slots);
int lre_amaze_tiles(int w, int h);
int lre_amaze_multi_fill(const lre_demosaic_job *jobs, int n, void *h_table);
int lre_launch_amaze_multi(const lre_demosaic_job *jobs, int n, const void *d_table, int total_tiles,
                           int slots, void *d_scratch, hipStream_t s);
size_t lre_xtrans_multi_table_bytes(int n);
size_t lre_xtrans_multi_scratch_bytes(int slots);
int lre_xtrans_tiles(int w, int h);
int lre_xtrans_multi_fill(const lre_demosaic_job *jobs, int n, void *h_table);
// h_table: the filled host table (read for the per-image border launches)
int lre_launch_xtrans_multi(const lre_demosaic_job *jobs, int n, const void *h_table, void *d_table,
                            int total_tiles, int slots, void *d_scratch, hipStream_t s);
// Output epilogue: optional 3x4 matrix (transform12 NULL skips it), optional gamma
// (output_color_space < 0 skips it), then float (bits 32) or quantized (8, 16) store.
// The matrix / gamma part of the epilogue as one parameter set (transform12
// NULL: no matrix; output_color_space < 0: no gamma).
lre::Epilogue lre_make_epilogue(const float *transform12, float gamma_power, float gamma_slope,
                                int output_color_space);
int lre_launch_epilogue(const float *d_in, void *d_out, size_t npx, const float *transform12,
                        float gamma_power, float gamma_slope, int output_color_space, int bits,
                        hipStream_t s);
// Green matching (d_src != d_dst unless X-Trans) and zero removal on uint16 x 4 images.
int lre_launch_green_matching(const uint16_t *d_src, uint16_t *d_dst, int w, int h, uint32_t filters,
                              int maximum, hipStream_t s);
size_t lre_remove_zeroes_scratch_bytes(int w, int h);
// The same two stages on a one-sample-per-site uint16 plane (Bayer), and the
// black subtraction and clip of such a plane in place; the zero removal
// scratch is lre_remove_zeroes_scratch_bytes.
int lre_launch_remove_zeroes_plane(uint16_t *d_plane, int w, int h, uint32_t filters, void *d_scratch,
                                   hipStream_t s);
int lre_launch_green_matching_plane(const uint16_t *d_src, uint16_t *d_dst, int w, int h, uint32_t filters,
                                    int maximum, hipStream_t s);
int lre_launch_black_plane(uint16_t *d_plane, int w, int h, uint32_t filters, const signed char *xtrans,
                           const int cblack[4], int black, int maximum, hipStream_t s);
int lre_launch_remove_zeroes(uint16_t *d_img4, int w, int h, uint32_t filters, void *d_scratch,
                             hipStream_t s);
// Automatic white balance (d_scratch: 64 bytes) and legacy radial aberration.
int lre_launch_auto_wb(const uint16_t *d_img4, int w, int h, uint32_t filters,
                       const signed char *xtrans, float *d_pre_mul, void *d_scratch, hipStream_t s);
size_t lre_scale_aber_scratch_bytes(int w, int h);
int lre_launch_scale_aber(uint16_t *d_img4, int w, int h, const float aber[4], void *d_scratch,
                          hipStream_t s);
// adjust_maximum on the device: d_maxv = rule(data_max, maximum, threshold)
int lre_launch_adjust_maximum(const float *d_data_max, float maximum, float threshold,
                              float *d_maxv, hipStream_t s);
int lre_launch_border_interpolate(const float *dummy, float *d_rgb, int w, int h,
                                  uint32_t filters, int border, hipStream_t s);
int lre_launch_tone_mapping(const float *d_in, float *d_out, size_t n, float after_scale,
                            hipStream_t s);
int lre_launch_detail_tonemap(const float *d_in, float *d_out, int w, int h, void *d_scratch,
                              hipStream_t s, const lre::Epilogue *epi = nullptr);
// d_in != d_out; wb_rgb is a host array of 3 or null; d_mask, d_max_val may be null
int lre_launch_recover_highlights(const float *d_in, float *d_out, int w, int h, float sat, float *d_mask,
                                  const float *wb_rgb, float *d_max_val, void *d_scratch, hipStream_t s);
int lre_launch_micro_contrast(const float *d_in, float *d_out, int w, int h, float threshold,
                              float strength, float target_contrast, const float *d_mask,
                              void *d_scratch, hipStream_t s);
int lre_launch_defringe(const float *d_in, float *d_out, int w, int h, float radius,
                        float edge_threshold, float chroma_threshold, float strength,
                        int defringe_green, float green_strength_scale, void *d_scratch,
                        hipStream_t s, int blur_rows = 0);
int lre_launch_xtrans3(lre_ctx *ctx, const float *d_cfa, float *d_rgb, int w, int h,
                       const signed char xtrans36[36], float maximum_value, void *d_scratch,
                       hipStream_t s);
int lre_launch_axial_ca(const float *d_in, float *d_out, int w, int h, int radius, float epsilon,
                        float strength, void *d_scratch, hipStream_t s);
int lre_launch_lateral_ca(const float *d_in, float *d_out, int w, int h, int cell_size,
                          int max_iterations, float max_shift, float min_confidence,
                          int pyramid_levels, void *d_scratch, hipStream_t s);
int lre_launch_black_level(uint16_t *d_img4, int w, int h, const int cblack[4], const int *d_pos,
                           int pos_rows, int pos_cols, int black, int maximum, hipStream_t s);
int lre_launch_ingest_u16(const uint16_t *d_raw, float *d_cfa, int w, int h, uint32_t filters,
                          const signed char *xtrans, const int cblack[4], int black, int maximum,
                          const float wb[4], float *d_data_max, hipStream_t s);
int lre_launch_white_balance(const uint16_t *d_img4, float *d_rgb, int w, int h,
                             const float wb[4], uint32_t filters, const signed char *xtrans,
                             hipStream_t s);
int lre_launch_rgb3_to_cfa(const float *d_rgb3, float *d_cfa, int w, int h, uint32_t filters,
                           const signed char *xtrans, hipStream_t s);
