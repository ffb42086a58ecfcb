/*
 * TEST INFRASTRUCTURE ONLY — CPU restatement ("oracle") of the reference's
 * develop path.  Only tests/, the smoke check and bench.py's CPU-baseline leg
 * may load liblre_oracle.so, and only as the checker or as a timed baseline;
 * the product never calls it.
 *
 * Every function names the reference file:line it restates.  Images are
 * caller-owned, row-major, interleaved float RGB (H x W x 3) unless stated.
 */
#ifndef LRE_ORACLE_H
#define LRE_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Scratch semantics of the tiled demosaics.
 *  ORACLE_TILES_CLEAN:      every tile starts from zeroed scratch (deterministic,
 *                           thread-count independent; the contract of the HIP path).
 *  ORACLE_TILES_SEQUENTIAL: one scratch buffer, zeroed once, reused by the tiles
 *                           in raster order — what the reference does with one
 *                           OpenMP thread (its per-thread buffer is reused and
 *                           some reads see the previous tile's values).
 *  ORACLE_TILES_POISON_NAN / _BIG (AMaZE only): every tile starts from scratch
 *                           filled with NaN / a large value, and only the cells
 *                           the HIP kernel clears are zeroed; equal to CLEAN
 *                           exactly when that clear set covers every cell a
 *                           tile reads before writing. */
enum { ORACLE_TILES_CLEAN = 0, ORACLE_TILES_SEQUENTIAL = 1, ORACLE_TILES_POISON_NAN = 2,
       ORACLE_TILES_POISON_BIG = 3 };

int oracle_amaze(const float *raw3, float *rgb3, int w, int h, uint32_t filters,
                 const float *cam_mul4, float maximum_value, int tile_mode);
int oracle_xtrans3(const float *raw3, float *rgb3, int w, int h, const signed char *xtrans36,
                   float maximum_value, int tile_mode);
void oracle_border_interpolate(float *rgb3, int w, int h, uint32_t filters, int border);
int oracle_tone_mapping(const float *in3, float *out3, int w, int h, float after_scale);
int oracle_detail_tonemap(float *rgb3, int w, int h);
/* mask == NULL selects the threshold gate on green; in == out is allowed */
int oracle_micro_contrast(const float *in3, float *out3, int w, int h, float threshold,
                          float strength, float target_contrast, const float *mask);
void oracle_mc_kernel(float k[81]);
/* in place; mask_out and wb_rgb may be NULL; counts3 = highlight, white and
 * LoCA-adjusted pixel counts (may be NULL).  loca_order selects the reference's
 * raster-order in-place LoCA pass or a snapshot reading of it. */
#define ORACLE_LOCA_RASTER 0
#define ORACLE_LOCA_SNAPSHOT 1
#define ORACLE_LOCA_OFF 2
int oracle_recover_highlights(float *rgb3, int w, int h, float sat, float *mask_out,
                              const float *wb_rgb, float *max_val_out, long long *counts3,
                              int loca_order);
/* in == out selects the in-place semantics of the reference */
int oracle_defringe(const float *in3, float *out3, int w, int h, float radius, float edge_threshold,
                    float chroma_threshold, float strength, int defringe_green,
                    float green_strength_scale);
int oracle_axial_ca(const float *in3, float *out3, int w, int h, int radius, float epsilon,
                    float strength);
int oracle_lateral_ca(const float *in3, float *out3, int w, int h, int cell_size, int max_iterations,
                      float max_shift, float min_confidence, int pyramid_levels);
int oracle_black_level(uint16_t *img4, int w, int h, const int *cblack4, const int *pos_table,
                       int pos_rows, int pos_cols, int black, int maximum);
int oracle_ingest_u16(const uint16_t *raw, float *cfa, int w, int h, uint32_t filters,
                      const signed char *xtrans36, const int *cblack4, int black, int maximum,
                      const float *wb4, float *data_max);
float oracle_adjust_maximum(float data_max, float maximum, float threshold);
int oracle_white_balance(const uint16_t *img4, float *rgb3, int w, int h, const float *wb4,
                         uint32_t filters, const signed char *xtrans36);

int oracle_green_matching(uint16_t *img4, int w, int h, uint32_t filters, int maximum);
int oracle_remove_zeroes(uint16_t *img4, int w, int h, uint32_t filters);
int oracle_auto_wb(const uint16_t *img4, int w, int h, uint32_t filters, const signed char *xtrans36,
                   float *pre_mul4);
int oracle_scale_aber(uint16_t *img4, int w, int h, const float *aber4);

/* output epilogue (epilogue.c) */
int oracle_convert_color_space(const float *in3, float *out3, int w, int h, const float *transform12);
float oracle_gamma_value(float v, float gamma_power, float gamma_slope, int output_color_space);
int oracle_gamma_correct(const float *in3, float *out3, int w, int h, float gamma_power,
                         float gamma_slope, int output_color_space);
int oracle_quantize(const float *in, void *out, size_t n, int bits);

#ifdef __cplusplus
}
#endif
#endif
