/*
 * CPU oracle for the sliding-window correlation path of imageproc 0.27.0.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: it is linked or
 * loaded by tests/, by the smoke check in __graft_entry__.py and by the CPU
 * baseline leg of bench.py. The product (imageproc_amd/) never calls it.
 *
 * Two modes:
 *   - reference semantics: same loop order as the reference, sequential
 *     non-fused f32 (`acc = acc + p * w`), truncating Clamp casts, the
 *     P-typed intermediate between separable passes;
 *   - FP64 (functions suffixed _f64): exact integer sums and double
 *     normalisation, for error reporting.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the imageproc source tree).
 */
#ifndef IPH_ORACLE_H
#define IPH_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Element type codes shared with the product C ABI. */
enum { OR_U8 = 0, OR_U16 = 1, OR_I16 = 2, OR_I32 = 3, OR_F32 = 4 };

/* MatchTemplateMethod, template_matching.rs:8-64 (declaration order). */
enum { OR_SSE = 0, OR_SSE_NORMED = 1, OR_CCORR = 2, OR_CCORR_NORMED = 3 };

/* gaussian_kernel_f32, filter/mod.rs:131-145. Returns the tap count, writes
 * min(count, cap) taps. */
int or_gaussian_kernel_f32(float sigma, float *out, int cap);

/* horizontal_filter / vertical_filter, filter/mod.rs:228-405.
 * p_type in {OR_U8, OR_F32}, k_type in {OR_F32, OR_I32}. Returns 0, or -1 for a
 * (pixel, kernel) pair the reference does not type-check. */
int or_horizontal_filter(const void *src, void *dst, uint32_t w, uint32_t h, int p_type,
                         const void *kern, uint32_t klen, int k_type);
int or_vertical_filter(const void *src, void *dst, uint32_t w, uint32_t h, int p_type,
                       const void *kern, uint32_t klen, int k_type);
/* separable_filter, filter/mod.rs:166-183 (H pass then V pass on the
 * P-typed intermediate). Returns -2 if the kernel lengths differ. */
int or_separable_filter(const void *src, void *dst, uint32_t w, uint32_t h, int p_type,
                        const void *hk, uint32_t hlen, const void *vk, uint32_t vlen,
                        int k_type);
/* gaussian_blur_f32, filter/mod.rs:147-164. Returns -3 for sigma <= 0. */
int or_gaussian_blur_f32(const void *src, void *dst, uint32_t w, uint32_t h, int p_type,
                         float sigma);

/* filter_clamped, filter/mod.rs:197-210 via filter, filter/mod.rs:24-69.
 * Returns -1 for a (P, K, S) triple the reference does not type-check. */
int or_filter_clamped(const void *src, int p_type, uint32_t w, uint32_t h, const void *kern,
                      int k_type, uint32_t kw, uint32_t kh, void *dst, int s_type);
/* filter_clamped_parallel, filter/mod.rs:212-226: rows split over threads. */
int or_filter_clamped_parallel(const void *src, int p_type, uint32_t w, uint32_t h,
                               const void *kern, int k_type, uint32_t kw, uint32_t kh,
                               void *dst, int s_type, int nthreads);

/* box_filter, filter/box_filter.rs:8-70 with row_running_sum and
 * column_running_sum, integral_image.rs:318-463. */
void or_box_filter(const uint8_t *src, uint8_t *dst, uint32_t w, uint32_t h, uint32_t rx,
                   uint32_t ry);

/* match_template, template_matching.rs:66-92, 165-293, 378-425 (f32,
 * reference order). out is (w-tw+1) x (h-th+1). Returns -1 if the template is
 * larger than the image. */
int or_match_template_u8(const uint8_t *img, uint32_t w, uint32_t h, const uint8_t *tpl,
                         uint32_t tw, uint32_t th, int method, float *out);
/* match_template_parallel, template_matching.rs:94-110, 183-203. */
int or_match_template_u8_parallel(const uint8_t *img, uint32_t w, uint32_t h,
                                  const uint8_t *tpl, uint32_t tw, uint32_t th, int method,
                                  float *out, int nthreads);
/* match_template_with_mask, template_matching.rs:112-140, 295-386, 427-461.
 * Returns -2 if the mask size differs from the template size. */
int or_match_template_with_mask_u8(const uint8_t *img, uint32_t w, uint32_t h,
                                   const uint8_t *tpl, uint32_t tw, uint32_t th,
                                   const uint8_t *mask, uint32_t mw, uint32_t mh, int method,
                                   float *out);
/* FP64 oracle for u8 images: exact integer sums, double normalisation. */
int or_match_template_u8_f64(const uint8_t *img, uint32_t w, uint32_t h, const uint8_t *tpl,
                             uint32_t tw, uint32_t th, int method, double *out);
/* The same four formulas over Luma<f32> images (an extension: the reference
 * only takes GrayImage). Reference order in f32, and an FP64 version. */
int or_match_template_f32(const float *img, uint32_t w, uint32_t h, const float *tpl,
                          uint32_t tw, uint32_t th, int method, float *out);
int or_match_template_f32_parallel(const float *img, uint32_t w, uint32_t h, const float *tpl,
                                   uint32_t tw, uint32_t th, int method, float *out,
                                   int nthreads);
int or_match_template_f32_f64(const float *img, uint32_t w, uint32_t h, const float *tpl,
                              uint32_t tw, uint32_t th, int method, double *out);

/* find_extremes, template_matching.rs:463-511. Writes
 * {max, min} values and {max_x, max_y, min_x, min_y}. Returns -1 if empty. */
int or_find_extremes_f32(const float *img, uint32_t w, uint32_t h, float vals[2],
                         uint32_t locs[4]);
int or_find_extremes_u8(const uint8_t *img, uint32_t w, uint32_t h, uint8_t vals[2],
                        uint32_t locs[4]);

/* integral_image::<_, u32> / integral_squared_image::<_, u32>,
 * integral_image.rs:50-152: (w+1) x (h+1), zero first row and column,
 * wrapping u32 arithmetic. */
void or_integral_image_u32(const uint8_t *img, uint32_t w, uint32_t h, uint32_t *out,
                           int squared);

/* adaptive_threshold, contrast.rs:19-47 (radii below 2^31 whose clipped
 * block end does not overflow u32). Returns -1 for block_radius 0. */
int or_adaptive_threshold(const uint8_t *img, uint32_t w, uint32_t h, uint32_t block_radius,
                          int32_t delta, uint8_t *out);

/* Direction class (0, 45, 90, 135) of non_maximum_suppression, edges.rs:62-80. */
int or_canny_direction(int gx, int gy);
/* or_canny_direction and hypotf over [-r, r]^2 (row-major by gy, then gx). */
void or_canny_direction_grid(int r, uint8_t *cls, float *mag);
/* canny, edges.rs:30-155. Returns -1 for high < low (or NaN), -3 for an empty
 * image, -4 where the reference's search indexes outside the image. */
int or_canny(const uint8_t *img, uint32_t w, uint32_t h, float low, float high, uint8_t *out);
/* canny's stages: thinned magnitudes (edges.rs:34-50), and hysteresis
 * (edges.rs:110-155) over them with a stack of 2 * w * h entries. */
int or_canny_thinned(const uint8_t *img, uint32_t w, uint32_t h, float *thin);
int or_canny_hysteresis(const float *thin, uint32_t w, uint32_t h, float low, float high,
                        uint32_t *stack, uint8_t *out);

#ifdef __cplusplus
}
#endif
#endif
