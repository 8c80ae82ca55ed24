/*
 * imageproc_hip.h — C ABI of the MI355X (gfx950) sliding-window correlation
 * library: imageproc's filter module (gaussian_blur_f32, separable_filter,
 * horizontal/vertical_filter, filter_clamped, box_filter) and template
 * matching (match_template, find_extremes).
 *
 * Each entry point names the imageproc 0.27.0 function it replaces
 * (file:line in the imageproc source tree). Images are row-major,
 * single-channel (Luma), stride = width, exactly as ImageBuffer stores
 * them. Element type codes are IPH_U8 (GrayImage) and IPH_F32
 * (Image<Luma<f32>>); kernel type codes IPH_F32 / IPH_I32.
 *
 * Host entry points (iph_*) take host pointers, block until dst is valid and
 * allocate device memory themselves. Device entry points (iph_dev_*) take
 * device pointers, enqueue on `stream` (a hipStream_t, NULL = default
 * stream) and return without waiting; their temporaries live in a
 * caller-provided workspace whose size the matching *_workspace query
 * returns.
 *
 * Errors: every function returns an iph_status. Where the reference panics,
 * the status is IPH_EINVAL and iph_last_error() holds the reference's panic
 * message. A (pixel, kernel, output) type combination that does not
 * type-check in the reference returns IPH_EUNSUPPORTED. Nothing here falls
 * back to the CPU: without a usable GPU the status is IPH_EDEVICE.
 */
#ifndef IMAGEPROC_HIP_H
#define IMAGEPROC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IPH_API __attribute__((visibility("default")))

typedef int32_t iph_status;
enum {
    IPH_OK = 0,
    IPH_EINVAL = 1,       /* a reference assertion failed (message in iph_last_error) */
    IPH_ENOMEM = 2,       /* device or host allocation failed */
    IPH_EDEVICE = 3,      /* HIP runtime error / no usable gfx950 device */
    IPH_EUNSUPPORTED = 4  /* type combination the reference does not accept */
};

/* Element type codes. */
enum { IPH_U8 = 0, IPH_U16 = 1, IPH_I16 = 2, IPH_I32 = 3, IPH_F32 = 4 };

/* MatchTemplateMethod, template_matching.rs:8-64, in declaration order. */
enum {
    IPH_SSE = 0,          /* SumOfSquaredErrors */
    IPH_SSE_NORMED = 1,   /* SumOfSquaredErrorsNormalized */
    IPH_CCORR = 2,        /* CrossCorrelation */
    IPH_CCORR_NORMED = 3  /* CrossCorrelationNormalized */
};

/* Correlation engine for match_template. */
enum {
    IPH_PATH_AUTO = 0,    /* MFMA path for templates up to 128 x 128, else the direct path:
                             every template the reference accepts (no larger than the
                             image, template_matching.rs:394-404) runs */
    IPH_PATH_MFMA = 1,    /* u8: int8 MFMA, exact i32 sums; f32: BF16x3 MFMA, FP32
                             accumulation; templates up to 128 x 128 (larger ->
                             IPH_EUNSUPPORTED) */
    IPH_PATH_DIRECT = 2   /* FP32 per-window loop in the reference's order (bit-exact) */
};

/* Extremes<T>, template_matching.rs:463-475. Values are stored in the
 * element type of the image (u8 in the first byte, or an f32). */
typedef struct {
    union { float f32; uint8_t u8; uint32_t raw; } max_value;
    union { float f32; uint8_t u8; uint32_t raw; } min_value;
    uint32_t max_x, max_y, min_x, min_y;
} iph_extremes;

/* ---- library ---------------------------------------------------------- */

/* Message of the last failed call on this thread ("" if none). */
IPH_API const char *iph_last_error(void);
/* Device name, CU count and gcn arch of the current device. */
IPH_API iph_status iph_device_info(char *buf, size_t cap);

/* gaussian_kernel_f32, filter/mod.rs:131-145 (host only; glibc expf).
 * Returns the tap count 2*ceil(2*sigma)+1 (or -1), writes min(count, cap). */
IPH_API int32_t iph_gaussian_kernel_f32(float sigma, float *out, int32_t cap);

/* ---- host-pointer entry points (synchronous) --------------------------- */

/* gaussian_blur_f32, filter/mod.rs:156-164. p_type IPH_U8 or IPH_F32.
 * sigma <= 0 -> IPH_EINVAL "sigma must be > 0.0". */
IPH_API iph_status iph_gaussian_blur_f32(const void *src, void *dst, uint32_t width,
                                         uint32_t height, int32_t p_type, float sigma);

/* separable_filter, filter/mod.rs:169-183 (separable_filter_equal,
 * :188-195, passes the same kernel twice). (p_type, k_type) in
 * {(U8,F32), (U8,I32), (F32,F32)}. Unequal lengths -> IPH_EINVAL
 * "the two 1D kernels must be the same length". */
IPH_API iph_status iph_separable_filter(const void *src, void *dst, uint32_t width,
                                        uint32_t height, int32_t p_type, const void *h_kernel,
                                        uint32_t h_len, const void *v_kernel, uint32_t v_len,
                                        int32_t k_type);

/* horizontal_filter, filter/mod.rs:232-314; vertical_filter, :319-405. */
IPH_API iph_status iph_horizontal_filter(const void *src, void *dst, uint32_t width,
                                         uint32_t height, int32_t p_type, const void *kernel,
                                         uint32_t len, int32_t k_type);
IPH_API iph_status iph_vertical_filter(const void *src, void *dst, uint32_t width,
                                       uint32_t height, int32_t p_type, const void *kernel,
                                       uint32_t len, int32_t k_type);

/* filter_clamped, filter/mod.rs:202-210, and filter_clamped_parallel,
 * :214-226 (same entry: there is one GPU path). kernel is a row-major
 * Kernel<K> (kernel.rs:24-33): kw, kh > 0. Accepted (P, K, S):
 * (U8|F32, F32, U8|U16|F32) and (U8, I32, U8|U16|I16|I32). */
IPH_API iph_status iph_filter_clamped(const void *src, int32_t p_type, uint32_t width,
                                      uint32_t height, const void *kernel, int32_t k_type,
                                      uint32_t kw, uint32_t kh, void *dst, int32_t s_type);

/* integral_image::<_, u32> (squared = 0) and integral_squared_image::<_, u32>
 * (squared != 0), integral_image.rs:50-57, 88-95, 98-152, of a GrayImage:
 * out holds (width+1) x (height+1) u32, zero top row and left column, every
 * sum modulo 2^32 (the reference's u32 additions wrap in release builds). */
IPH_API iph_status iph_integral_image(const uint8_t *src, uint32_t width, uint32_t height,
                                      int32_t squared, uint32_t *out);
/* adaptive_threshold, contrast.rs:19-47: 255 where pixel >= floor(block
 * mean) - delta over the clipped (2r+1)^2 block, else 0. block_radius 0 ->
 * IPH_EINVAL "assertion failed: block_radius > 0"; block_radius above 2^31 - 1
 * (where the reference's i32 cast of it goes negative) -> IPH_EINVAL. */
IPH_API iph_status iph_adaptive_threshold(const uint8_t *src, uint8_t *dst, uint32_t width,
                                          uint32_t height, uint32_t block_radius, int32_t delta);
/* canny, edges.rs:30-155: 255 on edge pixels, 0 elsewhere. high < low (or a
 * NaN threshold) -> IPH_EINVAL "assertion failed: high_threshold >=
 * low_threshold"; an empty image -> IPH_EINVAL (the reference's height - 1
 * overflows); an edge that reaches the image border (possible only for
 * low_threshold <= 0) -> IPH_EINVAL, where the reference's search indexes
 * outside the image and panics. */
IPH_API iph_status iph_canny(const uint8_t *src, uint8_t *dst, uint32_t width, uint32_t height,
                             float low_threshold, float high_threshold);
/* gradients_grayscale / gradients, gradients.rs:15-21, 99-160, for a
 * GrayImage: filter_clamped::<_, i32, i16> with each kernel (row-major i32,
 * at most 49 taps each), then (dx^2 + dy^2).sqrt() as u16 per pixel
 * (sobel_gradients, :198-204, and prewitt_gradients, :207-213, pass the
 * kernel.rs constants). */
IPH_API iph_status iph_gradients(const uint8_t *src, uint32_t width, uint32_t height,
                                 const int32_t *h_kernel, uint32_t hkw, uint32_t hkh,
                                 const int32_t *v_kernel, uint32_t vkw, uint32_t vkh,
                                 uint16_t *out);
/* sharpen3x3, filter/sharpen.rs:13-16 (and sharpen3x3_parallel, :20-24):
 * filter_clamped with the i32 kernel identity minus the 3x3 Laplacian. */
IPH_API iph_status iph_sharpen3x3(const uint8_t *src, uint8_t *dst, uint32_t width,
                                  uint32_t height);
/* sharpen_gaussian, filter/sharpen.rs:31-38: the image as f32, its f32
 * gaussian_blur_f32 q, then clamp_u8((1 + amount) * p - amount * q). */
IPH_API iph_status iph_sharpen_gaussian(const uint8_t *src, uint8_t *dst, uint32_t width,
                                        uint32_t height, float sigma, float amount);

/* box_filter, filter/box_filter.rs:19-70 (GrayImage only). */
IPH_API iph_status iph_box_filter(const uint8_t *src, uint8_t *dst, uint32_t width,
                                  uint32_t height, uint32_t x_radius, uint32_t y_radius);

/* match_template, template_matching.rs:78-92, and match_template_parallel,
 * :96-110. p_type IPH_U8 (the reference's GrayImage) or IPH_F32 (an
 * extension: the same four formulas over Image<Luma<f32>>). out holds
 * (width-tw+1) x (height-th+1) f32. A template larger than the image ->
 * IPH_EINVAL "image width must be greater than or equal to template width"
 * (or height). */
IPH_API iph_status iph_match_template(const void *image, int32_t p_type, uint32_t width,
                                      uint32_t height, const void *tmpl, uint32_t tw,
                                      uint32_t th, int32_t method, int32_t path, float *out);

/* match_template_with_mask, template_matching.rs:125-140, and
 * match_template_with_mask_parallel, :142-163, over GrayImages (the masked
 * methods of :295-376). path IPH_PATH_DIRECT: the reference's operation order
 * (bit-identical f32); IPH_PATH_MFMA (CrossCorrelation and
 * CrossCorrelationNormalized, templates up to 128x128): sum t i m^2 and
 * sum (i m)^2 as BF16x3 MFMA correlations of the image and its square with
 * T M^2 and M^2, then the reference's f32 finish (the SSE methods ->
 * IPH_EUNSUPPORTED: their decomposed sum cancels at a match);
 * IPH_PATH_AUTO: MFMA where it applies, else direct. mask is mw x mh; unequal
 * to the template -> IPH_EINVAL "... the template and mask must be the same
 * size ..." (ImageTemplateMask::new, :430-437), checked before the
 * template-vs-image asserts. out holds (width-tw+1) x (height-th+1) f32. */
IPH_API iph_status iph_match_template_with_mask(const uint8_t *image, uint32_t width,
                                                uint32_t height, const uint8_t *tmpl,
                                                uint32_t tw, uint32_t th, const uint8_t *mask,
                                                uint32_t mw, uint32_t mh, int32_t method,
                                                int32_t path, float *out);

/* find_extremes, template_matching.rs:479-511, for Image<Luma<u8>> and
 * Image<Luma<f32>>. Empty image -> IPH_EINVAL "image must be non-empty". */
IPH_API iph_status iph_find_extremes(const void *image, int32_t p_type, uint32_t width,
                                     uint32_t height, iph_extremes *out);

/* ---- device-pointer entry points (asynchronous on `stream`) ------------ */

/* Batched gaussian_blur_f32: n images of width x height, image i at
 * src + i*image_stride elements (likewise dst). */
IPH_API size_t iph_dev_gaussian_blur_workspace(uint32_t width, uint32_t height, uint32_t n,
                                               int32_t p_type, float sigma);
IPH_API iph_status iph_dev_gaussian_blur_f32(const void *src, void *dst, uint32_t width,
                                             uint32_t height, uint32_t n, size_t image_stride,
                                             int32_t p_type, float sigma, void *workspace,
                                             size_t workspace_bytes, void *stream);

/* Batched separable_filter (same contract as iph_separable_filter). */
IPH_API size_t iph_dev_separable_filter_workspace(uint32_t width, uint32_t height, uint32_t n,
                                                  int32_t p_type, uint32_t len,
                                                  int32_t k_type);
IPH_API iph_status iph_dev_separable_filter(const void *src, void *dst, uint32_t width,
                                            uint32_t height, uint32_t n, size_t image_stride,
                                            int32_t p_type, const void *h_kernel,
                                            uint32_t h_len, const void *v_kernel,
                                            uint32_t v_len, int32_t k_type, void *workspace,
                                            size_t workspace_bytes, void *stream);

/* filter_clamped on device memory (kernel stays a host pointer). */
IPH_API size_t iph_dev_filter_clamped_workspace(int32_t k_type, uint32_t kw, uint32_t kh);
IPH_API iph_status iph_dev_filter_clamped(const void *src, int32_t p_type, uint32_t width,
                                          uint32_t height, const void *kernel, int32_t k_type,
                                          uint32_t kw, uint32_t kh, void *dst, int32_t s_type,
                                          void *workspace, size_t workspace_bytes,
                                          void *stream);

/* box_filter on device memory. */
IPH_API size_t iph_dev_box_filter_workspace(uint32_t width, uint32_t height);
IPH_API iph_status iph_dev_box_filter(const uint8_t *src, uint8_t *dst, uint32_t width,
                                      uint32_t height, uint32_t x_radius, uint32_t y_radius,
                                      void *workspace, size_t workspace_bytes, void *stream);

/* match_template_with_mask on device memory. */
/* iph_integral_image / iph_adaptive_threshold on device buffers; workspace of
 * the bytes the _workspace call returns, out (w+1) x (h+1) u32. */
IPH_API size_t iph_dev_integral_image_workspace(uint32_t width, uint32_t height);
IPH_API iph_status iph_dev_integral_image(const uint8_t *src, uint32_t width, uint32_t height,
                                          int32_t squared, uint32_t *out, void *workspace,
                                          size_t workspace_bytes, void *stream);
/* 0 for radii up to 64 (block sums in LDS); larger radii need an integral
 * table of (w+1) x (h+1) u32 plus its scratch. */
IPH_API size_t iph_dev_adaptive_threshold_workspace(uint32_t width, uint32_t height,
                                                    uint32_t block_radius);
IPH_API iph_status iph_dev_adaptive_threshold(const uint8_t *src, uint8_t *dst, uint32_t width,
                                              uint32_t height, uint32_t block_radius,
                                              int32_t delta, void *workspace,
                                              size_t workspace_bytes, void *stream);
/* iph_canny on device buffers. Blocks the calling thread: the hysteresis
 * reads a convergence flag back every few launches on `stream`. */
IPH_API size_t iph_dev_canny_workspace(uint32_t width, uint32_t height);
IPH_API iph_status iph_dev_canny(const uint8_t *src, uint8_t *dst, uint32_t width,
                                 uint32_t height, float low_threshold, float high_threshold,
                                 void *workspace, size_t workspace_bytes, void *stream);
/* iph_gradients and iph_sharpen_gaussian on device buffers (kernels of
 * iph_dev_gradients stay host arrays, copied into the launch). */
IPH_API iph_status iph_dev_gradients(const uint8_t *src, uint32_t width, uint32_t height,
                                     const int32_t *h_kernel, uint32_t hkw, uint32_t hkh,
                                     const int32_t *v_kernel, uint32_t vkw, uint32_t vkh,
                                     uint16_t *out, void *stream);
/* 0 when the blur has at most 48 taps (sigma <= 11.5: one fused pass, the
 * workspace may be NULL); else two f32 images plus the blur's workspace. */
IPH_API size_t iph_dev_sharpen_gaussian_workspace(uint32_t width, uint32_t height, float sigma);
IPH_API iph_status iph_dev_sharpen_gaussian(const uint8_t *src, uint8_t *dst, uint32_t width,
                                            uint32_t height, float sigma, float amount,
                                            void *workspace, size_t workspace_bytes, void *stream);
IPH_API size_t iph_dev_match_template_with_mask_workspace(uint32_t width, uint32_t height,
                                                          uint32_t tw, uint32_t th,
                                                          int32_t method, int32_t path);
IPH_API iph_status iph_dev_match_template_with_mask(const uint8_t *image, uint32_t width,
                                                    uint32_t height, const uint8_t *tmpl,
                                                    uint32_t tw, uint32_t th,
                                                    const uint8_t *mask, uint32_t mw,
                                                    uint32_t mh, int32_t method, int32_t path,
                                                    float *out, void *workspace,
                                                    size_t workspace_bytes, void *stream);

/* match_template on device memory. */
IPH_API size_t iph_dev_match_template_workspace(int32_t p_type, uint32_t width,
                                                uint32_t height, uint32_t tw, uint32_t th,
                                                int32_t method, int32_t path);
IPH_API iph_status iph_dev_match_template(const void *image, int32_t p_type, uint32_t width,
                                          uint32_t height, const void *tmpl, uint32_t tw,
                                          uint32_t th, int32_t method, int32_t path, float *out,
                                          void *workspace, size_t workspace_bytes,
                                          void *stream);

/* find_extremes on device memory; *out is device memory. */
IPH_API size_t iph_dev_find_extremes_workspace(void);
IPH_API iph_status iph_dev_find_extremes(const void *image, int32_t p_type, uint32_t width,
                                         uint32_t height, iph_extremes *out, void *workspace,
                                         size_t workspace_bytes, void *stream);

/* find_extremes on device memory without allocating: no counterpart in the
 * reference. For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). The
 * per-workgroup partial results live in the library's own 16 KiB device
 * array (one per device, held until the process ends); every image the
 * workspace entry above takes (up to 2^32 - 1 pixels) fits, larger ->
 * IPH_EINVAL. Calls no hipMalloc/hipFree. All calls on one device share the
 * array: calls that overlap on one device (other streams or threads) race.
 * Same result as iph_dev_find_extremes; *out is device memory. */
IPH_API iph_status iph_dev_find_extremes_noalloc(const void *image, int32_t p_type,
                                                 uint32_t width, uint32_t height,
                                                 iph_extremes *out, void *stream);

/*
 * Batches of uneven GrayImages. The reference has no batch API: a batch
 * replaces a loop of per-image calls of gaussian_blur_f32
 * (filter/mod.rs:156-164), filter_clamped::<Luma<u8>, f32, u8>
 * (filter/mod.rs:202-210) and match_template (template_matching.rs:78-92),
 * and each job's output is bit-identical to that call's. All pointers are
 * device memory that stays valid while the batch is used; outputs must not
 * overlap inputs of other jobs.
 *
 * iph_batch_plan allocates the batch's device memory, autotunes the launch
 * shapes when `autotune` is non-zero (it runs the jobs a few times to time
 * them), and captures every job into one hipGraph. iph_batch_run replays that
 * graph on `stream`; iph_batch_run_eager enqueues the same launches without
 * the graph. Adding a job after iph_batch_plan discards the plan. A batch is
 * not thread-safe, and runs of one batch must not overlap.
 */
typedef struct iph_batch iph_batch;
IPH_API iph_status iph_batch_create(iph_batch **out);
IPH_API void iph_batch_destroy(iph_batch *batch);
/* sigma in (0, 8] (at most 33 taps). */
IPH_API iph_status iph_batch_add_gaussian_blur(iph_batch *batch, const uint8_t *src,
                                               uint8_t *dst, uint32_t width, uint32_t height,
                                               float sigma);
/* kernel: kw * kh f32 weights, row-major, host memory; at most 49 taps. */
IPH_API iph_status iph_batch_add_filter_clamped(iph_batch *batch, const uint8_t *src,
                                                uint8_t *dst, uint32_t width, uint32_t height,
                                                const float *kernel, uint32_t kw, uint32_t kh);
/* any template the single call takes (above 128x128 on the direct path);
 * out: (width - tw + 1) x (height - th + 1) f32. */
IPH_API iph_status iph_batch_add_match_template(iph_batch *batch, const uint8_t *image,
                                                uint32_t width, uint32_t height,
                                                const uint8_t *tmpl, uint32_t tw, uint32_t th,
                                                int32_t method, float *out);
IPH_API iph_status iph_batch_plan(iph_batch *batch, int32_t autotune, void *stream);
IPH_API iph_status iph_batch_run(iph_batch *batch, void *stream);
IPH_API iph_status iph_batch_run_eager(iph_batch *batch, void *stream);
/* The autotuner's measurements and choices, as text. */
IPH_API iph_status iph_batch_tuning(iph_batch *batch, char *buf, size_t cap);

/*
 * Learned match proposal: no counterpart in the reference, whose only search
 * is the full score map of match_template (template_matching.rs:78-92) read
 * by find_extremes (template_matching.rs:479-511). These are the building
 * blocks of a cheaper search that ends in the same argmax: a 4x-downsampled
 * coarse map, its best tiles, a small convolutional network that ranks
 * candidate places, and exact scores in windows around the best of them
 * (values identical to the full map's at the same places). All pointers are
 * device memory; everything is enqueued on `stream`.
 *
 * Network tensors: activations and gradients are bf16 (uint16_t bit
 * patterns), NHWC, IPH_CNN_PATCH x IPH_CNN_PATCH pixels by IPH_CNN_CHANNELS
 * channels per patch; a layer's weights are bf16
 * w[co][(ky * 3 + kx) * IPH_CNN_CHANNELS + ci], its bias f32[channels].
 */
#define IPH_CNN_CHANNELS 32
#define IPH_CNN_PATCH 16
#define IPH_CNN_LAYERS 4
/* y = conv3x3(x, w) + bias (zero padding) over n_patches patches; z (f32,
 * may be NULL) gets y, a (bf16, may be NULL) relu(y) if relu else y. */
IPH_API iph_status iph_dev_cnn_conv3x3(const uint16_t *x, const uint16_t *w, const float *bias,
                                       int32_t relu, int32_t n_patches, float *z, uint16_t *a,
                                       void *stream);
/* Weight and bias gradients of that layer from its input x and output
 * gradient dz (f32 dw[co][k], db[co]), summed over n_partials workgroups. */
IPH_API size_t iph_dev_cnn_conv3x3_dw_workspace(int32_t n_partials);
IPH_API iph_status iph_dev_cnn_conv3x3_dw(const uint16_t *x, const uint16_t *dz,
                                          int32_t n_patches, int32_t n_partials, void *workspace,
                                          size_t workspace_bytes, float *dw, float *db,
                                          void *stream);
/* Weights of the layer whose forward pass is the input gradient of w. */
IPH_API iph_status iph_dev_cnn_flip_weights(const uint16_t *w, uint16_t *w_flipped, void *stream);
/* dz = bf16(dx * (z > 0)) over n elements. */
IPH_API iph_status iph_dev_cnn_relu_backward(const float *dx, const float *z, size_t n,
                                             uint16_t *dz, void *stream);
/* Softmax cross-entropy of each pair over the channel-0 logits of its
 * patches_per_pair patches; target[pair] = patch * 256 + pixel or -1. dz_last
 * gets (softmax - onehot) * scale in channel 0; loss[pair] the loss. */
IPH_API iph_status iph_dev_cnn_softmax_xent(const float *z_last, int32_t n_pairs,
                                            int32_t patches_per_pair, const int32_t *target,
                                            float scale, float *loss, uint16_t *dz_last,
                                            void *stream);
/* One Adam step (step >= 1) on n f32 parameters. */
IPH_API iph_status iph_dev_cnn_adam(float *param, const float *grad, float *m, float *v,
                                    int32_t n, float lr, float beta1, float beta2, float eps,
                                    int32_t step, void *stream);
IPH_API iph_status iph_dev_cnn_to_bf16(const float *src, int32_t n, uint16_t *dst, void *stream);
/* dst (w/4) x (h/4): rounded mean of each full 4x4 block. */
IPH_API iph_status iph_dev_downsample4_u8(const uint8_t *src, uint32_t width, uint32_t height,
                                          uint8_t *dst, void *stream);
/* The k best (1 <= k <= 256) 16x16-tile maxima of a coarse map (larger
 * first; ties: lower index; NaN last): cand[2i], cand[2i+1] = x, y;
 * cand_score[i]; (0, 0) and -inf past the number of tiles. */
IPH_API size_t iph_dev_proposal_candidates_workspace(uint32_t cw, uint32_t ch);
IPH_API iph_status iph_dev_proposal_candidates(const float *coarse, uint32_t cw, uint32_t ch,
                                               int32_t k, int32_t *cand, float *cand_score,
                                               void *workspace, size_t workspace_bytes,
                                               void *stream);
/* Network input: one 16x16 patch per candidate, centred on it. */
IPH_API iph_status iph_dev_proposal_features(const float *coarse, uint32_t cw, uint32_t ch,
                                             const uint8_t *small, uint32_t sw, uint32_t sh,
                                             uint32_t small_tw, const int32_t *cand,
                                             const float *cand_score, int32_t k,
                                             uint16_t *patches, void *stream);
/* The n_top (1 to 64) best logits (channel 0 of z_last, k <= 256 patches)
 * at distinct coarse places inside the map (centers: x, y; ties: lower
 * patch * 256 + pixel) and their softmax probabilities over all k * 256
 * logits; (0, 0) and 0 past the last such place. */
IPH_API size_t iph_dev_proposal_top_workspace(int32_t k, int32_t n_top);
IPH_API iph_status iph_dev_proposal_top(const float *z_last, int32_t k, const int32_t *cand,
                                        uint32_t cw, uint32_t ch, int32_t n_top, int32_t *centers,
                                        float *probs, void *workspace, size_t workspace_bytes,
                                        void *stream);
/* Exact match_template scores (the full map's values) at every place within
 * 8 pixels of each centre * scale, then result[0..3] = x, y of the best
 * (find_extremes' order), its score, and the best score more than `sep`
 * pixels from it. For the correlation methods (larger is better). */
IPH_API size_t iph_dev_proposal_refine_workspace(int32_t n_centers);
IPH_API iph_status iph_dev_proposal_refine(const uint8_t *image, uint32_t width, uint32_t height,
                                           const uint8_t *tmpl, uint32_t tw, uint32_t th,
                                           int32_t method, const int32_t *centers,
                                           int32_t n_centers, int32_t scale, int32_t sep,
                                           float *result, void *workspace, size_t workspace_bytes,
                                           void *stream);
/* The whole search on a u8 image and a 4..128 px template (CrossCorrelation
 * or CrossCorrelationNormalized), enqueued on `stream`: downsampling, the
 * coarse map, k candidates, the network (weights: IPH_CNN_LAYERS blocks of
 * 32 x 288 bf16, bias: IPH_CNN_LAYERS x 32 f32; both NULL ranks by coarse
 * score instead), exact windows around the n_top (<= k) best places.
 * result[0..4] = x, y, best score, best score more than `sep` px away, top
 * probability (1 without the network): the composition of the functions
 * above, value for value. */
/* Copies n floats of a result (e.g. the search's 5) to host memory once the
 * stream's work before it has run, and waits for the copy. */
IPH_API iph_status iph_dev_read_result(const float *result, float *host, size_t n, void *stream);
IPH_API size_t iph_dev_proposal_search_workspace(uint32_t width, uint32_t height, uint32_t tw,
                                                 uint32_t th, int32_t method, int32_t k,
                                                 int32_t n_top);
IPH_API iph_status iph_dev_proposal_search(const uint8_t *image, uint32_t width, uint32_t height,
                                           const uint8_t *tmpl, uint32_t tw, uint32_t th,
                                           int32_t method, const uint16_t *weights,
                                           const float *bias, int32_t k, int32_t n_top,
                                           int32_t sep, float *result, void *workspace,
                                           size_t workspace_bytes, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* IMAGEPROC_HIP_H */
