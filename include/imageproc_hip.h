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
    IPH_PATH_AUTO = 0,    /* MFMA path */
    IPH_PATH_MFMA = 1,    /* u8: int8 MFMA, exact i32 sums; f32: BF16x3 MFMA, FP32 accumulation */
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
/* templates up to 128x128; out: (width - tw + 1) x (height - th + 1) f32. */
IPH_API iph_status iph_batch_add_match_template(iph_batch *batch, const uint8_t *image,
                                                uint32_t width, uint32_t height,
                                                const uint8_t *tmpl, uint32_t tw, uint32_t th,
                                                int32_t method, float *out);
IPH_API iph_status iph_batch_plan(iph_batch *batch, int32_t autotune, void *stream);
IPH_API iph_status iph_batch_run(iph_batch *batch, void *stream);
IPH_API iph_status iph_batch_run_eager(iph_batch *batch, void *stream);
/* The autotuner's measurements and choices, as text. */
IPH_API iph_status iph_batch_tuning(iph_batch *batch, char *buf, size_t cap);

#ifdef __cplusplus
}
#endif
#endif /* IMAGEPROC_HIP_H */
