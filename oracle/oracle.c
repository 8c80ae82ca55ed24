/*
 * CPU oracle for the sliding-window correlation path of imageproc 0.27.0.
 *
 * TEST INFRASTRUCTURE ONLY: the checker used by tests/, by the smoke check and
 * by the CPU-baseline leg of the benchmark. The product never calls it.
 *
 * Build: oracle/Makefile (gcc -O3 -march=x86-64 -ffp-contract=off, no
 * fast-math), so every f32 operation below is one IEEE binary32 operation in
 * source order, like the reference's Rust code (rustc never contracts a*b+c
 * into an FMA). expf/sqrtf come from glibc libm, which is what Rust's
 * f32::exp / f32::sqrt call on x86_64-unknown-linux-gnu.
 *
 * File:line citations refer to the imageproc 0.27.0 source tree.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* Clamp<T> casts, definitions.rs:114-190                              */
/* ------------------------------------------------------------------ */

/* implement_clamp!(f32, u8, ...), definitions.rs:120-136,153:
 * x < 255 ? (x > 0 ? x as u8 : 0) : 255. `as u8` truncates toward zero; a NaN
 * fails the first test and becomes 255. */
static inline uint8_t clamp_f32_u8(float x) {
    if (x < 255.0f) {
        if (x > 0.0f) return (uint8_t)x;
        return 0;
    }
    return 255;
}
static inline uint16_t clamp_f32_u16(float x) { /* definitions.rs:165-172 */
    if (x < 65535.0f) {
        if (x > 0.0f) return (uint16_t)x;
        return 0;
    }
    return 65535;
}
static inline uint8_t clamp_i32_u8(int32_t x) { /* definitions.rs:151 */
    if (x < 255) {
        if (x > 0) return (uint8_t)x;
        return 0;
    }
    return 255;
}
static inline uint16_t clamp_i32_u16(int32_t x) { /* definitions.rs:156-163 */
    if (x < 65535) {
        if (x > 0) return (uint16_t)x;
        return 0;
    }
    return 65535;
}
static inline int16_t clamp_i32_i16(int32_t x) { /* definitions.rs:181-188 */
    if (x < 32767) {
        if (x > -32768) return (int16_t)x;
        return -32768;
    }
    return 32767;
}

/* Wrapping i32 multiply-add: the reference's `acc + c.into() * weight` on i32
 * (filter/mod.rs:407-416) wraps in a release build. */
static inline int32_t wrap_mac_i32(int32_t acc, int32_t c, int32_t w) {
    return (int32_t)((uint32_t)acc + (uint32_t)c * (uint32_t)w);
}

static inline int64_t clampi64(int64_t v, int64_t lo, int64_t hi) {
    return v < lo ? lo : (v > hi ? hi : v);
}

/* ------------------------------------------------------------------ */
/* Gaussian kernel, filter/mod.rs:125-145                              */
/* ------------------------------------------------------------------ */

/* gaussian_pdf, filter/mod.rs:125-129:
 * (sigma * (2 PI).sqrt()).recip() * (-x.powi(2) / (2.0 * sigma.powi(2))).exp()
 * with PI = f32::consts::PI; powi(2) is a single multiply. */
static float gaussian_pdf(float x, float sigma) {
    const float pi_f32 = 3.14159274101257324219f;
    float two_pi = 2.0f * pi_f32;
    float norm = sigma * sqrtf(two_pi);
    float recip = 1.0f / norm;
    float x2 = x * x;
    float s2 = sigma * sigma;
    float den = 2.0f * s2;
    float arg = -x2 / den;
    return recip * expf(arg);
}

int or_gaussian_kernel_f32(float sigma, float *out, int cap) {
    /* filter/mod.rs:133: radius = (2.0 * sigma).ceil() as usize */
    float r_f = ceilf(2.0f * sigma);
    if (!(r_f >= 0.0f)) r_f = 0.0f;
    if (r_f > 1.0e7f) return -1;
    int radius = (int)r_f;
    int len = 2 * radius + 1;
    float *data = (float *)malloc(sizeof(float) * (size_t)len);
    if (!data) return -1;
    for (int i = 0; i <= radius; i++) { /* filter/mod.rs:135-139 */
        float v = gaussian_pdf((float)i, sigma);
        data[radius + i] = v;
        data[radius - i] = v;
    }
    /* filter/mod.rs:140: iter().sum() is a sequential f32 fold starting at -0.0 */
    float sum = -0.0f;
    for (int i = 0; i < len; i++) sum = sum + data[i];
    for (int i = 0; i < len; i++) data[i] = data[i] / sum; /* filter/mod.rs:141 */
    int n = len < cap ? len : cap;
    if (out && n > 0) memcpy(out, data, sizeof(float) * (size_t)n);
    free(data);
    return len;
}

/* ------------------------------------------------------------------ */
/* 1-D filters, filter/mod.rs:228-405                                  */
/* ------------------------------------------------------------------ */
/* The reference splits each row into left margin / interior / right margin
 * (filter/mod.rs:269-311) and has a fully clamped branch for kernels at least
 * as long as the side (filter/mod.rs:250-267). All branches read
 * image[clamp(x + i - len/2, 0, W-1)], i ascending; the reference's own tests
 * pin that equivalence against horizontal_filter_reference
 * (filter/mod.rs:516-613), which is the form restated here. */

#define DEFINE_HFILTER(NAME, PT, KT, ACC_ZERO, MAC, CLAMP)                                      \
    static void NAME(const PT *src, PT *dst, uint32_t w, uint32_t h, const KT *k,            \
                     uint32_t klen) {                                                       \
        int64_t half = (int64_t)(klen / 2);                                                  \
        for (uint32_t y = 0; y < h; y++) {                                                   \
            const PT *row = src + (size_t)y * w;                                             \
            for (uint32_t x = 0; x < w; x++) {                                               \
                KT acc = ACC_ZERO;                                                           \
                for (uint32_t i = 0; i < klen; i++) {                                        \
                    int64_t xp = clampi64((int64_t)x + (int64_t)i - half, 0, (int64_t)w - 1); \
                    acc = MAC(acc, row[xp], k[i]);                                           \
                }                                                                            \
                dst[(size_t)y * w + x] = CLAMP(acc);                                         \
            }                                                                                \
        }                                                                                    \
    }

#define DEFINE_VFILTER(NAME, PT, KT, ACC_ZERO, MAC, CLAMP)                                      \
    static void NAME(const PT *src, PT *dst, uint32_t w, uint32_t h, const KT *k,            \
                     uint32_t klen) {                                                       \
        int64_t half = (int64_t)(klen / 2);                                                  \
        for (uint32_t y = 0; y < h; y++) {                                                   \
            for (uint32_t x = 0; x < w; x++) {                                               \
                KT acc = ACC_ZERO;                                                           \
                for (uint32_t i = 0; i < klen; i++) {                                        \
                    int64_t yp = clampi64((int64_t)y + (int64_t)i - half, 0, (int64_t)h - 1); \
                    acc = MAC(acc, src[(size_t)yp * w + x], k[i]);                           \
                }                                                                            \
                dst[(size_t)y * w + x] = CLAMP(acc);                                         \
            }                                                                                \
        }                                                                                    \
    }

/* accumulate, filter/mod.rs:407-416: *a = *a + c.into() * weight */
#define MAC_F32(acc, c, w) ((acc) + (float)(c) * (w))
#define MAC_I32(acc, c, w) wrap_mac_i32((acc), (int32_t)(c), (w))
#define ID_F32(x) (x)

DEFINE_HFILTER(hfilter_u8_f32, uint8_t, float, 0.0f, MAC_F32, clamp_f32_u8)
DEFINE_VFILTER(vfilter_u8_f32, uint8_t, float, 0.0f, MAC_F32, clamp_f32_u8)
DEFINE_HFILTER(hfilter_u8_i32, uint8_t, int32_t, 0, MAC_I32, clamp_i32_u8)
DEFINE_VFILTER(vfilter_u8_i32, uint8_t, int32_t, 0, MAC_I32, clamp_i32_u8)
DEFINE_HFILTER(hfilter_f32_f32, float, float, 0.0f, MAC_F32, ID_F32)
DEFINE_VFILTER(vfilter_f32_f32, float, float, 0.0f, MAC_F32, ID_F32)

int or_horizontal_filter(const void *src, void *dst, uint32_t w, uint32_t h, int p_type,
                         const void *kern, uint32_t klen, int k_type) {
    if (p_type == OR_U8 && k_type == OR_F32)
        hfilter_u8_f32(src, dst, w, h, kern, klen);
    else if (p_type == OR_U8 && k_type == OR_I32)
        hfilter_u8_i32(src, dst, w, h, kern, klen);
    else if (p_type == OR_F32 && k_type == OR_F32)
        hfilter_f32_f32(src, dst, w, h, kern, klen);
    else
        return -1;
    return 0;
}

int or_vertical_filter(const void *src, void *dst, uint32_t w, uint32_t h, int p_type,
                       const void *kern, uint32_t klen, int k_type) {
    if (p_type == OR_U8 && k_type == OR_F32)
        vfilter_u8_f32(src, dst, w, h, kern, klen);
    else if (p_type == OR_U8 && k_type == OR_I32)
        vfilter_u8_i32(src, dst, w, h, kern, klen);
    else if (p_type == OR_F32 && k_type == OR_F32)
        vfilter_f32_f32(src, dst, w, h, kern, klen);
    else
        return -1;
    return 0;
}

static size_t elem_size(int t) {
    switch (t) {
    case OR_U8: return 1;
    case OR_U16: case OR_I16: return 2;
    default: return 4;
    }
}

int or_separable_filter(const void *src, void *dst, uint32_t w, uint32_t h, int p_type,
                        const void *hk, uint32_t hlen, const void *vk, uint32_t vlen,
                        int k_type) {
    /* filter/mod.rs:174-180: assert_eq!(h_kernel.len(), v_kernel.len()) */
    if (hlen != vlen) return -2;
    size_t n = (size_t)w * h * elem_size(p_type);
    void *mid = malloc(n ? n : 1);
    if (!mid) return -4;
    /* filter/mod.rs:181-182: H pass, then V pass on the P-typed result */
    int rc = or_horizontal_filter(src, mid, w, h, p_type, hk, hlen, k_type);
    if (rc == 0) rc = or_vertical_filter(mid, dst, w, h, p_type, vk, vlen, k_type);
    free(mid);
    return rc;
}

int or_gaussian_blur_f32(const void *src, void *dst, uint32_t w, uint32_t h, int p_type,
                         float sigma) {
    /* filter/mod.rs:161: assert!(sigma > 0.0, "sigma must be > 0.0") */
    if (!(sigma > 0.0f)) return -3;
    int len = or_gaussian_kernel_f32(sigma, NULL, 0);
    if (len < 0) return -4;
    float *k = (float *)malloc(sizeof(float) * (size_t)len);
    if (!k) return -4;
    or_gaussian_kernel_f32(sigma, k, len);
    /* filter/mod.rs:163 -> separable_filter_equal -> separable_filter(k, k) */
    int rc = or_separable_filter(src, dst, w, h, p_type, k, (uint32_t)len, k, (uint32_t)len,
                                 OR_F32);
    free(k);
    return rc;
}

/* ------------------------------------------------------------------ */
/* 2-D filter, filter/mod.rs:24-69 and 197-226                          */
/* ------------------------------------------------------------------ */

static int clamped_triple_ok(int p, int k, int s) {
    /* filter_clamped needs K: From<P::Subpixel> and S: Clamp<K>
     * (filter/mod.rs:202-208, definitions.rs:151-190). */
    if (k == OR_F32) return (p == OR_U8 || p == OR_F32) && (s == OR_U8 || s == OR_U16 || s == OR_F32);
    if (k == OR_I32) return p == OR_U8 && (s == OR_U8 || s == OR_U16 || s == OR_I16 || s == OR_I32);
    return 0;
}

static inline float load_f(const void *p, int t, size_t i) {
    return t == OR_U8 ? (float)((const uint8_t *)p)[i] : ((const float *)p)[i];
}

static void store_s_f32(void *dst, int s, size_t i, float acc) {
    switch (s) {
    case OR_U8: ((uint8_t *)dst)[i] = clamp_f32_u8(acc); break;
    case OR_U16: ((uint16_t *)dst)[i] = clamp_f32_u16(acc); break;
    default: ((float *)dst)[i] = acc; break;
    }
}
static void store_s_i32(void *dst, int s, size_t i, int32_t acc) {
    switch (s) {
    case OR_U8: ((uint8_t *)dst)[i] = clamp_i32_u8(acc); break;
    case OR_U16: ((uint16_t *)dst)[i] = clamp_i32_u16(acc); break;
    case OR_I16: ((int16_t *)dst)[i] = clamp_i32_i16(acc); break;
    default: ((int32_t *)dst)[i] = acc; break;
    }
}

/* One output row of filter(), filter/mod.rs:47-66: ky outer, kx inner,
 * y_p = clamp(y + ky - kh/2), x_p = clamp(x + kx - kw/2), row-major kernel. */
static void filter_row(const void *src, int p, uint32_t w, uint32_t h, const void *kern, int k,
                       uint32_t kw, uint32_t kh, void *dst, int s, uint32_t y) {
    int64_t W = w, H = h, KW = kw, KH = kh;
    for (uint32_t x = 0; x < w; x++) {
        if (k == OR_F32) {
            float acc = 0.0f;
            for (int64_t ky = 0; ky < KH; ky++) {
                int64_t yp = clampi64((int64_t)y + ky - KH / 2, 0, H - 1);
                for (int64_t kx = 0; kx < KW; kx++) {
                    int64_t xp = clampi64((int64_t)x + kx - KW / 2, 0, W - 1);
                    acc = acc + load_f(src, p, (size_t)yp * w + xp) *
                                    ((const float *)kern)[ky * KW + kx];
                }
            }
            store_s_f32(dst, s, (size_t)y * w + x, acc);
        } else {
            int32_t acc = 0;
            for (int64_t ky = 0; ky < KH; ky++) {
                int64_t yp = clampi64((int64_t)y + ky - KH / 2, 0, H - 1);
                for (int64_t kx = 0; kx < KW; kx++) {
                    int64_t xp = clampi64((int64_t)x + kx - KW / 2, 0, W - 1);
                    acc = wrap_mac_i32(acc, ((const uint8_t *)src)[(size_t)yp * w + xp],
                                       ((const int32_t *)kern)[ky * KW + kx]);
                }
            }
            store_s_i32(dst, s, (size_t)y * w + x, acc);
        }
    }
}

int or_filter_clamped(const void *src, int p_type, uint32_t w, uint32_t h, const void *kern,
                      int k_type, uint32_t kw, uint32_t kh, void *dst, int s_type) {
    if (!clamped_triple_ok(p_type, k_type, s_type)) return -1;
    if (kw == 0 || kh == 0) return -2; /* Kernel::new, kernel.rs:25 */
    for (uint32_t y = 0; y < h; y++)
        filter_row(src, p_type, w, h, kern, k_type, kw, kh, dst, s_type, y);
    return 0;
}

typedef struct {
    const void *src; int p; uint32_t w, h; const void *kern; int k; uint32_t kw, kh;
    void *dst; int s; int tid, nthreads;
} filter_job;

static void *filter_worker(void *arg) {
    filter_job *j = (filter_job *)arg;
    for (uint32_t y = (uint32_t)j->tid; y < j->h; y += (uint32_t)j->nthreads)
        filter_row(j->src, j->p, j->w, j->h, j->kern, j->k, j->kw, j->kh, j->dst, j->s, y);
    return NULL;
}

int or_filter_clamped_parallel(const void *src, int p_type, uint32_t w, uint32_t h,
                               const void *kern, int k_type, uint32_t kw, uint32_t kh,
                               void *dst, int s_type, int nthreads) {
    if (!clamped_triple_ok(p_type, k_type, s_type)) return -1;
    if (kw == 0 || kh == 0) return -2;
    if (nthreads < 1) nthreads = 1;
    pthread_t th[256];
    filter_job jobs[256];
    if (nthreads > 256) nthreads = 256;
    for (int t = 0; t < nthreads; t++) {
        jobs[t] = (filter_job){src, p_type, w, h, kern, k_type, kw, kh, dst, s_type, t, nthreads};
        pthread_create(&th[t], NULL, filter_worker, &jobs[t]);
    }
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    return 0;
}

/* ------------------------------------------------------------------ */
/* box_filter, filter/box_filter.rs:19-70                              */
/* ------------------------------------------------------------------ */

/* row_running_sum, integral_image.rs:348-380: running sum of the row padded
 * by `padding` copies of the first and last pixel; wrapping u32. */
static void row_running_sum(const uint8_t *img, uint32_t w, uint32_t row, uint32_t *buf,
                            uint32_t padding) {
    const uint8_t *r = img + (size_t)row * w;
    uint32_t first = r[0], last = r[w - 1], sum = 0;
    size_t k = 0;
    for (uint32_t i = 0; i < padding; i++) { sum += first; buf[k++] = sum; }
    for (uint32_t i = 0; i < w; i++) { sum += r[i]; buf[k++] = sum; }
    for (uint32_t i = 0; i < padding; i++) { sum += last; buf[k++] = sum; }
}

/* column_running_sum, integral_image.rs:413-463 */
static void column_running_sum(const uint8_t *img, uint32_t w, uint32_t h, uint32_t col,
                               uint32_t *buf, uint32_t padding) {
    uint32_t first = img[col], last = img[(size_t)(h - 1) * w + col], sum = 0;
    size_t k = 0;
    for (uint32_t i = 0; i < padding; i++) { sum += first; buf[k++] = sum; }
    for (uint32_t y = 0; y < h; y++) { sum += img[(size_t)y * w + col]; buf[k++] = sum; }
    for (uint32_t i = 0; i < padding; i++) { sum += last; buf[k++] = sum; }
}

void or_box_filter(const uint8_t *src, uint8_t *dst, uint32_t w, uint32_t h, uint32_t rx,
                   uint32_t ry) {
    if (w == 0 || h == 0) return; /* box_filter.rs:22-24 */
    uint32_t kw = 2 * rx + 1, kh = 2 * ry + 1;
    uint32_t *rb = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)w + 2 * (size_t)rx));
    for (uint32_t y = 0; y < h; y++) { /* box_filter.rs:30-48 */
        row_running_sum(src, w, y, rb, rx);
        dst[(size_t)y * w] = (uint8_t)(rb[2 * (size_t)rx] / kw);
        for (uint32_t x = 1; x < w; x++) {
            size_t u = (size_t)x + 2 * (size_t)rx, l = (size_t)x - 1;
            dst[(size_t)y * w + x] = (uint8_t)((rb[u] - rb[l]) / kw);
        }
    }
    free(rb);
    uint32_t *cb = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)h + 2 * (size_t)ry));
    for (uint32_t x = 0; x < w; x++) { /* box_filter.rs:50-67, in place on `out` */
        column_running_sum(dst, w, h, x, cb, ry);
        dst[x] = (uint8_t)(cb[2 * (size_t)ry] / kh);
        for (uint32_t y = 1; y < h; y++) {
            size_t u = (size_t)y + 2 * (size_t)ry, l = (size_t)y - 1;
            dst[(size_t)y * w + x] = (uint8_t)((cb[u] - cb[l]) / kh);
        }
    }
    free(cb);
}

/* ------------------------------------------------------------------ */
/* Template matching, template_matching.rs                              */
/* ------------------------------------------------------------------ */

/* square_sum, template_matching.rs:378-380: sequential f32 sum of x*x. */
static float square_sum_u8(const uint8_t *t, size_t n) {
    float s = -0.0f;
    for (size_t i = 0; i < n; i++) s = s + (float)t[i] * (float)t[i];
    return s;
}
static float square_sum_f32(const float *t, size_t n) {
    float s = -0.0f;
    for (size_t i = 0; i < n; i++) s = s + t[i] * t[i];
    return s;
}

/* score_at for the four unmasked methods, template_matching.rs:213-293, with
 * slide_window_at's dy-outer/dx-inner walk, template_matching.rs:405-417. */
#define DEFINE_SCORE(NAME, PT)                                                                \
    static float NAME(const PT *img, uint32_t w, const PT *tpl, uint32_t tw, uint32_t th,    \
                      int method, float tss, uint32_t x, uint32_t y) {                       \
        float score = 0.0f, ii = 0.0f;                                                       \
        for (uint32_t dy = 0; dy < th; dy++) {                                               \
            const PT *irow = img + (size_t)(y + dy) * w + x;                                 \
            const PT *trow = tpl + (size_t)dy * tw;                                          \
            for (uint32_t dx = 0; dx < tw; dx++) {                                           \
                float i = (float)irow[dx], t = (float)trow[dx];                              \
                switch (method) {                                                            \
                case OR_SSE: { float d = t - i; score = score + d * d; } break;              \
                case OR_SSE_NORMED: { float d = t - i; score = score + d * d;                \
                                      ii = ii + i * i; } break;                              \
                case OR_CCORR: score = score + i * t; break;                                 \
                default: score = score + i * t; ii = ii + i * i; break;                      \
                }                                                                            \
            }                                                                                \
        }                                                                                    \
        if (method == OR_SSE_NORMED || method == OR_CCORR_NORMED) {                          \
            /* template_matching.rs:249-250, 289-291 */                                      \
            float norm = sqrtf(ii * tss);                                                    \
            return norm > 0.0f ? score / norm : score;                                       \
        }                                                                                    \
        return score;                                                                        \
    }

DEFINE_SCORE(score_u8, uint8_t)
DEFINE_SCORE(score_f32, float)

typedef struct {
    const void *img; uint32_t w; const void *tpl; uint32_t tw, th; int method; float tss;
    float *out; uint32_t ow, oh; int is_f32; int tid, nthreads;
} match_job;

static void match_rows(match_job *j) {
    for (uint32_t y = (uint32_t)j->tid; y < j->oh; y += (uint32_t)j->nthreads)
        for (uint32_t x = 0; x < j->ow; x++)
            j->out[(size_t)y * j->ow + x] =
                j->is_f32 ? score_f32(j->img, j->w, j->tpl, j->tw, j->th, j->method, j->tss, x, y)
                          : score_u8(j->img, j->w, j->tpl, j->tw, j->th, j->method, j->tss, x, y);
}
static void *match_worker(void *arg) { match_rows((match_job *)arg); return NULL; }

static int match_generic(const void *img, uint32_t w, uint32_t h, const void *tpl, uint32_t tw,
                         uint32_t th, int method, float *out, int nthreads, int is_f32) {
    /* ImageTemplate::new, template_matching.rs:394-404 */
    if (w < tw || h < th) return -1;
    uint32_t ow = w - tw + 1, oh = h - th + 1; /* template_matching.rs:419-425 */
    size_t n = (size_t)tw * th;
    /* SseNormalized/CcorrNormalized::init, template_matching.rs:235-238,276-279 */
    float tss = is_f32 ? square_sum_f32(tpl, n) : square_sum_u8(tpl, n);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    match_job jobs[256];
    pthread_t ths[256];
    for (int t = 0; t < nthreads; t++)
        jobs[t] = (match_job){img, w, tpl, tw, th, method, tss, out, ow, oh, is_f32, t, nthreads};
    if (nthreads == 1) {
        match_rows(&jobs[0]);
        return 0;
    }
    for (int t = 0; t < nthreads; t++) pthread_create(&ths[t], NULL, match_worker, &jobs[t]);
    for (int t = 0; t < nthreads; t++) pthread_join(ths[t], NULL);
    return 0;
}

int or_match_template_u8(const uint8_t *img, uint32_t w, uint32_t h, const uint8_t *tpl,
                         uint32_t tw, uint32_t th, int method, float *out) {
    return match_generic(img, w, h, tpl, tw, th, method, out, 1, 0);
}
int or_match_template_u8_parallel(const uint8_t *img, uint32_t w, uint32_t h,
                                  const uint8_t *tpl, uint32_t tw, uint32_t th, int method,
                                  float *out, int nthreads) {
    return match_generic(img, w, h, tpl, tw, th, method, out, nthreads, 0);
}
int or_match_template_f32(const float *img, uint32_t w, uint32_t h, const float *tpl,
                          uint32_t tw, uint32_t th, int method, float *out) {
    return match_generic(img, w, h, tpl, tw, th, method, out, 1, 1);
}
int or_match_template_f32_parallel(const float *img, uint32_t w, uint32_t h, const float *tpl,
                                   uint32_t tw, uint32_t th, int method, float *out,
                                   int nthreads) {
    return match_generic(img, w, h, tpl, tw, th, method, out, nthreads, 1);
}

/* Masked methods, template_matching.rs:295-386 (closures at :303-306,
 * :322-327, :340-343, :358-363; mult_square_sum at :381-386). */
int or_match_template_with_mask_u8(const uint8_t *img, uint32_t w, uint32_t h,
                                   const uint8_t *tpl, uint32_t tw, uint32_t th,
                                   const uint8_t *mask, uint32_t mw, uint32_t mh, int method,
                                   float *out) {
    if (mw != tw || mh != th) return -2; /* template_matching.rs:432-437 */
    if (w < tw || h < th) return -1;
    uint32_t ow = w - tw + 1, oh = h - th + 1;
    float tmss = -0.0f;
    for (size_t k = 0; k < (size_t)tw * th; k++) {
        float p = (float)tpl[k] * (float)mask[k];
        tmss = tmss + p * p;
    }
    for (uint32_t y = 0; y < oh; y++) {
        for (uint32_t x = 0; x < ow; x++) {
            float score = 0.0f, imim = 0.0f;
            for (uint32_t dy = 0; dy < th; dy++) {
                for (uint32_t dx = 0; dx < tw; dx++) {
                    float i = (float)img[(size_t)(y + dy) * w + x + dx];
                    float t = (float)tpl[(size_t)dy * tw + dx];
                    float m = (float)mask[(size_t)dy * tw + dx];
                    switch (method) {
                    case OR_SSE: { float d = (t - i) * m; score = score + d * d; } break;
                    case OR_SSE_NORMED: {
                        float d = (t - i) * m; score = score + d * d;
                        float e = i * m; imim = imim + e * e;
                    } break;
                    case OR_CCORR: score = score + t * i * m * m; break;
                    default: {
                        score = score + t * i * m * m;
                        float e = i * m; imim = imim + e * e;
                    } break;
                    }
                }
            }
            float r = score;
            if (method == OR_SSE_NORMED || method == OR_CCORR_NORMED) {
                float norm = sqrtf(tmss * imim); /* template_matching.rs:329, :365 */
                r = norm > 0.0f ? score / norm : score;
            }
            out[(size_t)y * ow + x] = r;
        }
    }
    return 0;
}

/* FP64 oracle, u8 images: every sum is an exact integer (int64). */
int or_match_template_u8_f64(const uint8_t *img, uint32_t w, uint32_t h, const uint8_t *tpl,
                             uint32_t tw, uint32_t th, int method, double *out) {
    if (w < tw || h < th) return -1;
    uint32_t ow = w - tw + 1, oh = h - th + 1;
    int64_t tss = 0;
    for (size_t k = 0; k < (size_t)tw * th; k++) tss += (int64_t)tpl[k] * tpl[k];
    for (uint32_t y = 0; y < oh; y++) {
        for (uint32_t x = 0; x < ow; x++) {
            int64_t c = 0, ii = 0, sse = 0;
            for (uint32_t dy = 0; dy < th; dy++) {
                for (uint32_t dx = 0; dx < tw; dx++) {
                    int64_t i = img[(size_t)(y + dy) * w + x + dx];
                    int64_t t = tpl[(size_t)dy * tw + dx];
                    c += i * t; ii += i * i; sse += (t - i) * (t - i);
                }
            }
            double score = (method == OR_SSE || method == OR_SSE_NORMED) ? (double)sse : (double)c;
            if (method == OR_SSE_NORMED || method == OR_CCORR_NORMED) {
                double norm = sqrt((double)ii * (double)tss);
                score = norm > 0.0 ? score / norm : score;
            }
            out[(size_t)y * ow + x] = score;
        }
    }
    return 0;
}

/* Neumaier-compensated double sum, for the FP64 oracle of f32 images
 * (products of two f32 values are exact in double). */
typedef struct { double s, c; } ksum;
static inline void kadd(ksum *k, double v) {
    double t = k->s + v;
    if (fabs(k->s) >= fabs(v)) k->c += (k->s - t) + v; else k->c += (v - t) + k->s;
    k->s = t;
}
static inline double kval(const ksum *k) { return k->s + k->c; }

int or_match_template_f32_f64(const float *img, uint32_t w, uint32_t h, const float *tpl,
                              uint32_t tw, uint32_t th, int method, double *out) {
    if (w < tw || h < th) return -1;
    uint32_t ow = w - tw + 1, oh = h - th + 1;
    ksum tss = {0, 0};
    for (size_t k = 0; k < (size_t)tw * th; k++) kadd(&tss, (double)tpl[k] * (double)tpl[k]);
    double tssv = kval(&tss);
    for (uint32_t y = 0; y < oh; y++) {
        for (uint32_t x = 0; x < ow; x++) {
            ksum c = {0, 0}, ii = {0, 0}, sse = {0, 0};
            for (uint32_t dy = 0; dy < th; dy++) {
                for (uint32_t dx = 0; dx < tw; dx++) {
                    double i = img[(size_t)(y + dy) * w + x + dx];
                    double t = tpl[(size_t)dy * tw + dx];
                    double d = t - i; /* exact: both are f32 values */
                    kadd(&c, i * t); kadd(&ii, i * i);
                    /* d*d may round in double; split d exactly via fma */
                    double p = d * d;
                    kadd(&sse, p); kadd(&sse, fma(d, d, -p));
                }
            }
            double score = (method == OR_SSE || method == OR_SSE_NORMED) ? kval(&sse) : kval(&c);
            if (method == OR_SSE_NORMED || method == OR_CCORR_NORMED) {
                double norm = sqrt(kval(&ii) * tssv);
                score = norm > 0.0 ? score / norm : score;
            }
            out[(size_t)y * ow + x] = score;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* find_extremes, template_matching.rs:477-511                          */
/* ------------------------------------------------------------------ */
/* Initial value from (0, 0); strict '<' / '>' over enumerate_pixels()
 * (row-major), so the first occurrence wins ties and a NaN never replaces. */
int or_find_extremes_f32(const float *img, uint32_t w, uint32_t h, float vals[2],
                         uint32_t locs[4]) {
    if (w == 0 || h == 0) return -1; /* template_matching.rs:483-486 */
    float mn = img[0], mx = img[0];
    uint32_t mnx = 0, mny = 0, mxx = 0, mxy = 0;
    for (uint32_t y = 0; y < h; y++)
        for (uint32_t x = 0; x < w; x++) {
            float p = img[(size_t)y * w + x];
            if (p < mn) { mn = p; mnx = x; mny = y; }
            if (p > mx) { mx = p; mxx = x; mxy = y; }
        }
    vals[0] = mx; vals[1] = mn;
    locs[0] = mxx; locs[1] = mxy; locs[2] = mnx; locs[3] = mny;
    return 0;
}

int or_find_extremes_u8(const uint8_t *img, uint32_t w, uint32_t h, uint8_t vals[2],
                        uint32_t locs[4]) {
    if (w == 0 || h == 0) return -1;
    uint8_t mn = img[0], mx = img[0];
    uint32_t mnx = 0, mny = 0, mxx = 0, mxy = 0;
    for (uint32_t y = 0; y < h; y++)
        for (uint32_t x = 0; x < w; x++) {
            uint8_t p = img[(size_t)y * w + x];
            if (p < mn) { mn = p; mnx = x; mny = y; }
            if (p > mx) { mx = p; mxx = x; mxy = y; }
        }
    vals[0] = mx; vals[1] = mn;
    locs[0] = mxx; locs[1] = mxy; locs[2] = mnx; locs[3] = mny;
    return 0;
}

/* ------------------------------------------------------------------ */
/* integral_image_impl, integral_image.rs:97-152                        */
/* ------------------------------------------------------------------ */
void or_integral_image_u32(const uint8_t *img, uint32_t w, uint32_t h, uint32_t *out,
                           int squared) {
    size_t ow = (size_t)w + 1;
    memset(out, 0, sizeof(uint32_t) * ow * ((size_t)h + 1));
    for (uint32_t y = 0; y < h; y++) {
        uint32_t run = 0;
        for (uint32_t x = 0; x < w; x++) {
            uint32_t v = img[(size_t)y * w + x];
            run += squared ? v * v : v;
            out[((size_t)y + 1) * ow + x + 1] = out[(size_t)y * ow + x + 1] + run;
        }
    }
}

/* contrast.rs:19-47: mean over the clipped block by sum_image_pixels
 * (integral_image.rs:250-269, u32 wrapping adds/subs) divided by the pixel
 * count, then pixel (i32) >= mean (as i32) - delta, wrapping. */
int or_adaptive_threshold(const uint8_t *img, uint32_t w, uint32_t h, uint32_t block_radius,
                          int32_t delta, uint8_t *out) {
    if (block_radius == 0) return -1;
    if (w == 0 || h == 0) return 0;
    uint32_t *ii = malloc(sizeof(uint32_t) * ((size_t)w + 1) * ((size_t)h + 1));
    if (!ii) return -2;
    or_integral_image_u32(img, w, h, ii, 0);
    const size_t ow = (size_t)w + 1;
    for (uint32_t y = 0; y < h; y++) {
        const int64_t ylo64 = (int64_t)y - (int64_t)block_radius;
        const uint32_t ylo = ylo64 < 0 ? 0 : (uint32_t)ylo64;
        const uint64_t yhi64 = (uint64_t)y + block_radius;
        const uint32_t yhi = yhi64 > h - 1 ? h - 1 : (uint32_t)yhi64;
        for (uint32_t x = 0; x < w; x++) {
            const int64_t xlo64 = (int64_t)x - (int64_t)block_radius;
            const uint32_t xlo = xlo64 < 0 ? 0 : (uint32_t)xlo64;
            const uint64_t xhi64 = (uint64_t)x + block_radius;
            const uint32_t xhi = xhi64 > w - 1 ? w - 1 : (uint32_t)xhi64;
            const uint32_t n = (yhi - ylo + 1) * (xhi - xlo + 1);
            const uint32_t sum = ii[((size_t)yhi + 1) * ow + xhi + 1] + ii[(size_t)ylo * ow + xlo] -
                                 ii[(size_t)ylo * ow + xhi + 1] - ii[((size_t)yhi + 1) * ow + xlo];
            const uint32_t mean = sum / n;
            const int32_t thr = (int32_t)(mean - (uint32_t)delta);
            out[(size_t)y * w + x] = (int32_t)img[(size_t)y * w + x] >= thr ? 255 : 0;
        }
    }
    free(ii);
    return 0;
}

/* ---- canny, edges.rs:30-155 ------------------------------------------------------------ */

/* The direction class of non_maximum_suppression, edges.rs:62-80: atan2f in
 * degrees (f32 steps), negative angles moved up by 180, then the 22.5-degree
 * bins. Returns 0, 45, 90 or 135. */
int or_canny_direction(int gx, int gy) {
    const float rad_to_deg = 180.0f / 3.14159265358979323846f;
    float angle = atan2f((float)gy, (float)gx) * rad_to_deg;
    if (angle < 0.0f) angle += 180.0f;
    if (!(angle >= 22.5f && angle < 157.5f)) return 0;
    if (angle >= 22.5f && angle < 67.5f) return 45;
    if (angle >= 67.5f && angle < 112.5f) return 90;
    return 135;
}

/* or_canny_direction and hypotf over every (gx, gy) in [-r, r]^2, row-major
 * by gy then gx: the whole input domain of the Sobel responses of u8 images
 * at r = 1020. */
void or_canny_direction_grid(int r, uint8_t *cls, float *mag) {
    const int n = 2 * r + 1;
    for (int gy = -r; gy <= r; gy++)
        for (int gx = -r; gx <= r; gx++) {
            const size_t i = (size_t)(gy + r) * n + (gx + r);
            cls[i] = (uint8_t)or_canny_direction(gx, gy);
            mag[i] = hypotf((float)gx, (float)gy);
        }
}

/* canny(image, low, high), edges.rs:30-52: gaussian_blur_f32(1.4), Sobel by
 * filter_clamped::<_, i32, i16>, magnitude (gx as f32).hypot(gy as f32),
 * non_maximum_suppression (edges.rs:55-108) and hysteresis (edges.rs:110-155,
 * the same depth-first search over the same six neighbours). Returns -1 when
 * the reference panics on its assertion, -4 when its search would index
 * outside the image (a border pixel reached), -2 on allocation failure. */
int or_canny_thinned(const uint8_t *img, uint32_t w, uint32_t h, float *thin_out);
int or_canny_hysteresis(const float *thin, uint32_t w, uint32_t h, float low, float high,
                        uint32_t *stack, uint8_t *out);

int or_canny(const uint8_t *img, uint32_t w, uint32_t h, float low, float high, uint8_t *out) {
    if (!(high >= low)) return -1;
    if (w == 0 || h == 0) return -3;
    const size_t n = (size_t)w * h;
    float *thin = calloc(n, sizeof(float));
    uint32_t *stack = malloc(n * sizeof(uint32_t) * 2);
    int rc = thin && stack ? or_canny_thinned(img, w, h, thin) : -2;
    if (rc == 0) rc = or_canny_hysteresis(thin, w, h, low, high, stack, out);
    free(thin);
    free(stack);
    return rc;
}

/* Stages 1-3 of canny (edges.rs:34-50): the thinned magnitudes, zero on the
 * border. */
int or_canny_thinned(const uint8_t *img, uint32_t w, uint32_t h, float *thin) {
    if (w == 0 || h == 0) return -3;
    const size_t n = (size_t)w * h;
    static const int32_t sob_h[9] = {-1, 0, 1, -2, 0, 2, -1, 0, 1};
    static const int32_t sob_v[9] = {-1, -2, -1, 0, 0, 0, 1, 2, 1};
    uint8_t *blur = malloc(n);
    int16_t *gx = malloc(n * sizeof(int16_t)), *gy = malloc(n * sizeof(int16_t));
    float *g = malloc(n * sizeof(float));
    int rc = 0;
    if (!blur || !gx || !gy || !g) {
        rc = -2;
        goto done;
    }
    memset(thin, 0, n * sizeof(float));
    or_gaussian_blur_f32(img, blur, w, h, 0, 1.4f);
    or_filter_clamped(blur, 0, w, h, sob_h, 3, 3, 3, gx, 2);
    or_filter_clamped(blur, 0, w, h, sob_v, 3, 3, 3, gy, 2);
    for (size_t i = 0; i < n; i++) g[i] = hypotf((float)gx[i], (float)gy[i]);
    for (uint32_t y = 1; y + 1 < h; y++)
        for (uint32_t x = 1; x + 1 < w; x++) {
            const size_t i = (size_t)y * w + x;
            float c1, c2;
            switch (or_canny_direction(gx[i], gy[i])) {
            case 0: c1 = g[i - 1]; c2 = g[i + 1]; break;
            case 45: c1 = g[i + w + 1]; c2 = g[i - w - 1]; break;
            case 90: c1 = g[i - w]; c2 = g[i + w]; break;
            default: c1 = g[i + w - 1]; c2 = g[i - w + 1]; break;
            }
            thin[i] = (g[i] < c1 || g[i] < c2) ? 0.0f : g[i];
        }
done:
    free(blur);
    free(gx);
    free(gy);
    free(g);
    return rc;
}

/* hysteresis, edges.rs:110-155: the reference's depth-first search. stack
 * holds 2 * w * h entries. */
int or_canny_hysteresis(const float *thin, uint32_t w, uint32_t h, float low, float high,
                        uint32_t *stack, uint8_t *out) {
    int rc = 0;
    memset(out, 0, (size_t)w * h);
    static const int dx[6] = {1, 1, 0, -1, -1, -1}, dy[6] = {0, 1, 1, -1, 0, 1};
    for (uint32_t y = 1; y + 1 < h && rc == 0; y++)
        for (uint32_t x = 1; x + 1 < w && rc == 0; x++) {
            const size_t i = (size_t)y * w + x;
            if (!(thin[i] >= high) || out[i] != 0) continue;
            out[i] = 255;
            size_t top = 0;
            stack[top++] = x;
            stack[top++] = y;
            while (top && rc == 0) {
                const uint32_t ny = stack[--top], nx = stack[--top];
                if (nx == 0 || ny == 0 || nx + 1 >= w || ny + 1 >= h) {
                    rc = -4; /* a neighbour of this pixel lies outside the image */
                    break;
                }
                for (int k = 0; k < 6; k++) {
                    const size_t j = (size_t)(ny + dy[k]) * w + (nx + dx[k]);
                    if (thin[j] >= low && out[j] == 0) {
                        out[j] = 255;
                        stack[top++] = nx + dx[k];
                        stack[top++] = ny + dy[k];
                    }
                }
            }
        }
    return rc;
}
