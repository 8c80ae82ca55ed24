// gfx950 kernels for imageproc's template matching:
//   match_template (template_matching.rs:66-110, 165-293, 378-425)
//   find_extremes  (template_matching.rs:463-511)
//
// Two correlation engines:
//   * direct: one thread per output window, walking dy-outer/dx-inner in f32
//     with separately rounded multiplies and adds, i.e. the reference's own
//     operation sequence (slide_window_at, template_matching.rs:405-417);
//     bit-identical to it.
//   * MFMA (match_mfma.hip): exact integer correlation on int8 MFMA for u8
//     images, BF16x3 on MFMA for f32 images.
#include "iph_internal.hpp"

#include "match_common.hpp"

#include <algorithm>
#include <mutex>

namespace iph {
namespace {

constexpr int kSumChunk = 4096;  // products staged per round of the sequential sums

__device__ __forceinline__ float fmul(float a, float b) { return __fmul_rn(a, b); }
__device__ __forceinline__ float fadd(float a, float b) { return __fadd_rn(a, b); }
__device__ __forceinline__ float fsub(float a, float b) { return __fsub_rn(a, b); }

// square_sum, template_matching.rs:378-380: sequential f32 sum of t*t.
template <typename P>
__global__ __launch_bounds__(256) void square_sum_kernel(const P *__restrict__ tpl, size_t n,
                                                        float *__restrict__ out) {
    // the products in parallel (each rounded once, as in the reference), then
    // their sequential sum from LDS, kSumChunk at a time
    __shared__ float sq[kSumChunk];
    float s = -0.0f;
    for (size_t base = 0; base < n; base += kSumChunk) {
        const int m = (int)min((size_t)kSumChunk, n - base);
        for (int i = (int)threadIdx.x; i < m; i += (int)blockDim.x) {
            const float t = (float)tpl[base + i];
            sq[i] = fmul(t, t);
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (int i = 0; i < m; i++) s = fadd(s, sq[i]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s;
}

// One output per thread, reference operation order. The template element
// index is uniform across the wave, so template reads are scalar loads.
template <typename P, int METHOD>
__global__ __launch_bounds__(256) void match_direct_kernel(const P *__restrict__ img, int W,
                                                           const P *__restrict__ tpl, int tw,
                                                           int th, const float *__restrict__ tss,
                                                           float *__restrict__ out, int ow,
                                                           int oh) {
    int x = blockIdx.x * 64 + threadIdx.x;
    int y = blockIdx.y * 4 + threadIdx.y;
    if (x >= ow || y >= oh) return;
    float score = 0.0f, ii = 0.0f;
    for (int dy = 0; dy < th; dy++) {
        const P *irow = img + (size_t)(y + dy) * W + x;
        const P *trow = tpl + (size_t)dy * tw;
        for (int dx = 0; dx < tw; dx++) {
            float i = (float)irow[dx];
            float t = (float)trow[dx];
            if (METHOD == IPH_SSE || METHOD == IPH_SSE_NORMED) {
                float d = fsub(t, i);                      // (t - i).powi(2)
                score = fadd(score, fmul(d, d));
            } else {
                score = fadd(score, fmul(i, t));           // i * t
            }
            if (METHOD == IPH_SSE_NORMED || METHOD == IPH_CCORR_NORMED)
                ii = fadd(ii, fmul(i, i));
        }
    }
    if (METHOD == IPH_SSE_NORMED || METHOD == IPH_CCORR_NORMED) {
        // template_matching.rs:249-250 / 289-291. sqrtf and '/' are correctly
        // rounded (built with -fhip-fp32-correctly-rounded-divide-sqrt).
        float norm = sqrtf(fmul(ii, *tss));
        score = norm > 0.0f ? score / norm : score;
    }
    out[(size_t)y * ow + x] = score;
}

template <typename P>
iph_status direct_typed(const P *img, uint32_t w, const P *tpl, uint32_t tw, uint32_t th,
                        int32_t method, float *out, uint32_t ow, uint32_t oh, float *tss,
                        hipStream_t s) {
    square_sum_kernel<P><<<1, 256, 0, s>>>(tpl, (size_t)tw * th, tss);
    IPH_TRY(launch_status("square_sum_kernel"));
    dim3 grid((ow + 63) / 64, (oh + 3) / 4), block(64, 4);
#define LAUNCH(M)                                                                            \
    match_direct_kernel<P, M><<<grid, block, 0, s>>>(img, (int)w, tpl, (int)tw, (int)th, tss, \
                                                     out, (int)ow, (int)oh)
    switch (method) {
    case IPH_SSE: LAUNCH(IPH_SSE); break;
    case IPH_SSE_NORMED: LAUNCH(IPH_SSE_NORMED); break;
    case IPH_CCORR: LAUNCH(IPH_CCORR); break;
    default: LAUNCH(IPH_CCORR_NORMED); break;
    }
#undef LAUNCH
    return launch_status("match_direct_kernel");
}

// ---- match_template_with_mask ----------------------------------------------------
//
// template_matching.rs:125-163 with the masked methods of :295-376, one output
// per thread in the reference's operation order (slide_window_at, :442-457:
// dy outer, dx inner; every product and sum separately rounded f32):
//   SumOfSquaredErrors:           score += ((t - i) * m)^2
//   CrossCorrelation:             score += ((t * i) * m) * m
//   the normalised methods also:  im_im += (i * m)^2, then
//                                 score / sqrt(mult_square_sum(T, M) * im_im)
//                                 (the raw score if that norm is 0).

// mult_square_sum, template_matching.rs:381-386: sequential f32 sum of
// (t * m)^2 (Rust's f32 Sum starts from -0.0).
__global__ __launch_bounds__(256) void mult_square_sum_kernel(const uint8_t *__restrict__ tpl,
                                                             const uint8_t *__restrict__ mask,
                                                             size_t n, float *__restrict__ out) {
    // (t * m)^2 in parallel (each product rounded as in the reference), then
    // the sequential sum from LDS
    __shared__ float sq[kSumChunk];
    float s = -0.0f;
    for (size_t base = 0; base < n; base += kSumChunk) {
        const int m = (int)min((size_t)kSumChunk, n - base);
        for (int i = (int)threadIdx.x; i < m; i += (int)blockDim.x) {
            const float q = fmul((float)tpl[base + i], (float)mask[base + i]);
            sq[i] = fmul(q, q);
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (int i = 0; i < m; i++) s = fadd(s, sq[i]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s;
}

template <int METHOD>
__global__ __launch_bounds__(256) void match_masked_direct_kernel(
    const uint8_t *__restrict__ img, int W, const uint8_t *__restrict__ tpl,
    const uint8_t *__restrict__ mask, int tw, int th, const float *__restrict__ tms,
    float *__restrict__ out, int ow, int oh) {
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y * 4 + threadIdx.y;
    if (x >= ow || y >= oh) return;
    float score = 0.0f, imim = 0.0f;
    for (int dy = 0; dy < th; dy++) {
        const uint8_t *irow = img + (size_t)(y + dy) * W + x;
        const uint8_t *trow = tpl + (size_t)dy * tw;
        const uint8_t *mrow = mask + (size_t)dy * tw;
        for (int dx = 0; dx < tw; dx++) {
            const float i = (float)irow[dx], t = (float)trow[dx], m = (float)mrow[dx];
            if (METHOD == IPH_SSE || METHOD == IPH_SSE_NORMED) {
                const float d = fmul(fsub(t, i), m);
                score = fadd(score, fmul(d, d));
            } else {
                score = fadd(score, fmul(fmul(fmul(t, i), m), m));
            }
            if (METHOD == IPH_SSE_NORMED || METHOD == IPH_CCORR_NORMED) {
                const float e = fmul(i, m);
                imim = fadd(imim, fmul(e, e));
            }
        }
    }
    if (METHOD == IPH_SSE_NORMED || METHOD == IPH_CCORR_NORMED) {
        // template_matching.rs:329-330 / 365-366 (correctly rounded sqrt and /)
        const float norm = sqrtf(fmul(*tms, imim));
        score = norm > 0.0f ? score / norm : score;
    }
    out[(size_t)y * ow + x] = score;
}

// ---- find_extremes ------------------------------------------------------------
//
// Sequential semantics (template_matching.rs:488-503): start from (0,0), a
// later pixel replaces the current extreme only if strictly smaller/larger,
// so the result is the first pixel in row-major order holding the extreme
// value, NaNs never replace, and a NaN at (0,0) is never replaced.
// Parallel form: reduce 64-bit keys (ordered value << 32 | index), skipping
// NaNs; -0.0 and +0.0 map to the same key because the sequential compare
// treats them as equal. Each workgroup writes its (min, max) key pair; the
// finalize step reduces the pairs, reads the value stored at the winning
// index and applies the NaN-at-(0,0) rule.

__device__ __forceinline__ uint32_t order_key(float v) {
    uint32_t u = __float_as_uint(v);
    if ((u & 0x7fffffffu) == 0) u = 0;  // -0.0 == +0.0
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ uint32_t order_key(uint8_t v) { return v; }
__device__ __forceinline__ bool is_nan(float v) { return v != v; }
__device__ __forceinline__ bool is_nan(uint8_t) { return false; }

__device__ __forceinline__ unsigned long long shfl_xor_u64(unsigned long long v, int m) {
    uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
    lo = __shfl_xor(lo, m);
    hi = __shfl_xor(hi, m);
    return ((unsigned long long)hi << 32) | lo;
}

// Folds pixel i = value v into the running min and max keys.
template <typename P>
__device__ __forceinline__ void fold_extreme(P v, uint32_t i, unsigned long long &kmin,
                                             unsigned long long &kmax) {
    if (is_nan(v)) return;
    const unsigned long long k = (unsigned long long)order_key(v) << 32;
    const unsigned long long a = k | i, b = k | (0xffffffffu - i);
    kmin = a < kmin ? a : kmin;
    kmax = b > kmax ? b : kmax;
}

// Min and max of two key pairs over the 256 threads of a workgroup.
__device__ __forceinline__ void block_extremes(unsigned long long &kmin, unsigned long long &kmax) {
    __shared__ unsigned long long part[2][16];
    for (int m = 32; m >= 1; m >>= 1) {
        unsigned long long o = shfl_xor_u64(kmin, m);
        kmin = o < kmin ? o : kmin;
        o = shfl_xor_u64(kmax, m);
        kmax = o > kmax ? o : kmax;
    }
    if ((threadIdx.x & 63) == 0) {
        part[0][threadIdx.x >> 6] = kmin;
        part[1][threadIdx.x >> 6] = kmax;
    }
    __syncthreads();
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
        kmin = part[0][w] < kmin ? part[0][w] : kmin;
        kmax = part[1][w] > kmax ? part[1][w] : kmax;
    }
}

// Each thread reads 16-byte vectors (kUnroll in flight per iteration) when
// `vec` (16-byte aligned map); the last n % (16 / sizeof(P)) pixels, or all
// of them without `vec`, one at a time. pairs[2 b], pairs[2 b + 1] = the
// min and max key of workgroup b (~0 and 0 when it saw only NaNs).
constexpr int kExtUnroll = 4;
constexpr int kExtMaxBlocks = 1024;
template <typename P>
__global__ __launch_bounds__(256) void extremes_reduce_kernel(const P *__restrict__ img,
                                                              size_t n, int vec,
                                                              unsigned long long *pairs) {
    constexpr int kV = 16 / sizeof(P);
    unsigned long long kmin = ~0ull, kmax = 0ull;
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t step = (size_t)gridDim.x * blockDim.x;
    const size_t nv = vec ? n / kV : 0;
    const uint4 *iv = reinterpret_cast<const uint4 *>(img);
    for (size_t v0 = tid; v0 < nv; v0 += step * kExtUnroll) {
        uint4 q[kExtUnroll];
#pragma unroll
        for (int u = 0; u < kExtUnroll; u++) {
            const size_t v = v0 + step * u;
            q[u] = iv[v < nv ? v : nv - 1];
        }
#pragma unroll
        for (int u = 0; u < kExtUnroll; u++) {
            const size_t v = v0 + step * u;
            if (v >= nv) break;
            const P *e = reinterpret_cast<const P *>(&q[u]);
#pragma unroll
            for (int j = 0; j < kV; j++) fold_extreme(e[j], (uint32_t)(v * kV + j), kmin, kmax);
        }
    }
    for (size_t i = nv * kV + tid; i < n; i += step) fold_extreme(img[i], (uint32_t)i, kmin, kmax);
    block_extremes(kmin, kmax);
    if (threadIdx.x == 0) {
        pairs[2 * blockIdx.x] = kmin;
        pairs[2 * blockIdx.x + 1] = kmax;
    }
}

template <typename P>
__global__ __launch_bounds__(256) void extremes_finalize_kernel(const P *__restrict__ img,
                                                                uint32_t w,
                                                                const unsigned long long *pairs,
                                                                int nblocks, iph_extremes *out) {
    unsigned long long kmin = ~0ull, kmax = 0ull;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
        kmin = pairs[2 * b] < kmin ? pairs[2 * b] : kmin;
        kmax = pairs[2 * b + 1] > kmax ? pairs[2 * b + 1] : kmax;
    }
    block_extremes(kmin, kmax);
    if (threadIdx.x != 0) return;
    const unsigned long long keys[2] = {kmin, kmax};
    iph_extremes r;
    r.max_value.raw = 0;
    r.min_value.raw = 0;
    P v0 = img[0];
    uint32_t imin = 0, imax = 0;
    if (!is_nan(v0)) {
        imin = (uint32_t)(keys[0] & 0xffffffffu);
        imax = 0xffffffffu - (uint32_t)(keys[1] & 0xffffffffu);
    }
    P vmin = img[imin], vmax = img[imax];
    if constexpr (sizeof(P) == 1) {
        r.min_value.u8 = vmin;
        r.max_value.u8 = vmax;
    } else {
        r.min_value.f32 = vmin;
        r.max_value.f32 = vmax;
    }
    r.min_x = imin % w;
    r.min_y = imin / w;
    r.max_x = imax % w;
    r.max_y = imax / w;
    *out = r;
}

// ---- match_template_with_mask on MFMA ----------------------------------------------
//
// The masked cross-correlation sum t i m^2 (template_matching.rs:337-349) is
// the correlation of the image I with the template T M^2, and the normalised
// method's image term sum (i m)^2 (:351-374) the correlation of I^2 with M^2:
// two BF16x3 correlations of f32 planes (I and I^2 are exact in a bf16 hi +
// lo pair, so is M^2; T M^2 < 2^24 keeps 16 of its 24 bits), then the
// reference's f32 finish with its sequential mult_square_sum. The SSE
// methods are not taken: their sum t^2 m^2 - 2 sum t i m^2 + sum i^2 m^2
// cancels to nearly nothing at a match, far below the error of the parts.

// I and (optionally) I^2 of a u8 image as f32 planes.
__global__ __launch_bounds__(256) void mask_image_planes_kernel(const uint8_t *__restrict__ img,
                                                                size_t n, float *__restrict__ i1,
                                                                float *__restrict__ i2) {
    for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (size_t)gridDim.x * 256) {
        const float v = (float)img[k];
        i1[k] = v;
        if (i2) i2[k] = v * v;  // <= 65025: exact
    }
}

// T M^2 and (w2 not null: the normalised method) M^2, exact integers below
// 2^24 in f32.
__global__ __launch_bounds__(256) void mask_template_planes_kernel(const uint8_t *__restrict__ tpl,
                                                                   const uint8_t *__restrict__ mask,
                                                                   int n, float *__restrict__ w1,
                                                                   float *__restrict__ w2) {
    for (int k = threadIdx.x; k < n; k += 256) {
        const float m2 = (float)mask[k] * (float)mask[k];
        w1[k] = (float)tpl[k] * m2;
        if (w2) w2[k] = m2;
    }
}

// CrossCorrelationNormalizedWithMask's finish (:362-373) in f32: score /
// sqrt(mult_square_sum * im_im), the raw score when that norm is 0.
__global__ __launch_bounds__(256) void mask_ccorr_normed_finish_kernel(
    float *__restrict__ out, const float *__restrict__ imim, const float *__restrict__ tms,
    size_t n) {
    const float t = *tms;
    for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (size_t)gridDim.x * 256) {
        const float norm = sqrtf(t * imim[k]);
        const float sc = out[k];
        out[k] = norm > 0.0f ? sc / norm : sc;
    }
}

}  // namespace

size_t match_workspace(int32_t p_type, uint32_t w, uint32_t h, uint32_t tw, uint32_t th,
                       int32_t method, int32_t path) {
    if (path == IPH_PATH_DIRECT) return 256;
    return mfma_match_workspace(p_type, w, h, tw, th, method);
}

iph_status launch_match(const void *img, int32_t p_type, uint32_t w, uint32_t h,
                        const void *tpl, uint32_t tw, uint32_t th, int32_t method,
                        int32_t path, float *out, void *ws, size_t ws_bytes, hipStream_t s) {
    if (method < IPH_SSE || method > IPH_CCORR_NORMED) {
        set_error("unknown MatchTemplateMethod %d", method);
        return IPH_EINVAL;
    }
    uint32_t ow = w - tw + 1, oh = h - th + 1;
    if (ow == 0 || oh == 0) return IPH_OK;
    if (path == IPH_PATH_DIRECT) {
        Workspace wsp(ws, ws_bytes);
        float *tss = wsp.take<float>(1);
        if (!tss) {
            set_error("workspace too small for match_template");
            return IPH_EINVAL;
        }
        if (p_type == IPH_U8)
            return direct_typed(static_cast<const uint8_t *>(img), w,
                                static_cast<const uint8_t *>(tpl), tw, th, method, out, ow, oh,
                                tss, s);
        return direct_typed(static_cast<const float *>(img), w, static_cast<const float *>(tpl),
                            tw, th, method, out, ow, oh, tss, s);
    }
    return launch_mfma_match(img, p_type, w, h, tpl, tw, th, method, out, ws, ws_bytes, s);
}

// Workspace of the masked path: the mult_square_sum scalar, and on the MFMA
// path the f32 image planes, the template planes, the im_im map and the
// BF16x3 launcher's own workspace.
template <typename WS>
static void masked_plan(WS &wp, uint32_t w, uint32_t h, uint32_t tw, uint32_t th, int32_t method,
                        int32_t path) {
    wp.template take<float>(1);
    if (path != IPH_PATH_MFMA) return;
    const bool nrm = method == IPH_CCORR_NORMED;
    const size_t n = (size_t)w * h, nt = (size_t)tw * th, no = (size_t)(w - tw + 1) * (h - th + 1);
    wp.template take<float>(n);
    if (nrm) wp.template take<float>(n);
    wp.template take<float>(nt);
    if (nrm) wp.template take<float>(nt);
    if (nrm) wp.template take<float>(no);
    wp.template take<unsigned char>(match_workspace(IPH_F32, w, h, tw, th, IPH_CCORR,
                                                    IPH_PATH_MFMA));
}

size_t masked_match_workspace(uint32_t w, uint32_t h, uint32_t tw, uint32_t th, int32_t method,
                              int32_t path) {
    if (w < tw || h < th) return 256;
    WorkspacePlan wp;
    masked_plan(wp, w, h, tw, th, method, path);
    return std::max<size_t>(wp.used, 256);
}

iph_status launch_masked_match(const uint8_t *img, uint32_t w, uint32_t h, const uint8_t *tpl,
                               uint32_t tw, uint32_t th, const uint8_t *mask, int32_t method,
                               int32_t path, float *out, void *ws, size_t ws_bytes,
                               hipStream_t s) {
    if (path == IPH_PATH_MFMA && method != IPH_CCORR && method != IPH_CCORR_NORMED) {
        set_error("the MFMA path of match_template_with_mask takes CrossCorrelation and "
                  "CrossCorrelationNormalized; use the direct path");
        return IPH_EUNSUPPORTED;
    }
    if (path == IPH_PATH_MFMA && (tw > kMfmaMaxSide || th > kMfmaMaxSide || !mfma_match_fits(w, h))) {
        set_error("the MFMA path of match_template_with_mask takes templates up to %ux%u; use "
                  "the direct path", kMfmaMaxSide, kMfmaMaxSide);
        return IPH_EUNSUPPORTED;
    }
    Workspace wsp(ws, ws_bytes);
    float *tms = wsp.take<float>(1);
    if (!tms) {
        set_error("workspace too small for match_template_with_mask");
        return IPH_EINVAL;
    }
    const uint32_t ow = w - tw + 1, oh = h - th + 1;
    if ((size_t)ow * oh == 0) return IPH_OK;
    mult_square_sum_kernel<<<1, 256, 0, s>>>(tpl, mask, (size_t)tw * th, tms);
    IPH_TRY(launch_status("mult_square_sum_kernel"));
    if (path == IPH_PATH_MFMA) {
        const bool nrm = method == IPH_CCORR_NORMED;
        const size_t n = (size_t)w * h, nt = (size_t)tw * th, no = (size_t)ow * oh;
        float *i1 = wsp.take<float>(n);
        float *i2 = nrm ? wsp.take<float>(n) : nullptr;
        float *w1 = wsp.take<float>(nt);
        float *w2 = nrm ? wsp.take<float>(nt) : nullptr;
        float *imim = nrm ? wsp.take<float>(no) : nullptr;
        const size_t bw = match_workspace(IPH_F32, w, h, tw, th, IPH_CCORR, IPH_PATH_MFMA);
        unsigned char *bws = wsp.take<unsigned char>(bw);
        if (!i1 || !w1 || !bws || (nrm && (!i2 || !w2 || !imim))) {
            set_error("workspace too small for match_template_with_mask");
            return IPH_EINVAL;
        }
        const unsigned g = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
        mask_image_planes_kernel<<<g, 256, 0, s>>>(img, n, i1, i2);
        IPH_TRY(launch_status("mask_image_planes_kernel"));
        mask_template_planes_kernel<<<1, 256, 0, s>>>(tpl, mask, (int)nt, w1, w2);
        IPH_TRY(launch_status("mask_template_planes_kernel"));
        IPH_TRY(launch_match(i1, IPH_F32, w, h, w1, tw, th, IPH_CCORR, IPH_PATH_MFMA, out, bws,
                             bw, s));
        if (!nrm) return IPH_OK;
        IPH_TRY(launch_match(i2, IPH_F32, w, h, w2, tw, th, IPH_CCORR, IPH_PATH_MFMA, imim, bws,
                             bw, s));
        const unsigned go = (unsigned)std::min<size_t>((no + 255) / 256, 8192);
        mask_ccorr_normed_finish_kernel<<<go, 256, 0, s>>>(out, imim, tms, no);
        return launch_status("mask_ccorr_normed_finish_kernel");
    }
    dim3 grid((ow + 63) / 64, (oh + 3) / 4), block(64, 4);
#define LAUNCH(M)                                                                            \
    match_masked_direct_kernel<M><<<grid, block, 0, s>>>(img, (int)w, tpl, mask, (int)tw,      \
                                                         (int)th, tms, out, (int)ow, (int)oh)
    switch (method) {
    case IPH_SSE: LAUNCH(IPH_SSE); break;
    case IPH_SSE_NORMED: LAUNCH(IPH_SSE_NORMED); break;
    case IPH_CCORR: LAUNCH(IPH_CCORR); break;
    default: LAUNCH(IPH_CCORR_NORMED); break;
    }
#undef LAUNCH
    return launch_status("match_masked_direct_kernel");
}

size_t extremes_workspace() { return (size_t)kExtMaxBlocks * 2 * sizeof(unsigned long long); }

// ---- find_extremes without allocation -------------------------------------------
//
// The per-workgroup key pairs of find_extremes live in a fixed device array
// (kExtMaxBlocks pairs, 16 KiB per device, held until the process ends), so
// the caller passes only the image and the result. The reduction grid never
// exceeds kExtMaxBlocks workgroups, so every image up to 2^32 - 1 pixels fits.
// Each workgroup writes its pair before the finalize kernel reads it: nothing
// needs clearing between calls. All calls on one device share the array, so
// calls that overlap on one device (other streams, other threads) race.
static __device__ __align__(256) unsigned long long scratch[2 * kExtMaxBlocks];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device with that device current (each device holds its own copy).
iph_status scratch_address(void **ptr) {
    int dev = 0;
    IPH_TRY_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("device %d is beyond the %d devices the scratch table tracks", dev,
                  kMaxDevices);
        return IPH_EDEVICE;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            set_error("hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                      hipGetErrorString(e));
            return IPH_EDEVICE;
        }
        g_scratch_ptr[dev] = p;
    }
    *ptr = g_scratch_ptr[dev];
    return IPH_OK;
}
}  // namespace

iph_status launch_extremes_noalloc(const void *img, int32_t p_type, uint32_t w, uint32_t h,
                                   iph_extremes *out, hipStream_t s) {
    if ((uint64_t)w * h >= (1ull << 32)) {
        set_error("image of %ux%u pixels exceeds the 2^32 - 1 pixels find_extremes indexes",
                  w, h);
        return IPH_EINVAL;
    }
    void *ptr = nullptr;
    IPH_TRY(scratch_address(&ptr));
    return launch_extremes(img, p_type, w, h, out, ptr, sizeof(scratch), s);
}

iph_status launch_extremes(const void *img, int32_t p_type, uint32_t w, uint32_t h,
                           iph_extremes *out, void *ws, size_t ws_bytes, hipStream_t s) {
    Workspace wsp(ws, ws_bytes);
    unsigned long long *pairs = wsp.take<unsigned long long>(2 * kExtMaxBlocks);
    if (!pairs) {
        set_error("workspace too small for find_extremes");
        return IPH_EINVAL;
    }
    const size_t n = (size_t)w * h;
    const int vec = (reinterpret_cast<uintptr_t>(img) & 15) == 0;
    const size_t per_block = (size_t)256 * kExtUnroll * (p_type == IPH_U8 ? 16 : 4);
    const int blocks = (int)std::max<size_t>(
        1, std::min<size_t>((n + per_block - 1) / per_block, kExtMaxBlocks));
    if (p_type == IPH_U8) {
        const uint8_t *im = static_cast<const uint8_t *>(img);
        extremes_reduce_kernel<uint8_t><<<blocks, 256, 0, s>>>(im, n, vec, pairs);
        IPH_TRY(launch_status("extremes_reduce_kernel"));
        extremes_finalize_kernel<uint8_t><<<1, 256, 0, s>>>(im, w, pairs, blocks, out);
    } else {
        const float *im = static_cast<const float *>(img);
        extremes_reduce_kernel<float><<<blocks, 256, 0, s>>>(im, n, vec, pairs);
        IPH_TRY(launch_status("extremes_reduce_kernel"));
        extremes_finalize_kernel<float><<<1, 256, 0, s>>>(im, w, pairs, blocks, out);
    }
    return launch_status("extremes_finalize_kernel");
}

}  // namespace iph
