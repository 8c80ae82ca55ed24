// gfx950 kernels for imageproc's filter module:
//   separable_filter / gaussian_blur_f32   (filter/mod.rs:147-195, 228-429)
//   filter_clamped                         (filter/mod.rs:24-69, 197-226)
//   box_filter                             (filter/box_filter.rs:19-70)
//
// Parity rules kept by every kernel here:
//   * the sum for one output is one sequential chain `acc = acc + p * w`
//     over taps in ascending order, each step a separately rounded f32 mul
//     and add (__fmul_rn/__fadd_rn; the file is also built with
//     -ffp-contract=off), or wrapping i32 arithmetic for integer kernels;
//   * out-of-image taps read the nearest edge pixel (clamp-to-edge);
//   * between the horizontal and the vertical pass the intermediate is
//     stored in the pixel type, i.e. truncated and clamped to u8 for
//     GrayImage, exactly as the reference's Image<P> intermediate.
#include "iph_internal.hpp"

#include <algorithm>
#include <type_traits>
#include <vector>

namespace iph {
namespace {

// ---- element conversions and Clamp<K> casts (definitions.rs:114-190) ----

__device__ __forceinline__ float to_acc(uint8_t v, float) { return (float)v; }
__device__ __forceinline__ float to_acc(float v, float) { return v; }
__device__ __forceinline__ int32_t to_acc(uint8_t v, int32_t) { return (int32_t)v; }

// accumulate, filter/mod.rs:407-416: acc + c * w, no FMA contraction.
__device__ __forceinline__ float mac(float acc, float c, float w) {
    return __fadd_rn(acc, __fmul_rn(c, w));
}
// Integer kernels: i32 arithmetic that wraps like a release build.
__device__ __forceinline__ int32_t mac(int32_t acc, int32_t c, int32_t w) {
    return (int32_t)((uint32_t)acc + (uint32_t)c * (uint32_t)w);
}

template <typename S> struct ClampTo;
template <> struct ClampTo<uint8_t> {
    // x < 255 ? (x > 0 ? x as u8 : 0) : 255; NaN -> 255.
    __device__ static uint8_t from(float x) {
        return x < 255.0f ? (x > 0.0f ? (uint8_t)(uint32_t)x : (uint8_t)0) : (uint8_t)255;
    }
    __device__ static uint8_t from(int32_t x) {
        return x < 255 ? (x > 0 ? (uint8_t)x : (uint8_t)0) : (uint8_t)255;
    }
};
template <> struct ClampTo<uint16_t> {
    __device__ static uint16_t from(float x) {
        return x < 65535.0f ? (x > 0.0f ? (uint16_t)(uint32_t)x : (uint16_t)0) : (uint16_t)65535;
    }
    __device__ static uint16_t from(int32_t x) {
        return x < 65535 ? (x > 0 ? (uint16_t)x : (uint16_t)0) : (uint16_t)65535;
    }
};
template <> struct ClampTo<int16_t> {
    __device__ static int16_t from(int32_t x) {
        return x < 32767 ? (x > -32768 ? (int16_t)x : (int16_t)-32768) : (int16_t)32767;
    }
};
template <> struct ClampTo<int32_t> {
    __device__ static int32_t from(int32_t x) { return x; }
};
template <> struct ClampTo<float> {
    __device__ static float from(float x) { return x; }
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

// Copies a clamp-to-edge window of an image into LDS: element i of the
// R x C window is image[clamp(oy + i / C)][clamp(ox + i % C)]. Loads are
// issued in batches of 8 before their LDS stores, so global latency is paid
// once per batch rather than once per element.
template <typename P>
__device__ __forceinline__ void stage_window(const P *__restrict__ img, int W, int H, int oy,
                                             int ox, int R, int C, P *dst) {
    const int total = R * C;
    for (int base = threadIdx.x; base < total; base += 8 * (int)blockDim.x) {
        P v[8];
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const int i = min(base + b * (int)blockDim.x, total - 1);
            const int r = i / C, c = i - r * C;
            v[b] = img[(size_t)clampi(oy + r, 0, H - 1) * W + clampi(ox + c, 0, W - 1)];
        }
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const int i = base + b * (int)blockDim.x;
            if (i < total) dst[i] = v[b];
        }
    }
}

// stage_window for a u8 image into an f32 window (the conversion is exact).
__device__ __forceinline__ void stage_window_f32(const uint8_t *__restrict__ img, int W, int H,
                                                 int oy, int ox, int R, int C, float *dst) {
    const int total = R * C;
    for (int base = threadIdx.x; base < total; base += 8 * (int)blockDim.x) {
        uint32_t v[8];
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const int i = min(base + b * (int)blockDim.x, total - 1);
            const int r = i / C, c = i - r * C;
            v[b] = img[(size_t)clampi(oy + r, 0, H - 1) * W + clampi(ox + c, 0, W - 1)];
        }
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const int i = base + b * (int)blockDim.x;
            if (i < total) dst[i] = (float)v[b];
        }
    }
}

// ---- fused separable filter ------------------------------------------------
//
// One workgroup produces a TW x TH output tile of one image:
//   1. load the (TH+L-1) x (TW+L-1) input window, clamped to the image, into LDS;
//   2. horizontal pass over every window row -> P-typed intermediate in LDS;
//   3. vertical pass over the intermediate -> output.
// Window row j holds image row clamp(y0 - L/2 + j), so the horizontal pass of
// a replicated edge row equals the replicated horizontal result, and the
// vertical taps clamp exactly like vertical_filter's.

constexpr int kFusedMaxTaps = 48;
constexpr int kSepTW = 64;
constexpr int kSepTH = 32;
constexpr int kSepRun = 8;  // outputs per task with compile-time taps
// output rows per sep_fused_kernel tile: 64 for 21 taps or more (the halo
// rows then weigh less), else 32
__host__ __device__ constexpr int sep_tile_h(int lt) { return lt >= 21 ? 64 : kSepTH; }

template <typename K> struct FusedTaps {
    K h[kFusedMaxTaps];
    K v[kFusedMaxTaps];
};

// LT > 0: the tap count is the compile-time LT (loops unrolled, taps in
// scalar registers); LT == 0: Lrt at run time.
template <typename P, typename K, int LT>
__global__ __launch_bounds__(256) void sep_fused_kernel(const P *__restrict__ src,
                                                        P *__restrict__ dst, int W, int H,
                                                        size_t stride, int Lrt,
                                                        FusedTaps<K> taps) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int TH = sep_tile_h(LT);
    const int L = LT > 0 ? LT : Lrt;
    const int R = TH + L - 1;
    const int C = kSepTW + L - 1;
    P *in_s = reinterpret_cast<P *>(smem_raw);
    P *mid_s = in_s + R * C;

    const int half = L / 2;
    const int x0 = blockIdx.x * kSepTW;
    const int y0 = blockIdx.y * TH;
    const P *img = src + (size_t)blockIdx.z * stride;
    P *out = dst + (size_t)blockIdx.z * stride;

    stage_window(img, W, H, y0 - half, x0 - half, R, C, in_s);
    __syncthreads();

    if constexpr (LT > 0) {
        // compile-time taps: each task computes kSepRun outputs from the
        // kSepRun + L - 1 values it loads once (the same products and sums,
        // in the same order, as one output per thread)
        constexpr int NV = kSepRun + LT - 1;
        constexpr int HT = kSepTW / kSepRun;  // tasks per row
        for (int i = threadIdx.x; i < R * HT; i += blockDim.x) {
            const int r = i / HT, c0 = (i - r * HT) * kSepRun;
            const P *row = in_s + r * C + c0;
            P v[NV];
#pragma unroll
            for (int j = 0; j < NV; j++) v[j] = row[j];
#pragma unroll
            for (int o = 0; o < kSepRun; o++) {
                K acc = K(0);
#pragma unroll
                for (int t = 0; t < LT; t++) acc = mac(acc, to_acc(v[o + t], K(0)), taps.h[t]);
                mid_s[r * kSepTW + c0 + o] = ClampTo<P>::from(acc);
            }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < (TH / kSepRun) * kSepTW; i += blockDim.x) {
            const int rb = i / kSepTW, c = i - rb * kSepTW;
            const int r0 = rb * kSepRun, gx = x0 + c;
            const P *col = mid_s + r0 * kSepTW + c;
            P v[NV];
#pragma unroll
            for (int j = 0; j < NV; j++) v[j] = col[j * kSepTW];
#pragma unroll
            for (int o = 0; o < kSepRun; o++) {
                K acc = K(0);
#pragma unroll
                for (int t = 0; t < LT; t++) acc = mac(acc, to_acc(v[o + t], K(0)), taps.v[t]);
                const int gy = y0 + r0 + o;
                if (gy < H && gx < W) out[(size_t)gy * W + gx] = ClampTo<P>::from(acc);
            }
        }
        return;
    }

    for (int i = threadIdx.x; i < R * kSepTW; i += blockDim.x) {
        int r = i / kSepTW, c = i - r * kSepTW;
        const P *row = in_s + r * C + c;
        K acc = K(0);
#pragma unroll
        for (int t = 0; t < (LT > 0 ? LT : kFusedMaxTaps); t++) {
            if (LT == 0 && t >= L) break;
            acc = mac(acc, to_acc(row[t], K(0)), taps.h[t]);
        }
        mid_s[i] = ClampTo<P>::from(acc);
    }
    __syncthreads();

    for (int i = threadIdx.x; i < TH * kSepTW; i += blockDim.x) {
        int r = i / kSepTW, c = i - r * kSepTW;
        int gy = y0 + r, gx = x0 + c;
        if (gy >= H || gx >= W) continue;
        const P *col = mid_s + r * kSepTW + c;
        K acc = K(0);
#pragma unroll
        for (int t = 0; t < (LT > 0 ? LT : kFusedMaxTaps); t++) {
            if (LT == 0 && t >= L) break;
            acc = mac(acc, to_acc(col[t * kSepTW], K(0)), taps.v[t]);
        }
        out[(size_t)gy * W + gx] = ClampTo<P>::from(acc);
    }
}

// ---- one 1-D pass, any tap count (weights in device memory) ----------------

template <typename P, typename K, bool VERTICAL>
__global__ __launch_bounds__(256) void filter1d_kernel(const P *__restrict__ src,
                                                       P *__restrict__ dst, int W, int H,
                                                       size_t stride, const K *__restrict__ k,
                                                       int L) {
    int x = blockIdx.x * 64 + threadIdx.x;
    int y = blockIdx.y * 4 + threadIdx.y;
    if (x >= W || y >= H) return;
    const P *img = src + (size_t)blockIdx.z * stride;
    const int half = L / 2;
    K acc = K(0);
    if (VERTICAL) {
        for (int t = 0; t < L; t++) {
            int yy = clampi(y + t - half, 0, H - 1);
            acc = mac(acc, to_acc(img[(size_t)yy * W + x], K(0)), k[t]);
        }
    } else {
        const P *row = img + (size_t)y * W;
        for (int t = 0; t < L; t++) {
            int xx = clampi(x + t - half, 0, W - 1);
            acc = mac(acc, to_acc(row[xx], K(0)), k[t]);
        }
    }
    dst[(size_t)blockIdx.z * stride + (size_t)y * W + x] = ClampTo<P>::from(acc);
}

template <typename P, typename K>
__global__ void fill_kernel(P *dst, size_t count, K value) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t step = (size_t)gridDim.x * blockDim.x;
    P v = ClampTo<P>::from(value);
    for (; i < count; i += step) dst[i] = v;
}

// ---- 2-D correlation (filter, filter/mod.rs:24-69) --------------------------

constexpr int kF2TW = 64;
constexpr int kF2TH = 16;

template <typename P, typename K, typename S>
__global__ __launch_bounds__(256) void filter2d_tiled_kernel(const P *__restrict__ src,
                                                             S *__restrict__ dst, int W, int H,
                                                             int kw, int kh, Taps<K> k) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    P *in_s = reinterpret_cast<P *>(smem_raw);
    const int R = kF2TH + kh - 1, C = kF2TW + kw - 1;
    const int x0 = blockIdx.x * kF2TW, y0 = blockIdx.y * kF2TH;
    // Window origin: kernel centre (kw/2, kh/2), filter/mod.rs:50,53.
    const int ox = x0 - kw / 2, oy = y0 - kh / 2;
    stage_window(src, W, H, oy, ox, R, C, in_s);
    __syncthreads();
    for (int i = threadIdx.x; i < kF2TH * kF2TW; i += blockDim.x) {
        int r = i / kF2TW, c = i - r * kF2TW;
        int gy = y0 + r, gx = x0 + c;
        if (gy >= H || gx >= W) continue;
        K acc = K(0);
        int t = 0;
        for (int ky = 0; ky < kh; ky++) {  // ky outer, kx inner: filter/mod.rs:49-62
            const P *row = in_s + (r + ky) * C + c;
            for (int kx = 0; kx < kw; kx++, t++) acc = mac(acc, to_acc(row[kx], K(0)), k.w[t]);
        }
        dst[(size_t)gy * W + gx] = ClampTo<S>::from(acc);
    }
}

// filter_clamped<u8, f32, u8> with a KW x KH kernel known at compile time
// (3 or 5 per side): one 64 x 32 output tile per workgroup, the clamp-to-edge
// window staged once as f32 (rows 68 floats apart), each thread eight
// consecutive outputs of one row from three 16-byte reads per kernel row,
// taps in scalar registers, each output's chain in the reference's order
// (ky outer, kx inner, separately rounded multiply then add); eight threads
// write a row's 64 outputs as one contiguous run.
constexpr int kF2STW = 64, kF2STH = 32, kF2SP = 68;

// One 64 x 32 output tile of filter_clamped<u8, f32, u8> with a KW x KH
// kernel (the reference's ky-outer, kx-inner f32 chain, each product and sum
// rounded): the window staged as f32 in `win`, eight outputs per thread.
// tap(q) returns weight q (row-major).
template <int KW, int KH, typename TapF>
__device__ __forceinline__ void f2s_tile(const uint8_t *__restrict__ src,
                                         uint8_t *__restrict__ dst, int W, int H, int x0,
                                         int y0, float *win, TapF tap) {
    static_assert(kF2STW + KW - 1 <= kF2SP, "the window row fits the pitch");
    const int ox = x0 - KW / 2, oy = y0 - KH / 2;  // filter/mod.rs:50,53
    constexpr int R = kF2STH + KH - 1;
    for (int i = threadIdx.x; i < R * kF2SP; i += 256) {
        const int r = i / kF2SP, c = i - r * kF2SP;
        win[i] = (float)src[(size_t)clampi(oy + r, 0, H - 1) * W + clampi(ox + c, 0, W - 1)];
    }
    __syncthreads();
    const int cg = threadIdx.x & 7, r = threadIdx.x >> 3;
    float acc[8];
#pragma unroll
    for (int o = 0; o < 8; o++) acc[o] = 0.0f;
#pragma unroll
    for (int ky = 0; ky < KH; ky++) {
        const float4 *row = reinterpret_cast<const float4 *>(win + (r + ky) * kF2SP + 8 * cg);
        float v[12];
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const float4 f = row[q];
            v[4 * q] = f.x; v[4 * q + 1] = f.y; v[4 * q + 2] = f.z; v[4 * q + 3] = f.w;
        }
#pragma unroll
        for (int o = 0; o < 8; o++)
#pragma unroll
            for (int kx = 0; kx < KW; kx++) acc[o] = mac(acc[o], v[o + kx], tap(ky * KW + kx));
    }
    const int gy = y0 + r, gx = x0 + 8 * cg;
    if (gy >= H || gx >= W) return;
    uint8_t b[8];
#pragma unroll
    for (int o = 0; o < 8; o++) b[o] = ClampTo<uint8_t>::from(acc[o]);
    uint8_t *o8 = dst + (size_t)gy * W + gx;
    if (gx + 8 <= W && ((reinterpret_cast<uintptr_t>(o8) & 7) == 0)) {
        *reinterpret_cast<uint2 *>(o8) =
            make_uint2(b[0] | (b[1] << 8) | (b[2] << 16) | ((uint32_t)b[3] << 24),
                       b[4] | (b[5] << 8) | (b[6] << 16) | ((uint32_t)b[7] << 24));
    } else {
        for (int o = 0; o < 8 && gx + o < W; o++) o8[o] = b[o];
    }
}

template <int KW, int KH>
__global__ __launch_bounds__(256) void filter2d_small_kernel(const uint8_t *__restrict__ src,
                                                             uint8_t *__restrict__ dst, int W,
                                                             int H, Taps<float> k) {
    __shared__ __align__(16) float win[(kF2STH + KH - 1) * kF2SP];
    f2s_tile<KW, KH>(src, dst, W, H, blockIdx.x * kF2STW, blockIdx.y * kF2STH, win,
                     [&](int q) { return k.w[q]; });
}

template <typename P, typename K, typename S>
__global__ __launch_bounds__(256) void filter2d_naive_kernel(const P *__restrict__ src,
                                                             S *__restrict__ dst, int W, int H,
                                                             int kw, int kh,
                                                             const K *__restrict__ k) {
    int x = blockIdx.x * 64 + threadIdx.x;
    int y = blockIdx.y * 4 + threadIdx.y;
    if (x >= W || y >= H) return;
    K acc = K(0);
    for (int ky = 0; ky < kh; ky++) {
        int yp = clampi(y + ky - kh / 2, 0, H - 1);
        for (int kx = 0; kx < kw; kx++) {
            int xp = clampi(x + kx - kw / 2, 0, W - 1);
            acc = mac(acc, to_acc(src[(size_t)yp * W + xp], K(0)), k[(size_t)ky * kw + kx]);
        }
    }
    dst[(size_t)y * W + x] = ClampTo<S>::from(acc);
}

// ---- ragged 2-D filter for batches ------------------------------------------
//
// filter_clamped<u8, f32, u8> over many images of different sizes in one
// launch: the grid walks a global list of TW x TH output tiles; job j owns
// tiles [tile0, tile0 + tiles_x * tiles_y) of its image. The arithmetic is
// filter2d_tiled_kernel's (ky outer, kx inner, f32 mul-then-add).
struct F2Job {
    const uint8_t *src;
    uint8_t *dst;
    int W, H, kw, kh, tiles_x, tile0;
    float w[kF2MaxJobTaps];
};

template <int TW, int TH>
__global__ __launch_bounds__(256) void filter2d_jobs_kernel(const F2Job *__restrict__ jobs,
                                                            int njobs) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    uint8_t *in_s = smem_raw;
    // the job owning this tile: binary search over tile0 (uniform per workgroup)
    const int t = (int)blockIdx.x;
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].tile0 <= t) lo = mid; else hi = mid - 1;
    }
    const F2Job &jb = jobs[lo];
    const int lt = t - jb.tile0, ty = lt / jb.tiles_x, tx = lt - ty * jb.tiles_x;
    const int W = jb.W, H = jb.H, kw = jb.kw, kh = jb.kh;
    const int R = TH + kh - 1, C = TW + kw - 1;
    const int x0 = tx * TW, y0 = ty * TH;
    stage_window(jb.src, W, H, y0 - kh / 2, x0 - kw / 2, R, C, in_s);
    __syncthreads();
    for (int i = threadIdx.x; i < TH * TW; i += blockDim.x) {
        int r = i / TW, c = i - r * TW;
        int gy = y0 + r, gx = x0 + c;
        if (gy >= H || gx >= W) continue;
        float acc = 0.0f;
        int q = 0;
        for (int ky = 0; ky < kh; ky++) {
            const uint8_t *row = in_s + (r + ky) * C + c;
            for (int kx = 0; kx < kw; kx++, q++) acc = mac(acc, to_acc(row[kx], 0.0f), jb.w[q]);
        }
        jb.dst[(size_t)gy * W + gx] = ClampTo<uint8_t>::from(acc);
    }
}

// The job table of 64 x 32 tiles when every job's kernel is at most KW x KH:
// jobs of exactly KW x KH take f2s_tile, the others the generic loop over a
// u8 window (the same arithmetic either way).
template <int KW, int KH>
__global__ __launch_bounds__(256) void filter2d_jobs_small_kernel(const F2Job *__restrict__ jobs,
                                                                  int njobs) {
    static_assert((kF2STH + KH - 1) * (kF2STW + KW - 1) <= (kF2STH + KH - 1) * kF2SP * 4,
                  "a u8 window fits the f32 window's LDS");
    __shared__ __align__(16) float win[(kF2STH + KH - 1) * kF2SP];
    const int t = (int)blockIdx.x;
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].tile0 <= t) lo = mid; else hi = mid - 1;
    }
    const F2Job &jb = jobs[lo];
    const int lt = t - jb.tile0, ty = lt / jb.tiles_x, tx = lt - ty * jb.tiles_x;
    const int x0 = tx * kF2STW, y0 = ty * kF2STH;
    if (jb.kw == KW && jb.kh == KH) {
        f2s_tile<KW, KH>(jb.src, jb.dst, jb.W, jb.H, x0, y0, win,
                         [&](int q) { return jb.w[q]; });
        return;
    }
    const int W = jb.W, H = jb.H, kw = jb.kw, kh = jb.kh;
    const int R = kF2STH + kh - 1, C = kF2STW + kw - 1;
    uint8_t *in_s = reinterpret_cast<uint8_t *>(win);
    stage_window(jb.src, W, H, y0 - kh / 2, x0 - kw / 2, R, C, in_s);
    __syncthreads();
    for (int i = threadIdx.x; i < kF2STH * kF2STW; i += blockDim.x) {
        int r = i / kF2STW, c = i - r * kF2STW;
        int gy = y0 + r, gx = x0 + c;
        if (gy >= H || gx >= W) continue;
        float acc = 0.0f;
        int q = 0;
        for (int ky = 0; ky < kh; ky++) {
            const uint8_t *row = in_s + (r + ky) * C + c;
            for (int kx = 0; kx < kw; kx++, q++) acc = mac(acc, to_acc(row[kx], 0.0f), jb.w[q]);
        }
        jb.dst[(size_t)gy * W + gx] = ClampTo<uint8_t>::from(acc);
    }
}

// ---- box filter (exact u32 window sums) ---------------------------------------
//
// out = floor(sum over the clamped (2r+1)-window / (2r+1)), first along rows
// into a u8 intermediate, then along columns of it: box_filter.rs:30-67.
// Each thread slides a window over kBoxSeg outputs. Its first window sum is
// (taps left of the image) * first + (taps right of it) * last + the in-image
// run, all in wrapping u32, which equals the reference's running-sum
// difference buf[u] - buf[l] (integral_image.rs:348-463) modulo 2^32.

constexpr int kBoxSeg = 64;

__device__ __forceinline__ uint32_t window_sum_first(const uint8_t *p, size_t pitch, int n,
                                                     int c, int r) {
    long long lo = (long long)c - r, hi = (long long)c + r;
    uint32_t s = 0;
    long long nl = min(hi, -1LL) - lo + 1;
    if (nl > 0) s += (uint32_t)nl * (uint32_t)p[0];
    long long nr = hi - max(lo, (long long)n) + 1;
    if (nr > 0) s += (uint32_t)nr * (uint32_t)p[(size_t)(n - 1) * pitch];
    long long a = max(lo, 0LL), b = min(hi, (long long)n - 1);
    for (long long j = a; j <= b; j++) s += p[(size_t)j * pitch];
    return s;
}

// Slides along one line of n pixels (pitch apart), writing kBoxSeg outputs.
__device__ __forceinline__ void box_line(const uint8_t *in, uint8_t *out, size_t pitch, int n,
                                         int c0, int r, uint32_t kdiv) {
    uint32_t s = window_sum_first(in, pitch, n, c0, r);
    int c1 = min(c0 + kBoxSeg, n);
    for (int c = c0; c < c1; c++) {
        out[(size_t)c * pitch] = (uint8_t)(s / kdiv);
        int add = min(c + 1 + r, n - 1);
        int sub = max(c - r, 0);
        s += (uint32_t)in[(size_t)add * pitch];
        s -= (uint32_t)in[(size_t)sub * pitch];
    }
}

__global__ __launch_bounds__(256) void box_rows_kernel(const uint8_t *__restrict__ src,
                                                       uint8_t *__restrict__ mid, int W, int H,
                                                       int rx, uint32_t kw) {
    int nseg = (W + kBoxSeg - 1) / kBoxSeg;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nseg * H) return;
    int y = (int)(t / nseg), seg = (int)(t - (long long)y * nseg);
    box_line(src + (size_t)y * W, mid + (size_t)y * W, 1, W, seg * kBoxSeg, rx, kw);
}

// Rows pass for rows of up to kBoxLdsMaxSeg * kBoxSeg pixels: a workgroup
// loads whole rows (256 / segments of them) into LDS with coalesced dword
// loads, each thread slides its window over one kBoxSeg-pixel segment there,
// and the rows are stored back coalesced. Segments are kBoxSegPitch bytes
// apart in LDS (17 dwords), so the byte reads of a wave's threads, one
// segment each, fall in distinct banks. Same sums as box_line.
constexpr int kBoxSegPitch = kBoxSeg + 4;
constexpr int kBoxLdsMaxSeg = 256;

template <typename G>
__device__ __forceinline__ uint32_t window_sum_first_f(G get, int n, int c, int r) {
    long long lo = (long long)c - r, hi = (long long)c + r;
    uint32_t s = 0;
    long long nl = min(hi, -1LL) - lo + 1;
    if (nl > 0) s += (uint32_t)nl * get(0);
    long long nr = hi - max(lo, (long long)n) + 1;
    if (nr > 0) s += (uint32_t)nr * get(n - 1);
    long long a = max(lo, 0LL), b = min(hi, (long long)n - 1);
    for (long long j = a; j <= b; j++) s += get((int)j);
    return s;
}

__global__ __launch_bounds__(256) void box_rows_lds_kernel(const uint8_t *__restrict__ src,
                                                           uint8_t *__restrict__ mid, int W,
                                                           int H, int rx, uint32_t kw, int nseg,
                                                           int rows_per_wg, bool aligned) {
    extern __shared__ __align__(16) uint8_t box_sm[];
    const int rowp = nseg * kBoxSegPitch;
    uint8_t *in_s = box_sm, *out_s = box_sm + rows_per_wg * rowp;
    const int y0 = blockIdx.x * rows_per_wg;
    const int nrows = min(rows_per_wg, H - y0);
    auto at = [&](int r, int c) { return r * rowp + (c >> 6) * kBoxSegPitch + (c & (kBoxSeg - 1)); };
    const int dpr = (W + 3) >> 2;  // dwords per row
    for (int i = threadIdx.x; i < nrows * dpr; i += 256) {
        const int r = i / dpr, c = 4 * (i - r * dpr);
        const uint8_t *row = src + (size_t)(y0 + r) * W;
        uint32_t v;
        if (aligned && c + 4 <= W) {
            v = *reinterpret_cast<const uint32_t *>(row + c);
        } else {
            v = 0;
            for (int b = 0; b < 4 && c + b < W; b++) v |= (uint32_t)row[c + b] << (8 * b);
        }
        *reinterpret_cast<uint32_t *>(in_s + at(r, c)) = v;
    }
    __syncthreads();
    const int r = (int)threadIdx.x / nseg, seg = (int)threadIdx.x - r * nseg;
    if (r < nrows) {
        const uint8_t *in_r = in_s + r * rowp;
        uint8_t *out_r = out_s + r * rowp;
        auto get = [&](int c) { return (uint32_t)in_r[(c >> 6) * kBoxSegPitch + (c & (kBoxSeg - 1))]; };
        const int c0 = seg * kBoxSeg;
        uint32_t sum = window_sum_first_f(get, W, c0, rx);
        const int c1 = min(c0 + kBoxSeg, W);
        for (int c = c0; c < c1; c++) {
            out_r[(c >> 6) * kBoxSegPitch + (c & (kBoxSeg - 1))] = (uint8_t)(sum / kw);
            sum += get(min(c + 1 + rx, W - 1));
            sum -= get(max(c - rx, 0));
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nrows * dpr; i += 256) {
        const int rr = i / dpr, c = 4 * (i - rr * dpr);
        uint8_t *row = mid + (size_t)(y0 + rr) * W;
        const uint32_t v = *reinterpret_cast<const uint32_t *>(out_s + at(rr, c));
        if (aligned && c + 4 <= W) {
            *reinterpret_cast<uint32_t *>(row + c) = v;
        } else {
            for (int b = 0; b < 4 && c + b < W; b++) row[c + b] = (uint8_t)(v >> (8 * b));
        }
    }
}

__global__ __launch_bounds__(256) void box_cols_kernel(const uint8_t *__restrict__ mid,
                                                       uint8_t *__restrict__ dst, int W, int H,
                                                       int ry, uint32_t kh) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    int seg = blockIdx.y;
    if (x >= W) return;
    box_line(mid + x, dst + x, (size_t)W, H, seg * kBoxSeg, ry, kh);
}

// One-kernel box filter for radii up to kBoxTileMaxR: a workgroup stages the
// input rows and columns its 128 x 64 output tile meets (rows and columns
// clamped, as the reference's running sums pad with the first and last
// pixel), computes the tile's u8 row means for every staged row into LDS,
// then the column means of those. Both passes slide exact integer window
// sums (any order gives the same integers) and divide by multiplying with
// ceil(2^32 / k): for sums below 2^16 and k <= 255 the high word is the
// floor of sum / k. Same bytes as box_rows_* + box_cols_kernel.
constexpr int kBoxTX = 128, kBoxTY = 64, kBoxRun = 8;
constexpr int kBoxTileMaxR = 32;

// staged bytes per row: the tile's columns and rx either side, a multiple of 4
__host__ __device__ constexpr int box_tile_pin(int rx) { return (kBoxTX + 2 * rx + 3) & ~3; }
__host__ __device__ constexpr size_t box_tile_lds(int rx, int ry) {
    return (size_t)(kBoxTY + 2 * ry) * (size_t)(box_tile_pin(rx) + kBoxTX);
}

__device__ __forceinline__ uint32_t box_div(uint32_t s, uint32_t magic) {
    return magic ? __umulhi(s, magic) : s;  // magic 0: k = 1
}

template <bool VEC>
__global__ __launch_bounds__(256) void box_tile_kernel(const uint8_t *__restrict__ src,
                                                       uint8_t *__restrict__ dst, int W, int H,
                                                       int rx, int ry, uint32_t mw, uint32_t mh) {
    extern __shared__ __align__(16) uint8_t box_tsm[];
    const int x0 = (int)blockIdx.x * kBoxTX, y0 = (int)blockIdx.y * kBoxTY;
    const int NR = kBoxTY + 2 * ry, PIN = box_tile_pin(rx);
    uint8_t *in_s = box_tsm;               // NR x PIN: global columns cs + [0, PIN)
    uint8_t *mid = box_tsm + NR * PIN;     // NR x kBoxTX: row means of the tile's columns
    const int cs = max(0, x0 - rx);
    // staging: four bytes per task from two aligned dwords (rows need not be
    // 4-aligned); bytes past the image's last column are never read
    const uint32_t ib = (uint32_t)reinterpret_cast<uintptr_t>(src) & 3u;
    const auto rs = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint8_t *>(src - ib), 0,
                                                      (int)(((uint32_t)(W * H) + ib + 3u) & ~3u),
                                                      0x00020000);
    const int dpr = PIN >> 2;
    // task (k, d) = (staged row, dword), stepped by 256 tasks without a division
    const int sk = 256 / dpr, sd = 256 - sk * dpr;
    int k = (int)threadIdx.x / dpr, d = (int)threadIdx.x - k * dpr;
    for (; k < NR; k += sk, d += sd) {
        if (d >= dpr) {
            d -= dpr;
            k++;
            if (k >= NR) break;
        }
        const int gy = min(max(y0 - ry + k, 0), H - 1);
        const uint32_t a = (uint32_t)(gy * W + cs + 4 * d) + ib;
        const uint32_t lo = __builtin_amdgcn_raw_buffer_load_b32(rs, (int)(a & ~3u), 0, 0);
        const uint32_t hi = __builtin_amdgcn_raw_buffer_load_b32(rs, (int)((a & ~3u) + 4u), 0, 0);
        *reinterpret_cast<uint32_t *>(in_s + k * PIN + 4 * d) =
            __builtin_amdgcn_alignbyte(hi, lo, a & 3u);
    }
    __syncthreads();
    // row means: task = (staged row, run of kBoxRun output columns)
    // (tiles whose windows stay inside the image read without clamping)
    const bool inner = x0 - rx >= 0 && x0 + kBoxTX + rx <= W;
    auto col = [&](int g) { return min(max(g, 0), W - 1) - cs; };
    constexpr int RPR = kBoxTX / kBoxRun;
    for (int i = (int)threadIdx.x; i < NR * RPR; i += 256) {
        const int k = i / RPR, c0 = (i - k * RPR) * kBoxRun;
        const uint8_t *row = in_s + k * PIN;
        const int g0 = x0 + c0;
        uint32_t sum = 0;
        uint32_t b[kBoxRun];
        if (inner) {
            const uint8_t *p = row + (g0 - cs);  // column g0
#pragma unroll 4
            for (int j = -rx; j <= rx; j++) sum += p[j];
#pragma unroll
            for (int c = 0; c < kBoxRun; c++) {
                b[c] = box_div(sum, mw);
                if (c + 1 < kBoxRun) sum += (uint32_t)p[c + rx + 1] - (uint32_t)p[c - rx];
            }
        } else {
#pragma unroll 4
            for (int j = -rx; j <= rx; j++) sum += row[col(g0 + j)];
#pragma unroll
            for (int c = 0; c < kBoxRun; c++) {
                b[c] = box_div(sum, mw);
                if (c + 1 < kBoxRun) {
                    sum += (uint32_t)row[col(g0 + c + rx + 1)];
                    sum -= (uint32_t)row[col(g0 + c - rx)];
                }
            }
        }
        *reinterpret_cast<uint2 *>(mid + k * kBoxTX + c0) =
            make_uint2(b[0] | (b[1] << 8) | (b[2] << 16) | (b[3] << 24),
                       b[4] | (b[5] << 8) | (b[6] << 16) | (b[7] << 24));
    }
    __syncthreads();
    // column means: thread = (column, half of the tile's rows); staged row
    // r + j holds the row means of global row clamp(y0 + r + j - ry)
    const int c = (int)threadIdx.x & (kBoxTX - 1), r0 = ((int)threadIdx.x >> 7) * (kBoxTY / 2);
    const int gx = x0 + c;
    uint32_t sum = 0;
#pragma unroll 4
    for (int j = 0; j <= 2 * ry; j++) sum += mid[(r0 + j) * kBoxTX + c];
    const int rend = min(r0 + kBoxTY / 2, H - y0);
    if (!VEC) {
#pragma unroll 8
        for (int r = r0; r < rend; r++) {
            if (gx < W) dst[(size_t)(y0 + r) * W + gx] = (uint8_t)box_div(sum, mh);
            if (r + 1 < rend) {
                sum += (uint32_t)mid[(r + 2 * ry + 1) * kBoxTX + c];
                sum -= (uint32_t)mid[r * kBoxTX + c];
            }
        }
        return;
    }
    // rows 16-aligned: the tile's output bytes into LDS (over the staged
    // input, done with), then 16-byte row pieces out
    uint8_t *out_s = in_s;
#pragma unroll 8
    for (int r = r0; r < rend; r++) {
        out_s[r * kBoxTX + c] = (uint8_t)box_div(sum, mh);
        if (r + 1 < rend) {
            sum += (uint32_t)mid[(r + 2 * ry + 1) * kBoxTX + c];
            sum -= (uint32_t)mid[r * kBoxTX + c];
        }
    }
    __syncthreads();
    const int nrows = min(kBoxTY, H - y0);
    for (int i = (int)threadIdx.x; i < nrows * (kBoxTX / 16); i += 256) {
        const int r = i >> 3, cx = (i & 7) * 16;
        if (x0 + cx < W)
            *reinterpret_cast<uint4 *>(dst + (size_t)(y0 + r) * W + x0 + cx) =
                *reinterpret_cast<const uint4 *>(out_s + r * kBoxTX + cx);
    }
}

// ---- host-side dispatch helpers -----------------------------------------------

template <typename K> void copy_taps(K *dst, const void *src, uint32_t len) {
    const K *s = static_cast<const K *>(src);
    for (uint32_t i = 0; i < len; i++) dst[i] = s[i];
}

template <typename P, typename K>
iph_status separable_typed(const void *src, void *dst, uint32_t w, uint32_t h, uint32_t n,
                           size_t stride, const void *hk, const void *vk, uint32_t len,
                           void *ws, size_t ws_bytes, hipStream_t s) {
    const P *in = static_cast<const P *>(src);
    P *out = static_cast<P *>(dst);
    if (len == 0) {
        // Both passes sum nothing: every pixel is Clamp(0) (filter/mod.rs:418-429).
        for (uint32_t i = 0; i < n; i++) {
            fill_kernel<P, K><<<1024, 256, 0, s>>>(out + (size_t)i * stride, (size_t)w * h, K(0));
        }
        return launch_status("fill_kernel");
    }
    if constexpr (std::is_same<P, uint8_t>::value && std::is_same<K, float>::value) {
        const float *hf = static_cast<const float *>(hk), *vf = static_cast<const float *>(vk);
        if (blur_mfma_supported(hf, vf, len)) {
            for (uint32_t i0 = 0; i0 < n; i0 += 65535) {
                uint32_t cnt = std::min<uint32_t>(65535, n - i0);
                IPH_TRY(launch_blur_u8_mfma(in + (size_t)i0 * stride, out + (size_t)i0 * stride, w,
                                            h, cnt, stride, hf, vf, len, s));
            }
            return IPH_OK;
        }
    }
    if (len <= (uint32_t)kFusedMaxTaps) {
        FusedTaps<K> taps{};
        copy_taps(taps.h, hk, len);
        copy_taps(taps.v, vk, len);
        int L = (int)len;
#define IPH_SEP(LT)                                                                          \
    do {                                                                                     \
        constexpr int TH = sep_tile_h(LT);                                                   \
        const size_t lds = sizeof(P) * ((size_t)(TH + L - 1) * (kSepTW + L - 1) +           \
                                        (size_t)(TH + L - 1) * kSepTW);                     \
        const dim3 grid((w + kSepTW - 1) / kSepTW, (h + TH - 1) / TH, n);                    \
        sep_fused_kernel<P, K, LT><<<grid, 256, lds, s>>>(in, out, (int)w, (int)h, stride, L, \
                                                          taps);                             \
    } while (0)
        switch (L) {
        case 3: IPH_SEP(3); break;
        case 5: IPH_SEP(5); break;
        case 7: IPH_SEP(7); break;
        case 9: IPH_SEP(9); break;
        case 11: IPH_SEP(11); break;
        case 13: IPH_SEP(13); break;
        case 15: IPH_SEP(15); break;
        case 17: IPH_SEP(17); break;
        case 19: IPH_SEP(19); break;
        case 21: IPH_SEP(21); break;
        case 23: IPH_SEP(23); break;
        case 25: IPH_SEP(25); break;
        default: IPH_SEP(0); break;
        }
#undef IPH_SEP
        return launch_status("sep_fused_kernel");
    }
    // Long kernels: two passes through a P-typed intermediate in the workspace.
    Workspace wsp(ws, ws_bytes);
    K *dk = wsp.take<K>(2 * (size_t)len);
    P *mid = wsp.take<P>((size_t)w * h * n);
    if (!dk || !mid) {
        set_error("workspace too small for a %u-tap separable filter", len);
        return IPH_EINVAL;
    }
    IPH_TRY_HIP(hipMemcpyAsync(dk, hk, sizeof(K) * len, hipMemcpyHostToDevice, s));
    IPH_TRY_HIP(hipMemcpyAsync(dk + len, vk, sizeof(K) * len, hipMemcpyHostToDevice, s));
    dim3 grid((w + 63) / 64, (h + 3) / 4, n), block(64, 4);
    for (uint32_t i = 0; i < n; i++) {
        filter1d_kernel<P, K, false><<<dim3(grid.x, grid.y, 1), block, 0, s>>>(
            in + (size_t)i * stride, mid + (size_t)i * w * h, (int)w, (int)h, 0, dk, (int)len);
        filter1d_kernel<P, K, true><<<dim3(grid.x, grid.y, 1), block, 0, s>>>(
            mid + (size_t)i * w * h, out + (size_t)i * stride, (int)w, (int)h, 0, dk + len,
            (int)len);
    }
    return launch_status("filter1d_kernel");
}

template <typename P, typename K>
iph_status filter1d_typed(const void *src, void *dst, uint32_t w, uint32_t h, const void *k,
                          uint32_t len, bool vertical, void *ws, size_t ws_bytes,
                          hipStream_t s) {
    if (len == 0) {
        fill_kernel<P, K><<<1024, 256, 0, s>>>(static_cast<P *>(dst), (size_t)w * h, K(0));
        return launch_status("fill_kernel");
    }
    Workspace wsp(ws, ws_bytes);
    K *dk = wsp.take<K>(len);
    if (!dk) {
        set_error("workspace too small for a %u-tap filter", len);
        return IPH_EINVAL;
    }
    IPH_TRY_HIP(hipMemcpyAsync(dk, k, sizeof(K) * len, hipMemcpyHostToDevice, s));
    dim3 grid((w + 63) / 64, (h + 3) / 4, 1), block(64, 4);
    if (vertical)
        filter1d_kernel<P, K, true><<<grid, block, 0, s>>>(
            static_cast<const P *>(src), static_cast<P *>(dst), (int)w, (int)h, 0, dk, (int)len);
    else
        filter1d_kernel<P, K, false><<<grid, block, 0, s>>>(
            static_cast<const P *>(src), static_cast<P *>(dst), (int)w, (int)h, 0, dk, (int)len);
    return launch_status("filter1d_kernel");
}

constexpr size_t kMaxTiledLds = 64 * 1024;

template <typename P, typename K, typename S>
iph_status filter2d_typed(const void *src, uint32_t w, uint32_t h, const void *kernel,
                          uint32_t kw, uint32_t kh, void *dst, void *ws, size_t ws_bytes,
                          hipStream_t s) {
    size_t taps = (size_t)kw * kh;
    size_t lds = sizeof(P) * ((size_t)kF2TH + kh - 1) * ((size_t)kF2TW + kw - 1);
    if constexpr (std::is_same<P, uint8_t>::value && std::is_same<K, float>::value &&
                  std::is_same<S, uint8_t>::value) {
        if ((kw == 3 || kw == 5) && (kh == 3 || kh == 5)) {
            Taps<float> k{};
            copy_taps(k.w, kernel, (uint32_t)taps);
            dim3 grid((w + kF2STW - 1) / kF2STW, (h + kF2STH - 1) / kF2STH, 1);
            const uint8_t *sp = static_cast<const uint8_t *>(src);
            uint8_t *dp = static_cast<uint8_t *>(dst);
            if (kw == 3 && kh == 3) filter2d_small_kernel<3, 3><<<grid, 256, 0, s>>>(sp, dp, (int)w, (int)h, k);
            else if (kw == 3) filter2d_small_kernel<3, 5><<<grid, 256, 0, s>>>(sp, dp, (int)w, (int)h, k);
            else if (kh == 3) filter2d_small_kernel<5, 3><<<grid, 256, 0, s>>>(sp, dp, (int)w, (int)h, k);
            else filter2d_small_kernel<5, 5><<<grid, 256, 0, s>>>(sp, dp, (int)w, (int)h, k);
            return launch_status("filter2d_small_kernel");
        }
    }
    if (taps <= (size_t)kMaxArgTaps && lds <= kMaxTiledLds) {
        Taps<K> k{};
        copy_taps(k.w, kernel, (uint32_t)taps);
        dim3 grid((w + kF2TW - 1) / kF2TW, (h + kF2TH - 1) / kF2TH, 1);
        filter2d_tiled_kernel<P, K, S><<<grid, 256, lds, s>>>(
            static_cast<const P *>(src), static_cast<S *>(dst), (int)w, (int)h, (int)kw,
            (int)kh, k);
        return launch_status("filter2d_tiled_kernel");
    }
    Workspace wsp(ws, ws_bytes);
    K *dk = wsp.take<K>(taps);
    if (!dk) {
        set_error("workspace too small for a %ux%u kernel", kw, kh);
        return IPH_EINVAL;
    }
    IPH_TRY_HIP(hipMemcpyAsync(dk, kernel, sizeof(K) * taps, hipMemcpyHostToDevice, s));
    dim3 grid((w + 63) / 64, (h + 3) / 4, 1), block(64, 4);
    filter2d_naive_kernel<P, K, S><<<grid, block, 0, s>>>(static_cast<const P *>(src),
                                                          static_cast<S *>(dst), (int)w, (int)h,
                                                          (int)kw, (int)kh, dk);
    return launch_status("filter2d_naive_kernel");
}

}  // namespace

// ---- launchers ----------------------------------------------------------------

size_t separable_workspace(uint32_t w, uint32_t h, uint32_t n, int32_t p_type, uint32_t len,
                           int32_t k_type) {
    if (len <= (uint32_t)kFusedMaxTaps) return 0;
    WorkspacePlan p;
    p.take<int32_t>(2 * (size_t)len);
    p.used = align_up(p.used, 256) + align_up((size_t)w * h * n * elem_size(p_type), 256);
    return p.used;
}

iph_status launch_separable(const void *src, void *dst, uint32_t w, uint32_t h, uint32_t n,
                            size_t stride, int32_t p_type, const void *hk, const void *vk,
                            uint32_t len, int32_t k_type, void *ws, size_t ws_bytes,
                            hipStream_t s) {
    if (w == 0 || h == 0 || n == 0) return IPH_OK;
    if (p_type == IPH_U8 && k_type == IPH_F32)
        return separable_typed<uint8_t, float>(src, dst, w, h, n, stride, hk, vk, len, ws,
                                               ws_bytes, s);
    if (p_type == IPH_U8 && k_type == IPH_I32)
        return separable_typed<uint8_t, int32_t>(src, dst, w, h, n, stride, hk, vk, len, ws,
                                                 ws_bytes, s);
    if (p_type == IPH_F32 && k_type == IPH_F32)
        return separable_typed<float, float>(src, dst, w, h, n, stride, hk, vk, len, ws,
                                             ws_bytes, s);
    set_error("separable filter: pixel type %d with kernel type %d is not supported", p_type,
              k_type);
    return IPH_EUNSUPPORTED;
}

iph_status launch_filter1d(const void *src, void *dst, uint32_t w, uint32_t h, int32_t p_type,
                           const void *k, uint32_t len, int32_t k_type, bool vertical,
                           void *ws, size_t ws_bytes, hipStream_t s) {
    if (w == 0 || h == 0) return IPH_OK;
    if (p_type == IPH_U8 && k_type == IPH_F32)
        return filter1d_typed<uint8_t, float>(src, dst, w, h, k, len, vertical, ws, ws_bytes, s);
    if (p_type == IPH_U8 && k_type == IPH_I32)
        return filter1d_typed<uint8_t, int32_t>(src, dst, w, h, k, len, vertical, ws, ws_bytes,
                                                s);
    if (p_type == IPH_F32 && k_type == IPH_F32)
        return filter1d_typed<float, float>(src, dst, w, h, k, len, vertical, ws, ws_bytes, s);
    set_error("1-D filter: pixel type %d with kernel type %d is not supported", p_type, k_type);
    return IPH_EUNSUPPORTED;
}

size_t filter_clamped_workspace(int32_t k_type, uint32_t kw, uint32_t kh) {
    return align_up((size_t)kw * kh * 4, 256);
}

iph_status launch_filter_clamped(const void *src, int32_t p, uint32_t w, uint32_t h,
                                 const void *kernel, int32_t k, uint32_t kw, uint32_t kh,
                                 void *dst, int32_t st, void *ws, size_t ws_bytes,
                                 hipStream_t s) {
    if (w == 0 || h == 0) return IPH_OK;
#define F2(PT, KT, ST)                                                                       \
    return filter2d_typed<PT, KT, ST>(src, w, h, kernel, kw, kh, dst, ws, ws_bytes, s)
    if (k == IPH_F32 && p == IPH_U8) {
        if (st == IPH_U8) F2(uint8_t, float, uint8_t);
        if (st == IPH_U16) F2(uint8_t, float, uint16_t);
        if (st == IPH_F32) F2(uint8_t, float, float);
    } else if (k == IPH_F32 && p == IPH_F32) {
        if (st == IPH_U8) F2(float, float, uint8_t);
        if (st == IPH_U16) F2(float, float, uint16_t);
        if (st == IPH_F32) F2(float, float, float);
    } else if (k == IPH_I32 && p == IPH_U8) {
        if (st == IPH_U8) F2(uint8_t, int32_t, uint8_t);
        if (st == IPH_U16) F2(uint8_t, int32_t, uint16_t);
        if (st == IPH_I16) F2(uint8_t, int32_t, int16_t);
        if (st == IPH_I32) F2(uint8_t, int32_t, int32_t);
    }
#undef F2
    set_error("filter_clamped: (pixel %d, kernel %d, output %d) is not a combination the "
              "reference accepts",
              p, k, st);
    return IPH_EUNSUPPORTED;
}

iph_status launch_box_filter(const uint8_t *src, uint8_t *dst, uint32_t w, uint32_t h,
                             uint32_t rx, uint32_t ry, void *ws, size_t ws_bytes,
                             hipStream_t s) {
    if (w == 0 || h == 0) return IPH_OK;  // box_filter.rs:22-24
    Workspace wsp(ws, ws_bytes);
    uint8_t *mid = wsp.take<uint8_t>((size_t)w * h);
    if (!mid) {
        set_error("workspace too small for box_filter");
        return IPH_EINVAL;
    }
    uint32_t kw = 2u * rx + 1u, kh = 2u * ry + 1u;  // box_filter.rs:26-27 (u32)
    if (rx <= (uint32_t)kBoxTileMaxR && ry <= (uint32_t)kBoxTileMaxR &&
        (uint64_t)w * h + 512u < (1ull << 31)) {
        // ceil(2^32 / k), 0 for k = 1 (the quotient is the sum)
        auto magic = [](uint32_t k) {
            return k == 1u ? 0u : (uint32_t)(((1ull << 32) + k - 1u) / k);
        };
        dim3 grid((w + kBoxTX - 1) / kBoxTX, (h + kBoxTY - 1) / kBoxTY);
        const size_t lds = box_tile_lds((int)rx, (int)ry);
        if (w % 16 == 0 && (reinterpret_cast<uintptr_t>(dst) & 15) == 0)
            box_tile_kernel<true><<<grid, 256, lds, s>>>(src, dst, (int)w, (int)h, (int)rx,
                                                         (int)ry, magic(kw), magic(kh));
        else
            box_tile_kernel<false><<<grid, 256, lds, s>>>(src, dst, (int)w, (int)h, (int)rx,
                                                          (int)ry, magic(kw), magic(kh));
        return launch_status("box_tile_kernel");
    }
    const uint32_t nseg = (w + kBoxSeg - 1) / kBoxSeg;
    if (nseg <= (uint32_t)kBoxLdsMaxSeg) {
        const int rows_per_wg = (int)(256 / nseg);
        const size_t lds = 2 * (size_t)rows_per_wg * nseg * kBoxSegPitch;
        const bool aligned = (w % 4) == 0 && (reinterpret_cast<uintptr_t>(src) & 3) == 0 &&
                             (reinterpret_cast<uintptr_t>(mid) & 3) == 0;
        box_rows_lds_kernel<<<(unsigned)((h + rows_per_wg - 1) / rows_per_wg), 256, lds, s>>>(
            src, mid, (int)w, (int)h, (int)rx, kw, (int)nseg, rows_per_wg, aligned);
        IPH_TRY(launch_status("box_rows_lds_kernel"));
    } else {
        long long nrow = (long long)nseg * h;
        box_rows_kernel<<<(unsigned)((nrow + 255) / 256), 256, 0, s>>>(src, mid, (int)w, (int)h,
                                                                       (int)rx, kw);
        IPH_TRY(launch_status("box_rows_kernel"));
    }
    dim3 grid((w + 255) / 256, (h + kBoxSeg - 1) / kBoxSeg);
    box_cols_kernel<<<grid, 256, 0, s>>>(mid, dst, (int)w, (int)h, (int)ry, kh);
    return launch_status("box_cols_kernel");
}

// ---- ragged 2-D filter batches --------------------------------------------------

namespace {
constexpr int kF2Variants[3][2] = {{64, 16}, {64, 32}, {128, 16}};

template <int TW, int TH> int f2_tiles(const FilterSpec &f) {
    return (int)(((f.w + TW - 1) / TW) * ((f.h + TH - 1) / TH));
}
}  // namespace

int filter_jobs_variants() { return 3; }

const char *filter_jobs_variant_name(int v) {
    static const char *names[3] = {"64x16", "64x32", "128x16"};
    return names[std::max(0, std::min(2, v))];
}

size_t filter_jobs_table_bytes(int n) { return (size_t)std::max(n, 1) * sizeof(F2Job); }

iph_status filter_jobs_table(const FilterSpec *specs, int n, int variant, void *host_table,
                             int *total_tiles) {
    const int TW = kF2Variants[variant][0], TH = kF2Variants[variant][1];
    F2Job *jobs = static_cast<F2Job *>(host_table);
    long long tile = 0;
    for (int i = 0; i < n; i++) {
        const FilterSpec &f = specs[i];
        if (f.kw * f.kh > (uint32_t)kF2MaxJobTaps || f.kw == 0 || f.kh == 0) {
            set_error("batched filter_clamped takes kernels of 1 to %d taps", kF2MaxJobTaps);
            return IPH_EUNSUPPORTED;
        }
        F2Job j{};
        j.src = f.src;
        j.dst = f.dst;
        j.W = (int)f.w;
        j.H = (int)f.h;
        j.kw = (int)f.kw;
        j.kh = (int)f.kh;
        j.tiles_x = (int)((f.w + TW - 1) / TW);
        j.tile0 = (int)tile;
        for (uint32_t q = 0; q < f.kw * f.kh; q++) j.w[q] = f.k[q];
        jobs[i] = j;
        tile += (long long)j.tiles_x * ((f.h + TH - 1) / TH);
        if (tile > (1LL << 30)) {
            set_error("batch too large for one launch");
            return IPH_EINVAL;
        }
    }
    *total_tiles = (int)tile;
    return IPH_OK;
}

iph_status launch_filter_jobs(const void *table_dev, int n, int total_tiles, int variant,
                              uint32_t max_kw, uint32_t max_kh, hipStream_t s) {
    if (n == 0 || total_tiles == 0) return IPH_OK;
    const F2Job *jobs = static_cast<const F2Job *>(table_dev);
#define FV(TW, TH)                                                                           \
    do {                                                                                     \
        size_t lds = (size_t)(TH + max_kh - 1) * (TW + max_kw - 1);                          \
        filter2d_jobs_kernel<TW, TH><<<total_tiles, 256, lds, s>>>(jobs, n);                 \
    } while (0)
    // 64 x 32 tiles with kernels of at most 3x3 or 5x5: eight outputs per thread
    // from an f32 window for the jobs of exactly that size
    static_assert(kF2STW == 64 && kF2STH == 32, "variant 1's tiles are f2s_tile's");
    if (variant == 1 && max_kw <= 5 && max_kh <= 5) {
        if (max_kw <= 3 && max_kh <= 3)
            filter2d_jobs_small_kernel<3, 3><<<total_tiles, 256, 0, s>>>(jobs, n);
        else
            filter2d_jobs_small_kernel<5, 5><<<total_tiles, 256, 0, s>>>(jobs, n);
        return launch_status("filter2d_jobs_small_kernel");
    }
    switch (variant) {
    case 0: FV(64, 16); break;
    case 1: FV(64, 32); break;
    default: FV(128, 16); break;
    }
#undef FV
    return launch_status("filter2d_jobs_kernel");
}

// ---- sharpen_gaussian epilogues (filter/sharpen.rs:31-38) ------------------------

namespace {
// map_subpixels(image, |x| x as f32)
// Four elements per thread; VEC: src and out 4-byte aligned, so the bytes
// move as one dword (the f32 side is always workspace, 16-byte aligned).
template <bool VEC>
__global__ __launch_bounds__(256) void u8_to_f32_kernel(const uint8_t *__restrict__ src,
                                                        float *__restrict__ dst, size_t n) {
    const size_t i = ((size_t)blockIdx.x * 256 + threadIdx.x) * 4;
    if (VEC && i + 4 <= n) {
        const uint32_t w = *reinterpret_cast<const uint32_t *>(src + i);
        *reinterpret_cast<float4 *>(dst + i) =
            make_float4((float)(w & 0xffu), (float)((w >> 8) & 0xffu), (float)((w >> 16) & 0xffu),
                        (float)(w >> 24));
    } else {
        for (size_t k = i; k < i + 4 && k < n; k++) dst[k] = (float)src[k];
    }
}

// v = (1 + amount) * p - amount * q, each operation rounded separately,
// then Clamp<f32> for u8 (definitions.rs:120-136: truncation, NaN -> 255).
__device__ __forceinline__ uint32_t sharpen_one(float a1, float amount, uint32_t p, float q) {
    const float v = __fsub_rn(__fmul_rn(a1, (float)p), __fmul_rn(amount, q));
    return v < 255.0f ? (v > 0.0f ? (uint32_t)v : 0u) : 255u;
}

// sharpen_gaussian in one pass over a kSepTW x kSepTH output tile
// (sharpen.rs:27-38): the u8 window is staged once, as f32; the f32 blur of
// that conversion runs as sep_fused_kernel<float, float> does on the converted
// image (the same products and sums in the same order, so the same bits),
// its horizontal pass kept as f32 in LDS; the combine reads the centre pixel
// from the staged window.
template <int LT>
__global__ __launch_bounds__(256) void sharpen_fused_kernel(const uint8_t *__restrict__ src,
                                                            uint8_t *__restrict__ dst, int W,
                                                            int H, int Lrt,
                                                            FusedTaps<float> taps,
                                                            float amount) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int L = LT > 0 ? LT : Lrt;
    const int R = kSepTH + L - 1;
    const int C = kSepTW + L - 1;
    float *in_s = reinterpret_cast<float *>(smem_raw);  // the window as f32 (exact)
    float *mid_s = in_s + R * C;
    const int half = L / 2;
    const int x0 = blockIdx.x * kSepTW;
    const int y0 = blockIdx.y * kSepTH;

    stage_window_f32(src, W, H, y0 - half, x0 - half, R, C, in_s);
    __syncthreads();
    const float a1 = __fadd_rn(1.0f, amount);

    if constexpr (LT > 0) {
        // 4 outputs per item from 4 + LT - 1 window values held in registers
        constexpr int NV = 4 + LT - 1;
        for (int i = threadIdx.x; i < R * (kSepTW / 4); i += blockDim.x) {
            const int r = i / (kSepTW / 4), c0 = 4 * (i - r * (kSepTW / 4));
            const float *row = in_s + r * C + c0;
            float x[NV];
#pragma unroll
            for (int k = 0; k < NV; k++) x[k] = row[k];
            float acc[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                acc[j] = 0.0f;
#pragma unroll
                for (int t = 0; t < LT; t++) acc[j] = mac(acc[j], x[j + t], taps.h[t]);
            }
            *reinterpret_cast<float4 *>(mid_s + r * kSepTW + c0) =
                make_float4(acc[0], acc[1], acc[2], acc[3]);
        }
        __syncthreads();
        for (int i = threadIdx.x; i < (kSepTH / 4) * kSepTW; i += blockDim.x) {
            const int r0 = 4 * (i / kSepTW), c = i % kSepTW;
            const int gx = x0 + c;
            if (gx >= W) continue;
            float x[NV];
#pragma unroll
            for (int k = 0; k < NV; k++) x[k] = mid_s[(r0 + k) * kSepTW + c];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int gy = y0 + r0 + j;
                float q = 0.0f;
#pragma unroll
                for (int t = 0; t < LT; t++) q = mac(q, x[j + t], taps.v[t]);
                const uint32_t p = (uint32_t)in_s[(r0 + j + half) * C + c + half];
                if (gy < H) dst[(size_t)gy * W + gx] = (uint8_t)sharpen_one(a1, amount, p, q);
            }
        }
        return;
    }

    for (int i = threadIdx.x; i < R * kSepTW; i += blockDim.x) {
        int r = i / kSepTW, c = i - r * kSepTW;
        const float *row = in_s + r * C + c;
        float acc = 0.0f;
        for (int t = 0; t < L; t++) acc = mac(acc, row[t], taps.h[t]);
        mid_s[i] = acc;
    }
    __syncthreads();

    for (int i = threadIdx.x; i < kSepTH * kSepTW; i += blockDim.x) {
        int r = i / kSepTW, c = i - r * kSepTW;
        int gy = y0 + r, gx = x0 + c;
        if (gy >= H || gx >= W) continue;
        const float *col = mid_s + r * kSepTW + c;
        float q = 0.0f;
        for (int t = 0; t < L; t++) q = mac(q, col[t * kSepTW], taps.v[t]);
        const uint32_t p = (uint32_t)in_s[(r + half) * C + c + half];
        dst[(size_t)gy * W + gx] = (uint8_t)sharpen_one(a1, amount, p, q);
    }
}

template <bool VEC>
__global__ __launch_bounds__(256) void sharpen_combine_kernel(const uint8_t *__restrict__ p,
                                                              const float *__restrict__ q,
                                                              float amount, size_t n,
                                                              uint8_t *__restrict__ out) {
    const size_t i = ((size_t)blockIdx.x * 256 + threadIdx.x) * 4;
    const float a1 = __fadd_rn(1.0f, amount);
    if (VEC && i + 4 <= n) {
        const uint32_t w = *reinterpret_cast<const uint32_t *>(p + i);
        const float4 qq = *reinterpret_cast<const float4 *>(q + i);
        *reinterpret_cast<uint32_t *>(out + i) =
            sharpen_one(a1, amount, w & 0xffu, qq.x) |
            (sharpen_one(a1, amount, (w >> 8) & 0xffu, qq.y) << 8) |
            (sharpen_one(a1, amount, (w >> 16) & 0xffu, qq.z) << 16) |
            (sharpen_one(a1, amount, w >> 24, qq.w) << 24);
    } else {
        for (size_t k = i; k < i + 4 && k < n; k++)
            out[k] = (uint8_t)sharpen_one(a1, amount, p[k], q[k]);
    }
}
}  // namespace

size_t sharpen_gaussian_workspace(uint32_t w, uint32_t h, uint32_t taps) {
    if (taps <= (uint32_t)kFusedMaxTaps) return 0;  // one fused pass
    WorkspacePlan wp;
    wp.take<float>((size_t)w * h);
    wp.take<float>((size_t)w * h);
    wp.take<uint8_t>(separable_workspace(w, h, 1, IPH_F32, taps, IPH_F32));
    return wp.used;
}

iph_status launch_sharpen_gaussian(const uint8_t *src, uint8_t *dst, uint32_t w, uint32_t h,
                                   const float *taps, uint32_t n_taps, float amount, void *ws,
                                   size_t ws_bytes, hipStream_t s) {
    const size_t n = (size_t)w * h;
    if (n == 0) return IPH_OK;
    if (n_taps <= (uint32_t)kFusedMaxTaps) {
        FusedTaps<float> ft{};
        copy_taps(ft.h, taps, n_taps);
        copy_taps(ft.v, taps, n_taps);
        const int L = (int)n_taps;
        const size_t lds = sizeof(float) * ((size_t)(kSepTH + L - 1) * (kSepTW + L - 1) +
                                            (size_t)(kSepTH + L - 1) * kSepTW);
        const dim3 grid((w + kSepTW - 1) / kSepTW, (h + kSepTH - 1) / kSepTH);
#define IPH_SHARP(LT)                                                                       \
    sharpen_fused_kernel<LT><<<grid, 256, lds, s>>>(src, dst, (int)w, (int)h, L, ft, amount)
        switch (L) {
        case 3: IPH_SHARP(3); break;
        case 5: IPH_SHARP(5); break;
        case 7: IPH_SHARP(7); break;
        case 9: IPH_SHARP(9); break;
        case 11: IPH_SHARP(11); break;
        case 13: IPH_SHARP(13); break;
        case 17: IPH_SHARP(17); break;
        default: IPH_SHARP(0); break;
        }
#undef IPH_SHARP
        return launch_status("sharpen_fused_kernel");
    }
    Workspace wsp(ws, ws_bytes);
    float *img = wsp.take<float>(n);
    float *smooth = wsp.take<float>(n);
    const size_t sep_bytes = separable_workspace(w, h, 1, IPH_F32, n_taps, IPH_F32);
    uint8_t *sep_ws = wsp.take<uint8_t>(sep_bytes);
    if (!img || !smooth || !sep_ws) {
        set_error("workspace too small for sharpen_gaussian");
        return IPH_EINVAL;
    }
    const unsigned blocks = (unsigned)((n + 1023) / 1024);
    const bool vec = ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 3) == 0;
    if (vec)
        u8_to_f32_kernel<true><<<blocks, 256, 0, s>>>(src, img, n);
    else
        u8_to_f32_kernel<false><<<blocks, 256, 0, s>>>(src, img, n);
    IPH_TRY(launch_status("u8_to_f32_kernel"));
    IPH_TRY(launch_separable(img, smooth, w, h, 1, n, IPH_F32, taps, taps, n_taps, IPH_F32,
                             sep_ws, sep_bytes, s));
    if (vec)
        sharpen_combine_kernel<true><<<blocks, 256, 0, s>>>(src, smooth, amount, n, dst);
    else
        sharpen_combine_kernel<false><<<blocks, 256, 0, s>>>(src, smooth, amount, n, dst);
    return launch_status("sharpen_combine_kernel");
}

// ---- gradients (gradients.rs:99-160) -------------------------------------------

namespace {
constexpr int kGradMaxTaps = 49;
struct GradTaps {
    int32_t h[kGradMaxTaps], v[kGradMaxTaps];
    int hw, hh, vw, vh;
};

// One 64 x 32 output tile per workgroup: the image window the two kernels
// read (clamp-to-edge coordinates, filter/mod.rs:29-69, centre (kw/2, kh/2))
// staged once in LDS, then per output both filter_clamped::<_, i32, i16>
// sums (exact in i32, so their order is free) saturated to i16
// (definitions.rs:181-188), and out = (dx^2 + dy^2).sqrt() as u16
// (gradient_magnitude, gradients.rs:157-160): f32 squares and sum each rounded
// once, correctly rounded sqrt, Rust's saturating float -> u16 cast.
constexpr int kGradTW = 64, kGradTH = 32;
constexpr int kGradSW = kGradTW + kGradMaxTaps - 1, kGradSH = kGradTH + kGradMaxTaps - 1;

// K3: both kernels 3 x 3 (loops unrolled, taps in scalar registers); M24:
// every |tap| < 2^23, so the u8 x tap products are exact 24-bit multiplies.
// Otherwise runtime-sized loops with 32-bit products (wrapping like the
// reference's i32 accumulator).
template <bool K3, bool M24>
__device__ __forceinline__ int grad_sum(const int *win, int pitch, int r, int c,
                                        const int32_t *k, int kw, int kh) {
    // products and sums in uint32: two's-complement wrapping, defined in C++,
    // the same bits as the reference's wrapping i32 arithmetic
    uint32_t acc = 0;
    if (K3) {
#pragma unroll
        for (int ky = 0; ky < 3; ky++)
#pragma unroll
            for (int kx = 0; kx < 3; kx++) {
                const int p = win[(r + ky) * pitch + c + kx], t = k[ky * 3 + kx];
                acc += M24 ? (uint32_t)__mul24(p, t) : (uint32_t)p * (uint32_t)t;
            }
    } else {
        for (int ky = 0; ky < kh; ky++) {
            const int *row = win + (r + ky) * pitch + c;
            for (int kx = 0; kx < kw; kx++) {
                const int p = row[kx], t = k[ky * kw + kx];
                acc += M24 ? (uint32_t)__mul24(p, t) : (uint32_t)p * (uint32_t)t;
            }
        }
    }
    return min(max((int)acc, -32768), 32767);
}

template <bool K3, bool M24>
__global__ __launch_bounds__(256) void gradients_kernel(const uint8_t *__restrict__ src, int W,
                                                        int H, const GradTaps taps,
                                                        uint16_t *__restrict__ out) {
    // one int per pixel: byte elements let the compiler merge neighbouring
    // reads into misaligned 16-bit LDS accesses, which stall badly
    __shared__ int win[K3 ? (kGradTH + 2) * (kGradTW + 2) : kGradSH * kGradSW];
    // window origin: the farthest left / up either kernel reaches
    const int lx0 = max(taps.hw / 2, taps.vw / 2), ly0 = max(taps.hh / 2, taps.vh / 2);
    const int rx = max(taps.hw - 1 - taps.hw / 2, taps.vw - 1 - taps.vw / 2);
    const int ry = max(taps.hh - 1 - taps.hh / 2, taps.vh - 1 - taps.vh / 2);
    const int sw = kGradTW + lx0 + rx, sh = kGradTH + ly0 + ry;
    const int X0 = blockIdx.x * kGradTW, Y0 = blockIdx.y * kGradTH;
    // a wave per staged row, a lane per column; every load of the thread is
    // issued before the LDS stores (a load-store loop would wait for each)
    constexpr int kRowsPerWave = ((K3 ? kGradTH + 2 : kGradSH) + 3) / 4;
    constexpr int kColsPerLane = ((K3 ? kGradTW + 2 : kGradSW) + 63) / 64;
    int v[kRowsPerWave][kColsPerLane];
#pragma unroll
    for (int k = 0; k < kRowsPerWave; k++) {
        const int r = (threadIdx.x >> 6) + 4 * k;
        const uint8_t *row = src + (size_t)min(max(Y0 - ly0 + r, 0), H - 1) * W;
#pragma unroll
        for (int u = 0; u < kColsPerLane; u++) {
            const int c = (threadIdx.x & 63) + 64 * u;
            v[k][u] = (r < sh && c < sw) ? row[min(max(X0 - lx0 + c, 0), W - 1)] : 0;
        }
    }
#pragma unroll
    for (int k = 0; k < kRowsPerWave; k++)
#pragma unroll
        for (int u = 0; u < kColsPerLane; u++) {
            const int r = (threadIdx.x >> 6) + 4 * k, c = (threadIdx.x & 63) + 64 * u;
            if (r < sh && c < sw) win[r * sw + c] = v[k][u];
        }
    __syncthreads();
    const int lx = threadIdx.x & 63, x = X0 + lx;
    if (x >= W) return;
    for (int r = threadIdx.x >> 6; r < kGradTH; r += 4) {
        const int y = Y0 + r;
        if (y >= H) break;
        // output (r, lx): kernel tap (0, 0) sits at window (r + ly0 - kh/2, lx + lx0 - kw/2)
        const float dx = (float)grad_sum<K3, M24>(win, sw, r + ly0 - taps.hh / 2,
                                                  lx + lx0 - taps.hw / 2, taps.h, taps.hw, taps.hh);
        const float dy = (float)grad_sum<K3, M24>(win, sw, r + ly0 - taps.vh / 2,
                                                  lx + lx0 - taps.vw / 2, taps.v, taps.vw, taps.vh);
        const float m = sqrtf(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)));
        out[(size_t)y * W + x] = m >= 65535.0f ? (uint16_t)65535 : (uint16_t)m;
    }
}
}  // namespace

iph_status launch_gradients(const uint8_t *src, uint32_t w, uint32_t h, const int32_t *hk,
                            uint32_t hkw, uint32_t hkh, const int32_t *vk, uint32_t vkw,
                            uint32_t vkh, uint16_t *out, hipStream_t s) {
    if (hkw * hkh > (uint32_t)kGradMaxTaps || vkw * vkh > (uint32_t)kGradMaxTaps) {
        set_error("gradients takes kernels of at most %d taps", kGradMaxTaps);
        return IPH_EUNSUPPORTED;
    }
    if ((size_t)w * h == 0) return IPH_OK;
    GradTaps t = {};
    for (uint32_t i = 0; i < hkw * hkh; i++) t.h[i] = hk[i];
    for (uint32_t i = 0; i < vkw * vkh; i++) t.v[i] = vk[i];
    t.hw = (int)hkw;
    t.hh = (int)hkh;
    t.vw = (int)vkw;
    t.vh = (int)vkh;
    if ((h + kGradTH - 1) / kGradTH > 65535u) {
        set_error("gradients supports heights up to %u", 65535u * kGradTH);
        return IPH_EUNSUPPORTED;
    }
    const bool k3 = hkw == 3 && hkh == 3 && vkw == 3 && vkh == 3;
    bool m24 = true;
    for (uint32_t i = 0; i < hkw * hkh; i++) m24 = m24 && hk[i] > -(1 << 23) && hk[i] < (1 << 23);
    for (uint32_t i = 0; i < vkw * vkh; i++) m24 = m24 && vk[i] > -(1 << 23) && vk[i] < (1 << 23);
    const dim3 grid((w + kGradTW - 1) / kGradTW, (h + kGradTH - 1) / kGradTH);
    if (k3 && m24)
        gradients_kernel<true, true><<<grid, 256, 0, s>>>(src, (int)w, (int)h, t, out);
    else if (k3)
        gradients_kernel<true, false><<<grid, 256, 0, s>>>(src, (int)w, (int)h, t, out);
    else if (m24)
        gradients_kernel<false, true><<<grid, 256, 0, s>>>(src, (int)w, (int)h, t, out);
    else
        gradients_kernel<false, false><<<grid, 256, 0, s>>>(src, (int)w, (int)h, t, out);
    return launch_status("gradients_kernel");
}

// ---- integral images (integral_image.rs:50-152) and adaptive_threshold ------------
//
// (W+1) x (H+1) u32 tables with a zero top row and left column, every sum
// modulo 2^32 (the reference's u32 additions wrap in release builds, and all
// arithmetic here is u32 too, so any order of additions gives the same bits).
// The image is cut into 64 x 64 tiles and read twice, written once:
//   1. tile_sums: per tile, its 64 column sums and 64 row sums;
//   2. scan_aux: column sums scanned down the tile rows (exclusive, A[ty][x]
//      = sum of column x above tile row ty), row sums scanned across the tile
//      columns (exclusive, R[tx][y] = sum of row y left of tile column tx),
//      tile totals scanned down the tile rows (exclusive, T[ty][tx]);
//   3. emit: per tile, out(y, x) = C + sum of A[ty][X0 .. x] + sum over the
//      tile's rows y' <= y of (R[tx][y'] + in-tile prefix of row y' up to
//      x), with C = sum over tx' < tx of T[ty][tx'] (everything above and
//      left of the tile).
// Workspace: tiles_y * W + tiles_x * H + tiles_x * tiles_y u32.

namespace {
constexpr int kIntTile = 64;

size_t int_tiles(uint32_t n) { return ((size_t)n + kIntTile - 1) / kIntTile; }

// Tile of work item b: tiles are dealt to the 8 XCDs (workgroup b runs on XCD
// b % 8 when the grid is a multiple of 8) in contiguous raster-order bands,
// so horizontally adjacent tiles, whose output rows share cache lines, meet
// in one L2.
__device__ inline size_t int_tile_of(size_t b, size_t ntiles) {
    const size_t per = (ntiles + 7) / 8;
    return (b % 8) * per + b / 8;
}

// Thread (r, q) of a tile: the 16 pixels of row y at columns X0 + 16 q ..
// X0 + 16 q + 15 (0 outside the image), squared for integral_squared_image.
// VEC: W % 16 == 0 and a 16-byte aligned image, so the 16 bytes are one
// aligned load, all inside or all outside the image.
template <bool SQ, bool VEC>
__device__ __forceinline__ void int_load16(const uint8_t *__restrict__ src, int W, int H, int X0,
                                           int y, int q, uint32_t v[16]) {
    const int x0 = X0 + q * 16;
    if (VEC) {
        uint4 d = make_uint4(0u, 0u, 0u, 0u);
        if (y < H && x0 < W) d = *reinterpret_cast<const uint4 *>(src + (size_t)y * W + x0);
        const uint32_t w4[4] = {d.x, d.y, d.z, d.w};
#pragma unroll
        for (int j = 0; j < 16; j++) v[j] = (w4[j >> 2] >> (8 * (j & 3))) & 0xffu;
    } else {
#pragma unroll
        for (int j = 0; j < 16; j++)
            v[j] = (y < H && x0 + j < W) ? (uint32_t)src[(size_t)y * W + x0 + j] : 0u;
    }
    if (SQ) {
#pragma unroll
        for (int j = 0; j < 16; j++) v[j] *= v[j];
    }
}

template <bool SQ, bool VEC>
__global__ __launch_bounds__(256) void integral_tile_sums_kernel(
    const uint8_t *__restrict__ src, int W, int H, int tiles_x, size_t ntiles,
    uint32_t *__restrict__ colsum, uint32_t *__restrict__ rowsum, uint32_t *__restrict__ tsum) {
    __shared__ uint4 tile4[kIntTile][kIntTile / 4 + 1];
    __shared__ uint32_t cpart[4][kIntTile];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = threadIdx.x >> 2, q = threadIdx.x & 3;
    const size_t nb = 8 * ((ntiles + 7) / 8);
    for (size_t b = blockIdx.x; b < nb; b += gridDim.x) {
        const size_t t = int_tile_of(b, ntiles);
        if (t >= ntiles) continue;  // the same for the whole workgroup
        const int tx = (int)(t % (size_t)tiles_x), ty = (int)(t / (size_t)tiles_x);
        const int X0 = tx * kIntTile, Y0 = ty * kIntTile;
        uint32_t v[16];
        int_load16<SQ, VEC>(src, W, H, X0, Y0 + r, q, v);
        uint32_t rs = 0;
#pragma unroll
        for (int j = 0; j < 16; j++) rs += v[j];
#pragma unroll
        for (int k = 0; k < 4; k++)
            tile4[r][q * 4 + k] = make_uint4(v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]);
        rs += __shfl_xor(rs, 1);
        rs += __shfl_xor(rs, 2);
        if (q == 0 && Y0 + r < H) rowsum[(size_t)tx * H + Y0 + r] = rs;
        __syncthreads();
        const uint32_t *col = reinterpret_cast<const uint32_t *>(&tile4[0][0]);
        uint32_t cs = 0;
#pragma unroll
        for (int i = 0; i < 16; i++) cs += col[(wave * 16 + i) * (kIntTile + 4) + lane];
        cpart[wave][lane] = cs;
        __syncthreads();
        if (wave == 0) {
            const uint32_t c = cpart[0][lane] + cpart[1][lane] + cpart[2][lane] + cpart[3][lane];
            if (X0 + lane < W) colsum[(size_t)ty * W + X0 + lane] = c;
            uint32_t tot = c;  // columns past W hold zeros
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o);
            if (lane == 0) tsum[t] = tot;
        }
        __syncthreads();
    }
}

// Threads [0, W) scan colsum[.][x] down the tile rows, threads [W, W + H)
// scan rowsum[.][y] across the tile columns, threads [W + H, W + H + tiles_x)
// scan tsum[.][tx] down the tile rows; all exclusive and in place, loads
// issued 32 at a time.
__global__ __launch_bounds__(256) void integral_scan_aux_kernel(uint32_t *__restrict__ colsum,
                                                                int W, int tiles_y,
                                                                uint32_t *__restrict__ rowsum,
                                                                int H, int tiles_x,
                                                                uint32_t *__restrict__ tsum) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    uint32_t *p;
    size_t stride;
    int n;
    if (i < (size_t)W) {
        p = colsum + i;
        stride = (size_t)W;
        n = tiles_y;
    } else if (i < (size_t)W + H) {
        p = rowsum + (i - W);
        stride = (size_t)H;
        n = tiles_x;
    } else if (i < (size_t)W + H + tiles_x) {
        p = tsum + (i - W - H);
        stride = (size_t)tiles_x;
        n = tiles_y;
    } else {
        return;
    }
    uint32_t run = 0;
    for (int k0 = 0; k0 < n; k0 += 32) {
        uint32_t v[32];
#pragma unroll
        for (int u = 0; u < 32; u++) v[u] = k0 + u < n ? p[(size_t)(k0 + u) * stride] : 0u;
#pragma unroll
        for (int u = 0; u < 32; u++)
            if (k0 + u < n) {
                p[(size_t)(k0 + u) * stride] = run;
                run += v[u];
            }
    }
}

// V4OUT: `out` is 16-byte aligned, so every output row is written as aligned
// 4-element groups from LDS plus at most three elements at each end.
template <bool SQ, bool VEC, bool V4OUT>
__global__ __launch_bounds__(256) void integral_emit_kernel(
    const uint8_t *__restrict__ src, int W, int H, int tiles_x, size_t ntiles,
    const uint32_t *__restrict__ above, const uint32_t *__restrict__ rleft,
    const uint32_t *__restrict__ tabove, uint32_t *__restrict__ out) {
    __shared__ uint4 tile4[kIntTile][kIntTile / 4 + 1];
    __shared__ uint32_t vpart[4][kIntTile];
    __shared__ uint32_t cpart[4];
    uint32_t *tile = reinterpret_cast<uint32_t *>(&tile4[0][0]);
    constexpr int P = kIntTile + 4;  // row pitch of `tile` in elements
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = threadIdx.x >> 2, q = threadIdx.x & 3;
    const size_t W1 = (size_t)W + 1;
    const size_t nb = 8 * ((ntiles + 7) / 8);
    for (size_t b = blockIdx.x; b < nb; b += gridDim.x) {
        const size_t t = int_tile_of(b, ntiles);
        if (t >= ntiles) continue;  // the same for the whole workgroup
        const int tx = (int)(t % (size_t)tiles_x), ty = (int)(t / (size_t)tiles_x);
        const int X0 = tx * kIntTile, Y0 = ty * kIntTile;
        const int cw = min(kIntTile, W - X0), chh = min(kIntTile, H - Y0);
        // row prefix of the thread's 16 pixels, the quarter-row offsets and
        // the part of the row left of the tile
        uint32_t v[16];
        int_load16<SQ, VEC>(src, W, H, X0, Y0 + r, q, v);
        // everything above the tile row in columns X0 .. X0 + lane: C (the
        // tile totals above-left, summed by the workgroup) plus the prefix of
        // A[ty][X0 ..] across the lanes
        uint32_t corner = 0;
        for (int k = threadIdx.x; k < tx; k += 256) corner += tabove[(size_t)ty * tiles_x + k];
        uint32_t col_above = lane < cw ? above[(size_t)ty * W + X0 + lane] : 0u;
#pragma unroll
        for (int j = 1; j < 16; j++) v[j] += v[j - 1];
        uint32_t a = v[15];
        uint32_t u = __shfl_up(a, 1);
        if (q >= 1) a += u;
        u = __shfl_up(a, 2);
        if (q >= 2) a += u;
        const uint32_t off = (a - v[15]) + (r < chh ? rleft[(size_t)tx * H + Y0 + r] : 0u);
#pragma unroll
        for (int k = 0; k < 4; k++)
            tile4[r][q * 4 + k] = make_uint4(v[4 * k] + off, v[4 * k + 1] + off,
                                             v[4 * k + 2] + off, v[4 * k + 3] + off);
        __syncthreads();
        // column prefix down the tile, plus everything above the tile row
        uint32_t c[16], run = 0;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            run += tile[(wave * 16 + i) * P + lane];
            c[i] = run;
        }
        vpart[wave][lane] = run;
        __syncthreads();
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) corner += __shfl_xor(corner, o);
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t u2 = __shfl_up(col_above, o);
            if (lane >= o) col_above += u2;
        }
        if (lane == 0) cpart[wave] = corner;
        __syncthreads();
        uint32_t base = cpart[0] + cpart[1] + cpart[2] + cpart[3] + col_above;
        for (int w2 = 0; w2 < wave; w2++) base += vpart[w2][lane];
        if (!V4OUT) {
            if (lane < cw) {
#pragma unroll
                for (int i = 0; i < 16; i++)
                    if (wave * 16 + i < chh)
                        out[(size_t)(Y0 + wave * 16 + i + 1) * W1 + X0 + lane + 1] = c[i] + base;
            }
        } else {
#pragma unroll
            for (int i = 0; i < 16; i++) tile[(wave * 16 + i) * P + lane] = c[i] + base;
            __syncthreads();
            // aligned groups of 4 fully inside the row segment
            for (int slot = threadIdx.x; slot < kIntTile * 16; slot += 256) {
                const int rr = slot >> 4, g = slot & 15;
                if (rr >= chh) continue;
                const size_t e0 = (size_t)(Y0 + rr + 1) * W1 + X0 + 1;
                const int head = (int)((4 - (e0 & 3)) & 3);
                const int nfull = cw > head ? (cw - head) >> 2 : 0;
                if (g >= nfull) continue;
                const int c0 = head + 4 * g;
                const uint32_t *s4 = tile + rr * P + c0;
                *reinterpret_cast<uint4 *>(out + e0 + c0) = make_uint4(s4[0], s4[1], s4[2], s4[3]);
            }
            // up to three elements before the first group and after the last
            for (int slot = threadIdx.x; slot < kIntTile * 6; slot += 256) {
                const int rr = slot / 6, k = slot % 6;
                if (rr >= chh) continue;
                const size_t e0 = (size_t)(Y0 + rr + 1) * W1 + X0 + 1;
                const int head = (int)((4 - (e0 & 3)) & 3);
                const int nfull = cw > head ? (cw - head) >> 2 : 0;
                const int col = k < 3 ? (k < head ? k : kIntTile) : head + 4 * nfull + (k - 3);
                if (col < cw) out[e0 + col] = tile[rr * P + col];
            }
        }
        // the zero top row and left column
        if (ty == 0 && wave == 0) {
            if (lane < cw) out[X0 + lane + 1] = 0u;
            if (tx == 0 && lane == 0) out[0] = 0u;
        }
        if (tx == 0 && lane == 0) {
            for (int i = 0; i < 16; i++)
                if (wave * 16 + i < chh) out[(size_t)(Y0 + wave * 16 + i + 1) * W1] = 0u;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void zero_u32_kernel(uint32_t *__restrict__ p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        p[i] = 0u;
}

// contrast.rs:19-47: the block mean from the integral image (sum_image_pixels,
// integral_image.rs:262-268, u32 wrapping, then u32 division) and
// pixel >= mean - delta (i32, wrapping) -> 255, else 0.
__global__ __launch_bounds__(256) void adaptive_threshold_kernel(const uint8_t *__restrict__ src,
                                                                 const uint32_t *__restrict__ ii,
                                                                 int W, int H, int r, int delta,
                                                                 uint8_t *__restrict__ dst) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    if (x >= W) return;
    for (int y = blockIdx.y * 4 + (threadIdx.x >> 6); y < H; y += gridDim.y * 4) {
    const int ylo = (int)max(0LL, (long long)y - r);
    const int yhi = (int)min((long long)H - 1, (long long)y + r);
    const int xlo = (int)max(0LL, (long long)x - r);
    const int xhi = (int)min((long long)W - 1, (long long)x + r);
    const size_t W1 = (size_t)W + 1;
    const uint32_t a = ii[(size_t)(yhi + 1) * W1 + xhi + 1], b = ii[(size_t)ylo * W1 + xlo];
    const uint32_t c = ii[(size_t)ylo * W1 + xhi + 1], d = ii[(size_t)(yhi + 1) * W1 + xlo];
    const uint32_t count = (uint32_t)(yhi - ylo + 1) * (uint32_t)(xhi - xlo + 1);
    const uint32_t mean = (a + b - c - d) / count;
    const int32_t thr = (int32_t)((uint32_t)mean - (uint32_t)delta);
    dst[(size_t)y * W + x] = (int32_t)src[(size_t)y * W + x] >= thr ? 255 : 0;
    }
}
}  // namespace

size_t integral_workspace(uint32_t w, uint32_t h) {
    const size_t n = int_tiles(h) * w + int_tiles(w) * h + int_tiles(w) * int_tiles(h);
    return (n * sizeof(uint32_t) + 255) & ~(size_t)255;
}

iph_status launch_integral_image(const uint8_t *src, uint32_t w, uint32_t h, int squared,
                                 uint32_t *out, void *ws, size_t wsb, hipStream_t s) {
    const size_t w1 = (size_t)w + 1;
    if (w == 0 || h == 0) {  // all zeros, integral_image.rs:110-112
        const size_t n = w1 * (h + 1);
        const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
        zero_u32_kernel<<<grid, 256, 0, s>>>(out, n);
        return launch_status("zero_u32_kernel");
    }
    if (wsb < integral_workspace(w, h)) {
        set_error("integral image workspace of %zu bytes is below the %zu needed", wsb,
                  integral_workspace(w, h));
        return IPH_EINVAL;
    }
    const int tx = (int)int_tiles(w), ty = (int)int_tiles(h);
    uint32_t *colsum = static_cast<uint32_t *>(ws);
    uint32_t *rowsum = colsum + (size_t)ty * w;
    uint32_t *tsum = rowsum + (size_t)tx * h;
    const size_t ntiles = (size_t)tx * ty;
    // one workgroup per tile up to the cap (a multiple of 8 keeps the XCD bands)
    const unsigned grid = (unsigned)std::min<size_t>(8 * ((ntiles + 7) / 8), 1u << 20);
    const bool vec = w % 16 == 0 && (reinterpret_cast<uintptr_t>(src) & 15) == 0;
    const bool v4out = (reinterpret_cast<uintptr_t>(out) & 15) == 0;
    const int W = (int)w, H = (int)h;
#define IPH_INT_SUMS(SQ, VEC)                                                                  \
    integral_tile_sums_kernel<SQ, VEC><<<grid, 256, 0, s>>>(src, W, H, tx, ntiles, colsum, rowsum, \
                                                            tsum)
    if (squared) {
        if (vec) IPH_INT_SUMS(true, true); else IPH_INT_SUMS(true, false);
    } else {
        if (vec) IPH_INT_SUMS(false, true); else IPH_INT_SUMS(false, false);
    }
#undef IPH_INT_SUMS
    IPH_TRY(launch_status("integral_tile_sums_kernel"));
    integral_scan_aux_kernel<<<(unsigned)(((size_t)w + h + tx + 255) / 256), 256, 0, s>>>(
        colsum, W, ty, rowsum, H, tx, tsum);
    IPH_TRY(launch_status("integral_scan_aux_kernel"));
#define IPH_INT_EMIT(SQ, VEC, V4)                                                               \
    integral_emit_kernel<SQ, VEC, V4><<<grid, 256, 0, s>>>(src, W, H, tx, ntiles, colsum, rowsum, \
                                                          tsum, out)
    if (squared) {
        if (vec) {
            if (v4out) IPH_INT_EMIT(true, true, true); else IPH_INT_EMIT(true, true, false);
        } else {
            if (v4out) IPH_INT_EMIT(true, false, true); else IPH_INT_EMIT(true, false, false);
        }
    } else {
        if (vec) {
            if (v4out) IPH_INT_EMIT(false, true, true); else IPH_INT_EMIT(false, true, false);
        } else {
            if (v4out) IPH_INT_EMIT(false, false, true); else IPH_INT_EMIT(false, false, false);
        }
    }
#undef IPH_INT_EMIT
    return launch_status("integral_emit_kernel");
}

// ---- adaptive_threshold for block radii up to 64: block sums in LDS ----------------
//
// One 64 x 64 output tile per workgroup. The tile's source window (radius r
// around it; rows and columns outside the image staged as zeros, which add
// nothing to a clipped block) is staged in LDS; vertical window sums of 2r+1
// rows are slid down each column four columns at a time (two u16 halves per
// u32: a sum of at most 129 bytes stays below 2^16) into registers, then
// written over the staged window (so the workgroup's LDS is the larger of the
// two, not their sum); horizontal window sums of 2r+1 of those are slid along
// each row. The block sums are exact
// (below 2^32), equal to the reference's wrapping four-corner sums, and the
// mean, threshold and comparison follow contrast.rs:19-47.
namespace {
constexpr int kAdT = 64;     // output tile side
constexpr int kAdMaxR = 64;  // largest radius of this path

struct AdShape {
    int r, off, sw, sr;  // radius; (X0 - r) - xs; staged columns (multiple of 4); staged rows
    // V rows are sw + 2 u16 apart: an odd dword pitch, so the 16 rows one
    // wave reads at a time fall in distinct banks
    size_t lds() const { return std::max((size_t)sw * sr, (size_t)kAdT * (sw + 2) * 2); }
};

AdShape ad_shape(int r) {
    AdShape a;
    a.r = r;
    a.off = (4 - (r & 3)) & 3;  // X0 is a multiple of 64, so (X0 - r) mod 4 = (-r) mod 4
    a.sw = (a.off + kAdT + 2 * r + 3) & ~3;
    a.sr = kAdT + 2 * r;
    return a;
}

template <bool VEC>
__global__ __launch_bounds__(256) void adaptive_block_kernel(const uint8_t *__restrict__ src,
                                                             int W, int H, AdShape a, int delta,
                                                             int tiles_x,
                                                             uint8_t *__restrict__ dst) {
    extern __shared__ __align__(16) unsigned char ad_smem[];
    uint8_t *S = ad_smem;                                               // sr x sw bytes
    uint16_t *V = reinterpret_cast<uint16_t *>(ad_smem);  // 64 x (sw + 2), after S is read
    const int r = a.r, sw = a.sw, sr = a.sr, off = a.off;
    const int X0 = (int)(blockIdx.x % (unsigned)tiles_x) * kAdT;
    const int Y0 = (int)(blockIdx.x / (unsigned)tiles_x) * kAdT;
    const int xs = X0 - r - off, ys = Y0 - r;  // image position of S[0][0] (xs % 4 == 0)
    const int quads = sw >> 2;
    // stage the window, one dword (4 columns) per item: thread (tr, tq) takes
    // column quad tq of rows tr, tr + rpp, ... (one division per thread, not
    // per item), twelve loads in flight before their LDS stores
    const int rpp = 256 / quads, tq = (int)threadIdx.x % quads, tr = (int)threadIdx.x / quads;
    const int x = xs + 4 * tq;
    const bool xin = VEC && x >= 0 && x + 4 <= W;
    for (int k0 = tr; k0 < sr; k0 += 12 * rpp) {
        uint32_t v[12];
#pragma unroll
        for (int u = 0; u < 12; u++) {
            const int k = k0 + u * rpp, y = ys + k;
            v[u] = 0;
            if (tr < rpp && k < sr && y >= 0 && y < H) {
                const uint8_t *row = src + (size_t)y * W;
                if (xin) {
                    v[u] = *reinterpret_cast<const uint32_t *>(row + x);
                } else {
#pragma unroll
                    for (int b = 0; b < 4; b++)
                        if (x + b >= 0 && x + b < W) v[u] |= (uint32_t)row[x + b] << (8 * b);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < 12; u++) {
            const int k = k0 + u * rpp;
            if (tr < rpp && k < sr) reinterpret_cast<uint32_t *>(S)[k * quads + tq] = v[u];
        }
    }
    __syncthreads();
    // vertical sums: item (quad, 16-row segment), at most one per thread
    // (quads <= 49), kept in registers until every item has read S
    const int vi = threadIdx.x, vq = vi % quads, vseg = vi / quads;
    uint32_t vlo[16], vhi[16];
    if (vi < quads * 4) {
        const int q = vq, seg = vseg;
        const uint32_t *col = reinterpret_cast<const uint32_t *>(S) + q;
        const int qs = sw >> 2;  // dwords per staged row
        uint32_t lo = 0, hi = 0;  // columns (0, 2) and (1, 3) of the quad
        const int y0 = 16 * seg;
#pragma unroll 8
        for (int k = y0; k <= y0 + 2 * r; k++) {
            const uint32_t w = col[k * qs];
            lo += w & 0x00ff00ffu;
            hi += (w >> 8) & 0x00ff00ffu;
        }
#pragma unroll
        for (int u = 0; u < 16; u++) {
            vlo[u] = lo;
            vhi[u] = hi;
            if (u < 15) {
                const uint32_t add = col[(y0 + u + 2 * r + 1) * qs], sub = col[(y0 + u) * qs];
                lo += (add & 0x00ff00ffu) - (sub & 0x00ff00ffu);
                hi += ((add >> 8) & 0x00ff00ffu) - ((sub >> 8) & 0x00ff00ffu);
            }
        }
    }
    // the tile's own pixels for the comparison, from global memory (S is
    // about to be overwritten): thread (row, 16-column segment)
    const int yy = threadIdx.x >> 2, seg = threadIdx.x & 3;
    const int y = Y0 + yy, x0 = X0 + 16 * seg;
    uint32_t pix[4] = {0u, 0u, 0u, 0u};
    if (y < H) {
        const uint8_t *srow = src + (size_t)y * W;
        if (VEC && x0 + 16 <= W) {
            const uint4 p4 = *reinterpret_cast<const uint4 *>(srow + x0);
            pix[0] = p4.x;
            pix[1] = p4.y;
            pix[2] = p4.z;
            pix[3] = p4.w;
        } else {
#pragma unroll
            for (int u = 0; u < 16; u++)
                if (x0 + u < W) pix[u >> 2] |= (uint32_t)srow[x0 + u] << (8 * (u & 3));
        }
    }
    __syncthreads();  // every vertical item has read S
    if (vi < quads * 4) {
#pragma unroll
        for (int u = 0; u < 16; u++) {
            uint32_t *vr = reinterpret_cast<uint32_t *>(V + (size_t)(16 * vseg + u) * (sw + 2) + 4 * vq);
            vr[0] = (vlo[u] & 0xffffu) | (vhi[u] << 16);
            vr[1] = (vlo[u] >> 16) | (vhi[u] & 0xffff0000u);
        }
    }
    __syncthreads();
    // horizontal sums, mean and comparison: item (row, 16-column segment)
    if (y >= H) return;
    const uint16_t *vr = V + (size_t)yy * (sw + 2);
    const int j0 = 16 * seg + off;  // staged column of the first window's left edge
    // the first window's sum, two u16 per LDS read: pairs cover [j0 & ~1,
    // (j0 + 2r) | 1], then the one or two values outside the window come off
    const int ja = j0 & ~1, jb = (j0 + 2 * r) | 1;
    const uint32_t *vp = reinterpret_cast<const uint32_t *>(vr + ja);  // row pitch is even
    uint32_t sum = 0;
#pragma unroll 8
    for (int k = 0; k <= (jb - ja) >> 1; k++) {
        const uint32_t d = vp[k];
        sum += (d & 0xffffu) + (d >> 16);
    }
    if (ja < j0) sum -= vr[ja];
    if (jb > j0 + 2 * r) sum -= vr[jb];
    const uint32_t rows = (uint32_t)(min(H - 1, y + r) - max(0, y - r) + 1);
    uint32_t packed[4] = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int u = 0; u < 16; u++) {
        const int xl = 16 * seg + u, x = X0 + xl;
        const uint32_t cols = (uint32_t)(min(W - 1, x + r) - max(0, x - r) + 1);
        // floor(sum / n): sum < 2^23 and n < 2^15 are exact in f32 and the
        // quotient (a mean of bytes) is at most 255, so the f32 estimate is
        // within 1e-4 of it; one correction step each way makes it exact
        // (24-bit multiplies: n <= 129^2 and mean <= 256)
        const uint32_t n = __umul24(rows, cols);
        uint32_t mean = (uint32_t)((float)sum * __builtin_amdgcn_rcpf((float)n));
        mean -= __umul24(mean, n) > sum ? 1u : 0u;
        mean += __umul24(mean + 1, n) <= sum ? 1u : 0u;
        const int32_t thr = (int32_t)(mean - (uint32_t)delta);
        const int32_t p = (int32_t)((pix[u >> 2] >> (8 * (u & 3))) & 0xffu);
        const uint32_t o = p >= thr ? 255u : 0u;
        packed[u >> 2] |= o << (8 * (u & 3));
        if (u < 15) sum += (uint32_t)vr[j0 + u + 2 * r + 1] - (uint32_t)vr[j0 + u];
    }
    uint8_t *orow = dst + (size_t)y * W;
    if (VEC && x0 + 16 <= W) {
        *reinterpret_cast<uint4 *>(orow + x0) = make_uint4(packed[0], packed[1], packed[2], packed[3]);
    } else {
#pragma unroll
        for (int u = 0; u < 16; u++)
            if (x0 + u < W) orow[x0 + u] = (uint8_t)(packed[u >> 2] >> (8 * (u & 3)));
    }
}
}  // namespace

size_t adaptive_threshold_workspace(uint32_t w, uint32_t h, uint32_t block_radius) {
    if (std::min<uint32_t>(block_radius, std::max(w, h)) <= (uint32_t)kAdMaxR) return 0;
    const size_t table = (((size_t)w + 1) * ((size_t)h + 1) * sizeof(uint32_t) + 255) & ~(size_t)255;
    return table + integral_workspace(w, h);
}

iph_status launch_adaptive_threshold(const uint8_t *src, uint32_t w, uint32_t h,
                                     uint32_t block_radius, int32_t delta, uint8_t *dst, void *ws,
                                     size_t wsb, hipStream_t s) {
    if (w == 0 || h == 0) return IPH_OK;
    // A radius past the image's larger side clips to the same block.
    const int r = (int)std::min<uint32_t>(block_radius, std::max(w, h));
    if (r <= kAdMaxR) {
        const AdShape sh = ad_shape(r);
        const int tiles_x = (int)((w + kAdT - 1) / kAdT);
        const size_t ntiles = (size_t)tiles_x * ((h + kAdT - 1) / kAdT);
        if (ntiles > 0x7fffffffu) {
            set_error("adaptive threshold of %ux%u pixels is too large", w, h);
            return IPH_EUNSUPPORTED;
        }
        const bool vec = w % 16 == 0 && (reinterpret_cast<uintptr_t>(src) & 15) == 0 &&
                         (reinterpret_cast<uintptr_t>(dst) & 15) == 0;
        if (vec)
            adaptive_block_kernel<true><<<(unsigned)ntiles, 256, sh.lds(), s>>>(
                src, (int)w, (int)h, sh, delta, tiles_x, dst);
        else
            adaptive_block_kernel<false><<<(unsigned)ntiles, 256, sh.lds(), s>>>(
                src, (int)w, (int)h, sh, delta, tiles_x, dst);
        return launch_status("adaptive_block_kernel");
    }
    if (wsb < adaptive_threshold_workspace(w, h, block_radius)) {
        set_error("adaptive threshold workspace of %zu bytes is below the %zu needed", wsb,
                  adaptive_threshold_workspace(w, h, block_radius));
        return IPH_EINVAL;
    }
    const size_t table =
        (((size_t)w + 1) * ((size_t)h + 1) * sizeof(uint32_t) + 255) & ~(size_t)255;
    uint32_t *ii = static_cast<uint32_t *>(ws);
    IPH_TRY(launch_integral_image(src, w, h, 0, ii, static_cast<char *>(ws) + table, wsb - table,
                                  s));
    const unsigned gy = std::min((h + 3) / 4, 65535u);
    adaptive_threshold_kernel<<<dim3((w + 63) / 64, gy), 256, 0, s>>>(
        src, ii, (int)w, (int)h, r, delta, dst);
    return launch_status("adaptive_threshold_kernel");
}

}  // namespace iph
