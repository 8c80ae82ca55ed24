// Learned match proposal. Nothing here has a counterpart in the reference:
// its only search is the full score map of match_template
// (template_matching.rs:78-92) read by find_extremes (:479-511). This file
// holds the pieces of a cheaper search that ends in the same answer:
//
//   * a 3x3 convolution layer on 16x16 patches of 32 channels (bf16 NHWC in,
//     f32 accumulation on v_mfma_f32_32x32x16_bf16, implicit im2col from
//     LDS), its weight/bias gradient (an MFMA reduction over pixels) and
//     its input gradient (the same layer with flipped, transposed weights),
//     ReLU backward, softmax cross-entropy over a pair's candidate logits,
//     and Adam: the training and inference kernels of the proposal network;
//   * 4x downsampling of a GrayImage, coarse-map tile maxima and their top K,
//     per-candidate feature patches, the top candidates of the network's
//     logits, and exact scores in windows around them, computed from integer
//     sums with the full map's own normalisation, so a window's score is the
//     full map's value at that place.
#include "iph_internal.hpp"
#include "match_common.hpp"

#include <algorithm>
#include <cmath>

namespace iph {
namespace {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

constexpr int kC = IPH_CNN_CHANNELS;  // channels of every layer (padded)
constexpr int kPS = IPH_CNN_PATCH;    // patch side
constexpr int kPix = kPS * kPS;       // pixels per patch
constexpr int kKK = 9 * kC;           // im2col length of a 3x3 layer
constexpr int kXP = kPS + 2;          // patch side with its zero halo
constexpr int kXPitch = 40;           // LDS elements per staged pixel (80 bytes: spreads banks)
constexpr int kWPitch = kKK + 8;      // LDS elements per staged weight row
constexpr int kXTPitch = kXP * kXP + 4;  // transposed staging: elements per channel
constexpr int kDzPitch = kPix + 8;

static_assert(kC == 32 && kPS == 16, "the MFMA tiling assumes 32 channels and 16x16 patches");

__device__ __forceinline__ uint16_t to_bf16(float x) {
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);  // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// ---- network layers ------------------------------------------------------------

// Layer l's weights w[co][k] into LDS rows of kWPitch elements.
__device__ __forceinline__ void stage_weights(uint16_t *ws, const uint16_t *__restrict__ w, int tid) {
    for (int i = tid; i < kC * (kKK / 8); i += 256) {
        const int r = i / (kKK / 8), q = i - r * (kKK / 8);
        *reinterpret_cast<uint4 *>(ws + r * kWPitch + 8 * q) =
            *reinterpret_cast<const uint4 *>(w + r * kKK + 8 * q);
    }
}

// One 32-pixel x 32-channel tile (pixels 32 mt .. 32 mt + 31) of a 3x3
// layer from a halo-padded input patch xs and staged weights ws: A row m =
// lane n is the tile's pixel n (im2col taken straight from LDS), B column n
// the output channel n, 18 MFMAs of K = 16 over the 9 taps x 32 inputs.
__device__ __forceinline__ v16f conv_tile(const uint16_t *xs, const uint16_t *ws, int mt, int n,
                                          int g) {
    const int y = 2 * mt + (n >> 4), xq = n & 15;
    v16f acc = {};
#pragma unroll
    for (int s = 0; s < 2 * 9; s++) {
        const int tap = s >> 1, ky = tap / 3, kx = tap - 3 * ky, ci0 = 16 * (s & 1) + 8 * g;
        const v8bf av =
            *reinterpret_cast<const v8bf *>(xs + ((y + ky) * kXP + xq + kx) * kXPitch + ci0);
        const v8bf bv = *reinterpret_cast<const v8bf *>(ws + n * kWPitch + 16 * s + 8 * g);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);
    }
    return acc;
}

// D of conv_tile: register r of lane (n, g) holds output channel n of pixel
// 32 mt + 8 (r / 4) + 4 g + r % 4.
__device__ __forceinline__ int tile_pixel(int mt, int g, int r) {
    return 32 * mt + 8 * (r >> 2) + 4 * g + (r & 3);
}

// y = conv3x3(x) + bias over one 16x16 patch per workgroup (zero padding).
// Weights w[co][(ky*3 + kx)*32 + ci]. z (f32, optional) gets y, a (bf16,
// optional) relu(y) or y.
__global__ __launch_bounds__(256) void cnn_conv3x3_kernel(const uint16_t *__restrict__ x,
                                                          const uint16_t *__restrict__ w,
                                                          const float *__restrict__ bias, int relu,
                                                          float *__restrict__ z,
                                                          uint16_t *__restrict__ a) {
    __shared__ __align__(16) uint16_t xs[kXP * kXP * kXPitch];
    __shared__ __align__(16) uint16_t ws[kC * kWPitch];
    const int p = blockIdx.x, tid = threadIdx.x;
    stage_weights(ws, w, tid);
    const uint16_t *xp = x + (size_t)p * kPix * kC;
    for (int i = tid; i < kXP * kXP * 4; i += 256) {
        const int pix = i >> 2, q = i & 3;
        const int yy = pix / kXP - 1, xx = pix % kXP - 1;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (yy >= 0 && yy < kPS && xx >= 0 && xx < kPS)
            v = *reinterpret_cast<const uint4 *>(xp + (yy * kPS + xx) * kC + 8 * q);
        *reinterpret_cast<uint4 *>(xs + pix * kXPitch + 8 * q) = v;
    }
    __syncthreads();
    const int lane = tid & 63, wave = tid >> 6, n = lane & 31, g = lane >> 5;
    const float b = bias ? bias[n] : 0.0f;
    for (int mt = 2 * wave; mt < 2 * wave + 2; mt++) {
        const v16f acc = conv_tile(xs, ws, mt, n, g);
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int pix = tile_pixel(mt, g, r);
            const float v = acc[r] + b;
            const size_t o = ((size_t)p * kPix + pix) * kC + n;
            if (z) z[o] = v;
            if (a) a[o] = to_bf16(relu ? fmaxf(v, 0.0f) : v);
        }
    }
}

// wf[ci][(tap')*32 + co] = w[co][(8 - tap')*32 + ci]: the layer whose
// forward pass is the input gradient of w.
__global__ __launch_bounds__(256) void cnn_flip_kernel(const uint16_t *__restrict__ w,
                                                       uint16_t *__restrict__ wf) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= kC * kKK) return;
    const int ci = i / kKK, k = i - ci * kKK, tap = k / kC, co = k - tap * kC;
    wf[i] = w[co * kKK + (8 - tap) * kC + ci];
}

// Partial weight and bias gradients of one layer: workgroup b sums patches
// b, b + grid, ... into partial[b][co*288 + tap*32 + ci] and
// partial[b][9216 + co].
__global__ __launch_bounds__(256) void cnn_conv3x3_dw_kernel(const uint16_t *__restrict__ x,
                                                             const uint16_t *__restrict__ dz,
                                                             int n_patches,
                                                             float *__restrict__ partial) {
    __shared__ __align__(16) uint16_t xt[kC * kXTPitch];  // channel-major, halo-padded
    __shared__ __align__(16) uint16_t dzt[kC * kDzPitch];
    __shared__ float bsum[8][kC];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, n = lane & 31, g = lane >> 5;
    v16f acc[3] = {};
    float bacc = 0.0f;
    for (int p = blockIdx.x; p < n_patches; p += gridDim.x) {
        __syncthreads();
        const uint16_t *xp = x + (size_t)p * kPix * kC;
        for (int i = tid; i < kXP * kXP * kC; i += 256) {
            const int pix = i / kC, ch = i - pix * kC;
            const int yy = pix / kXP - 1, xx = pix % kXP - 1;
            uint16_t v = 0;
            if (yy >= 0 && yy < kPS && xx >= 0 && xx < kPS) v = xp[(yy * kPS + xx) * kC + ch];
            xt[ch * kXTPitch + pix] = v;
        }
        const uint16_t *dp = dz + (size_t)p * kPix * kC;
        for (int i = tid; i < kPix * kC; i += 256) {
            const int pix = i / kC, ch = i - pix * kC;
            dzt[ch * kDzPitch + pix] = dp[i];
        }
        __syncthreads();
        {
            const int ch = tid & 31, part = tid >> 5;
            float s2 = 0.0f;
            for (int q = 0; q < 32; q++) s2 += from_bf16(dzt[ch * kDzPitch + part * 32 + q]);
            bacc += s2;
        }
        // M = output channel (A = dz^T), N = input channel of tap wave + 4t (B = shifted x),
        // K = the 16 pixels of row s
        for (int s = 0; s < kPS; s++) {
            const v8bf av = *reinterpret_cast<const v8bf *>(dzt + n * kDzPitch + 16 * s + 8 * g);
#pragma unroll
            for (int t = 0; t < 3; t++) {
                const int tap = wave + 4 * t;
                if (tap > 8) break;
                const int ky = tap / 3, kx = tap - 3 * ky;
                const uint16_t *src = xt + n * kXTPitch + (s + ky) * kXP + 8 * g + kx;
                v8bf bv;
#pragma unroll
                for (int j = 0; j < 8; j++) bv[j] = __builtin_bit_cast(__bf16, src[j]);
                acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc[t], 0, 0, 0);
            }
        }
    }
    float *out = partial + (size_t)blockIdx.x * (kC * kKK + kC);
#pragma unroll
    for (int t = 0; t < 3; t++) {
        const int tap = wave + 4 * t;
        if (tap > 8) break;
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int co = 8 * (r >> 2) + 4 * g + (r & 3);
            out[co * kKK + tap * kC + n] = acc[t][r];
        }
    }
    bsum[tid >> 5][tid & 31] = bacc;
    __syncthreads();
    if (tid < kC) {
        float s2 = 0.0f;
        for (int q = 0; q < 8; q++) s2 += bsum[q][tid];
        out[kC * kKK + tid] = s2;
    }
}

__global__ __launch_bounds__(256) void cnn_reduce_partials_kernel(const float *__restrict__ partial,
                                                                  int n_partials,
                                                                  float *__restrict__ dw,
                                                                  float *__restrict__ db) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    constexpr int kN = kC * kKK + kC;
    if (i >= kN) return;
    float s = 0.0f;
    for (int q = 0; q < n_partials; q++) s += partial[(size_t)q * kN + i];
    if (i < kC * kKK) dw[i] = s;
    else db[i - kC * kKK] = s;
}

__global__ __launch_bounds__(256) void cnn_relu_backward_kernel(const float *__restrict__ dx,
                                                                const float *__restrict__ z,
                                                                size_t n,
                                                                uint16_t *__restrict__ dz) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) dz[i] = to_bf16(z[i] > 0.0f ? dx[i] : 0.0f);
}

// Softmax cross-entropy of one pair per workgroup over the logits (channel 0)
// of its kpp patches; target = flat index patch*256 + pixel, or -1 (no
// gradient). dz gets (softmax - onehot) * scale in channel 0, zeros elsewhere.
__global__ __launch_bounds__(256) void cnn_softmax_xent_kernel(const float *__restrict__ zl,
                                                               int kpp,
                                                               const int *__restrict__ target,
                                                               float scale,
                                                               float *__restrict__ loss,
                                                               uint16_t *__restrict__ dz) {
    __shared__ float red[256];
    const int pair = blockIdx.x, tid = threadIdx.x, n = kpp * kPix;
    const size_t base = (size_t)pair * n;
    const int tgt = target[pair];
    float mx = -INFINITY;
    for (int i = tid; i < n; i += 256) mx = fmaxf(mx, zl[(base + i) * kC]);
    red[tid] = mx;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o) red[tid] = fmaxf(red[tid], red[tid + o]);
        __syncthreads();
    }
    mx = red[0];
    __syncthreads();
    float se = 0.0f;
    for (int i = tid; i < n; i += 256) se += expf(zl[(base + i) * kC] - mx);
    red[tid] = se;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o) red[tid] += red[tid + o];
        __syncthreads();
    }
    se = red[0];
    if (tid == 0) loss[pair] = tgt >= 0 ? logf(se) + mx - zl[(base + tgt) * kC] : 0.0f;
    for (int i = tid; i < n; i += 256) {
        float gr = 0.0f;
        if (tgt >= 0) gr = (expf(zl[(base + i) * kC] - mx) / se - (i == tgt ? 1.0f : 0.0f)) * scale;
        uint16_t *d = dz + (base + i) * kC;
        d[0] = to_bf16(gr);
        for (int c = 1; c < kC; c++) d[c] = 0;
    }
}

__global__ __launch_bounds__(256) void cnn_adam_kernel(float *__restrict__ p,
                                                       const float *__restrict__ gr,
                                                       float *__restrict__ m, float *__restrict__ v,
                                                       int n, float lr, float b1, float b2, float eps,
                                                       float c1, float c2) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = gr[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ __launch_bounds__(256) void cnn_to_bf16_kernel(const float *__restrict__ src, int n,
                                                          uint16_t *__restrict__ dst) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = to_bf16(src[i]);
}

// ---- search pipeline ---------------------------------------------------------------

// dst[y][x] = rounded mean of the 4x4 block at (4x, 4y) (full blocks only).
// A thread makes 4 outputs of one row from 16 bytes of each of 4 source
// rows (one 16-byte load each when `vec`: 16-byte aligned rows).
__global__ __launch_bounds__(256) void downsample4_kernel(const uint8_t *__restrict__ src, int W,
                                                          int dw, int dh, int vec,
                                                          uint8_t *__restrict__ dst) {
    const int x0 = 4 * (blockIdx.x * 256 + threadIdx.x), y = blockIdx.y;
    if (x0 >= dw || y >= dh) return;
    uint32_t s[4] = {0u, 0u, 0u, 0u};
    if (vec && x0 + 4 <= dw) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const uint4 v =
                *reinterpret_cast<const uint4 *>(src + (size_t)(4 * y + r) * W + 4 * x0);
            s[0] = __builtin_amdgcn_udot4(v.x, 0x01010101u, s[0], false);
            s[1] = __builtin_amdgcn_udot4(v.y, 0x01010101u, s[1], false);
            s[2] = __builtin_amdgcn_udot4(v.z, 0x01010101u, s[2], false);
            s[3] = __builtin_amdgcn_udot4(v.w, 0x01010101u, s[3], false);
        }
    } else {
        for (int j = 0; j < 4 && x0 + j < dw; j++)
            for (int r = 0; r < 4; r++) {
                const uint8_t *row = src + (size_t)(4 * y + r) * W + 4 * (x0 + j);
                s[j] += row[0] + row[1] + row[2] + row[3];
            }
    }
    uint8_t *o = dst + (size_t)y * dw + x0;
    if (vec && x0 + 4 <= dw && (dw & 3) == 0) {
        *reinterpret_cast<uint32_t *>(o) = ((s[0] + 8) >> 4) | (((s[1] + 8) >> 4) << 8) |
                                           (((s[2] + 8) >> 4) << 16) | (((s[3] + 8) >> 4) << 24);
    } else {
        for (int j = 0; j < 4 && x0 + j < dw; j++) o[j] = (uint8_t)((s[j] + 8) >> 4);
    }
}

// Order of candidate entries (value, index): larger value first; NaN after
// every number; equal values: lower index. entry_key maps an entry to a
// 64-bit key that is larger exactly when the entry comes first (value bits
// made monotone, -0 taken as +0, NaN lowest; the complemented index below);
// key 0 is an empty slot, after everything.
__device__ __forceinline__ uint64_t entry_key(float v, int idx) {
    const uint32_t u = __float_as_uint(v + 0.0f);
    const uint32_t k = (v != v) ? 0u : ((u & 0x80000000u) ? ~u : (u | 0x80000000u));
    return ((uint64_t)k << 32) | (uint32_t)(0xffffffffu - (uint32_t)idx);
}
__device__ __forceinline__ int key_index(uint64_t key) {
    return (int)(0xffffffffu - (uint32_t)key);
}
// The value of a key made from a number.
__device__ __forceinline__ float key_value(uint64_t key) {
    const uint32_t k = (uint32_t)(key >> 32);
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}
__device__ __forceinline__ uint64_t max_key(uint64_t a, uint64_t b) { return a > b ? a : b; }
__device__ __forceinline__ uint64_t shfl_xor_key(uint64_t v, int o) {
    const uint32_t lo = __shfl_xor((uint32_t)v, o), hi = __shfl_xor((uint32_t)(v >> 32), o);
    return ((uint64_t)hi << 32) | lo;
}

// Best entry key of each 16x16 tile of the coarse map. A workgroup covers a
// band of 16 map rows by 256 columns (16 tiles): each lane folds one column
// of the band, then each group of 16 lanes folds into its tile.
__global__ __launch_bounds__(256) void tile_max_kernel(const float *__restrict__ map, int cw, int ch,
                                                       int tiles_x, uint64_t *__restrict__ tkey) {
    const int x = blockIdx.x * 256 + threadIdx.x, ty = blockIdx.y;
    uint64_t best = 0;
    if (x < cw) {
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int y = 16 * ty + r;
            if (y < ch) best = max_key(best, entry_key(map[(size_t)y * cw + x], y * cw + x));
        }
    }
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) best = max_key(best, shfl_xor_key(best, o));
    const int tx = x >> 4;
    if ((threadIdx.x & 15) == 0 && tx < tiles_x) tkey[ty * tiles_x + tx] = best;
}

constexpr int kChunk = 1024;
constexpr int kMaxCand = 256;

// Lane l's value of lane l ^ S: DPP moves inside a row of 16 lanes for
// S = 1, 2, 4 (row half mirror = xor 7, then quad reverse = xor 3) and 8
// (row rotate by 8), a shuffle otherwise.
template <int S> __device__ __forceinline__ uint32_t xor_lane(uint32_t v) {
    if constexpr (S == 1) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);
    else if constexpr (S == 2) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);
    else if constexpr (S == 4) {
        const int t = __builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, false);
        return (uint32_t)__builtin_amdgcn_update_dpp(0, t, 0x1B, 0xF, 0xF, false);
    } else if constexpr (S == 8) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xF, 0xF, false);
    else return __shfl_xor(v, S);
}
template <int S> __device__ __forceinline__ uint64_t xor_lane_key(uint64_t v) {
    return ((uint64_t)xor_lane<S>((uint32_t)(v >> 32)) << 32) | xor_lane<S>((uint32_t)v);
}

// One bitonic stage of sort1024: pair (tid, tid ^ STRIDE) inside the
// size-SIZE blocks.
template <int SIZE, int STRIDE>
__device__ __forceinline__ uint64_t bitonic_stage(uint64_t key, uint64_t *xs, int tid, int &flip) {
    uint64_t o;
    if constexpr (STRIDE >= 64) {
        uint64_t *buf = xs + flip * kChunk;
        flip ^= 1;
        buf[tid] = key;
        __syncthreads();
        o = buf[tid ^ STRIDE];
    } else {
        o = xor_lane_key<STRIDE>(key);
    }
    const bool keep_max = ((tid & STRIDE) == 0) == ((tid & SIZE) == 0);
    return keep_max ? max_key(key, o) : (key < o ? key : o);
}
template <int SIZE, int STRIDE>
__device__ __forceinline__ uint64_t bitonic_merge(uint64_t key, uint64_t *xs, int tid, int &flip) {
    key = bitonic_stage<SIZE, STRIDE>(key, xs, tid, flip);
    if constexpr (STRIDE > 1) key = bitonic_merge<SIZE, STRIDE / 2>(key, xs, tid, flip);
    return key;
}
template <int SIZE>
__device__ __forceinline__ uint64_t bitonic_sort(uint64_t key, uint64_t *xs, int tid, int &flip) {
    key = bitonic_merge<SIZE, SIZE / 2>(key, xs, tid, flip);
    if constexpr (SIZE < kChunk) key = bitonic_sort<SIZE * 2>(key, xs, tid, flip);
    return key;
}

// Sorts the kChunk = 1024 keys of a 1024-thread workgroup (thread t holds
// position t), largest first, with a bitonic network: pairs less than 64
// positions apart are exchanged between the lanes of a wave (DPP moves or
// shuffles), farther ones through LDS (xs: 2 x 1024 keys, alternating so
// one barrier per such stage suffices).
__device__ __forceinline__ uint64_t sort1024(uint64_t key, uint64_t *xs, int tid) {
    int flip = 0;
    return bitonic_sort<2>(key, xs, tid, flip);
}

// Chunk-wise top-k of n keys: workgroup b sorts keys [b * kChunk, ...) and
// keeps its first `keep` at out[b * keep] or, with cand set (a single
// chunk), writes them as the candidate list cand = (index % cw,
// index / cw), cscore = map[index], with (0, 0) and -inf for empty slots.
__global__ __launch_bounds__(kChunk) void topk_chunk_kernel(const uint64_t *__restrict__ in, int n,
                                                            int keep, uint64_t *__restrict__ out,
                                                            const float *__restrict__ map, int cw,
                                                            int *__restrict__ cand,
                                                            float *__restrict__ cscore) {
    __shared__ uint64_t xs[2 * kChunk];
    const int tid = threadIdx.x, g = blockIdx.x * kChunk + tid;
    const uint64_t key = sort1024(g < n ? in[g] : 0ull, xs, tid);
    if (tid >= keep) return;
    if (cand) {
        const int id = key_index(key);
        cand[2 * tid] = key ? id % cw : 0;
        cand[2 * tid + 1] = key ? id / cw : 0;
        cscore[tid] = key ? map[id] : -INFINITY;
    } else {
        out[(size_t)blockIdx.x * keep + tid] = key;
    }
}

// The k best of t keys in one workgroup, when t <= 1024 * kTopKeysPerThread
// and every thread's share of at most ceil(t / 1024) keys keeps the
// candidates below kChunk: the k-th
// largest of the 1024 per-thread maxima (keys are distinct) bounds the
// answer from below, only the k threads reaching it hold keys that large,
// so at most k * ceil(t / 1024) <= kChunk keys pass; each finds its rank
// by counting the passing keys above it (typically a few more than k).
constexpr int kTopKeysPerThread = 16;
__global__ __launch_bounds__(kChunk) void topk_single_kernel(const uint64_t *__restrict__ keys,
                                                             int t, int k,
                                                             const float *__restrict__ map,
                                                             int cw, int *__restrict__ cand,
                                                             float *__restrict__ cscore) {
    __shared__ uint64_t xs[2 * kChunk];
    __shared__ uint64_t thr;
    __shared__ int cnt;
    const int tid = threadIdx.x;
    uint64_t kr[kTopKeysPerThread];  // this thread's keys, all loads in flight at once
#pragma unroll
    for (int q = 0; q < kTopKeysPerThread; q++) {
        const int i = tid + q * kChunk;
        kr[q] = i < t ? keys[i] : 0ull;
    }
    uint64_t mx = 0;
#pragma unroll
    for (int q = 0; q < kTopKeysPerThread; q++) mx = max_key(mx, kr[q]);
    if (tid == 0) {
        thr = 0ull;
        cnt = 0;
    }
    const uint64_t sorted = sort1024(mx, xs, tid);
    __syncthreads();
    if (tid == k - 1) thr = sorted;
    __syncthreads();
    const uint64_t T = thr;
#pragma unroll
    for (int q = 0; q < kTopKeysPerThread; q++) {
        const uint64_t key = kr[q];
        if (key != 0ull && key >= T) {
            const int slot = atomicAdd(&cnt, 1);
            if (slot < kChunk) xs[slot] = key;
        }
    }
    __syncthreads();
    const int nc = min(cnt, kChunk);
    if (tid < nc) {
        const uint64_t key = xs[tid];
        int rank = 0;
#pragma unroll 8
        for (int j = 0; j < nc; j++) rank += xs[j] > key;
        if (rank < k) {
            const int id = key_index(key);
            cand[2 * rank] = id % cw;
            cand[2 * rank + 1] = id / cw;
            cscore[rank] = map[id];
        }
    }
    for (int r = nc + tid; r < k; r += kChunk) {  // fewer keys than k: empty slots
        cand[2 * r] = 0;
        cand[2 * r + 1] = 0;
        cscore[r] = -INFINITY;
    }
}

// The coarse map and its tile keys in one pass, for downsampled templates
// of at most kCoarseMax pixels a side (TWQ = words per template row). A
// workgroup covers 16 coarse rows by 256 columns (16 tiles); a thread one
// column of them. The workgroup downsamples the full template itself
// (rounded 4x4 means, as downsample4_kernel), stages its part of the small
// image in LDS as 4-byte words, and sums exactly: sum(i*t) by v_dot4
// against the template rows (image words realigned to the column with
// v_alignbyte), sum(i^2) from each staged row's v_dot4 with itself, both
// reused by every output row the image row falls in. Scores come from
// finish_score, as the full map's do. TR >= sth template rows are summed
// (rows past sth are zero words).
constexpr int kCoarseMax = 16;
constexpr int kCoarseW = 68;  // staged words per image row: 256 + 15 columns, rounded up
template <int METHOD, int TWQ, int TR>
__global__ __launch_bounds__(256) void coarse_tiles_kernel(
    const uint8_t *__restrict__ small, int sw, int sh, const uint8_t *__restrict__ tpl, int tw,
    int th, int cw, int ch, int tiles_x, int vec, float *__restrict__ coarse,
    uint64_t *__restrict__ tkey) {
    __shared__ uint8_t tb[kCoarseMax * kCoarseMax];
    __shared__ uint32_t ts[kCoarseMax * 4];
    __shared__ uint32_t iw[(16 + TR - 1) * kCoarseW];
    // per-thread prefix sums of row energies, rows 1 .. 16 + TR - 1
    // (row 0 is zero and never read)
    __shared__ uint32_t pre[(16 + TR - 1) * 256];
    __shared__ long long red[4];
    const int tid = threadIdx.x, x0 = blockIdx.x * 256, y0 = blockIdx.y * 16;
    const int stw = tw >> 2, sth = th >> 2, R = 16 + sth - 1;
    long long tsq = 0;
    for (int i = tid; i < sth * stw; i += 256) {
        const int r = i / stw, q = i - r * stw;
        uint32_t sum = 0;
#pragma unroll
        for (int a = 0; a < 4; a++) {
            const uint8_t *row = tpl + (size_t)(4 * r + a) * tw + 4 * q;
            sum += row[0] + row[1] + row[2] + row[3];
        }
        const uint32_t v = (sum + 8) >> 4;
        tb[r * kCoarseMax + q] = (uint8_t)v;
        tsq += v * v;
    }
    {  // every load of the window in flight at once, then the LDS stores
        constexpr int kWords = (16 + TR - 1) * kCoarseW;
        constexpr int kPer = (kWords + 255) / 256;
        uint32_t wv[kPer];
#pragma unroll
        for (int u = 0; u < kPer; u++) {
            const int i = tid + 256 * u;
            const int r = i / kCoarseW, q = i - r * kCoarseW;
            const int gy = y0 + r, gx = x0 + 4 * q;
            wv[u] = 0;
            if (i < kWords && r < R && gy < sh) {
                const uint8_t *row = small + (size_t)gy * sw;
                if (vec && gx + 4 <= sw) {
                    wv[u] = *reinterpret_cast<const uint32_t *>(row + gx);
                } else {
#pragma unroll
                    for (int b = 0; b < 4; b++)
                        if (gx + b < sw) wv[u] |= (uint32_t)row[gx + b] << (8 * b);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kPer; u++)
            if (tid + 256 * u < kWords) iw[tid + 256 * u] = wv[u];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) tsq += __shfl_xor(tsq, o);
    if ((tid & 63) == 0) red[tid >> 6] = tsq;
    __syncthreads();
    if (tid < kCoarseMax * 4) {
        const int r = tid >> 2, q = tid & 3;
        uint32_t wv = 0;
        if (r < sth)
#pragma unroll
            for (int b = 0; b < 4; b++)
                if (4 * q + b < stw) wv |= (uint32_t)tb[r * kCoarseMax + 4 * q + b] << (8 * b);
        ts[tid] = wv;
    }
    __syncthreads();
    const long long stt = red[0] + red[1] + red[2] + red[3];
    // the template rows, the same for every lane: scalar registers (rows past
    // sth are zero words, so they add nothing to sum(i*t))
    uint32_t tsw[TR][TWQ];
#pragma unroll
    for (int r = 0; r < TR; r++)
#pragma unroll
        for (int q = 0; q < TWQ; q++) tsw[r][q] = __builtin_amdgcn_readfirstlane(ts[r * 4 + q]);
    const int j0 = tid >> 2;
    const uint32_t shift = (uint32_t)(tid & 3);
    const uint32_t last = (stw & 3) ? (1u << (8 * (stw & 3))) - 1u : 0xffffffffu;
    // sum(i * t) per output row: at most 256 products below 2^16 each, so
    // a u32 accumulator is exact
    uint32_t sit[16], sii[16];
#pragma unroll
    for (int j = 0; j < 16; j++) sit[j] = 0u;
    // every (image row, template row) pair, without branches: staged rows
    // past R are zero. sum(i^2) of output row j is pre[j + sth] - pre[j],
    // pre[rr] the sum of the row energies above image row rr (stored at
    // row rr - 1).
    uint32_t run = 0;
#pragma unroll
    for (int rr = 0; rr < 16 + TR - 1; rr++) {
        const uint32_t *row = iw + rr * kCoarseW + j0;
        uint32_t w4[TWQ];
        uint32_t lo = row[0];
#pragma unroll
        for (int q = 0; q < TWQ; q++) {
            const uint32_t hi = row[q + 1];
            w4[q] = __builtin_amdgcn_alignbyte(hi, lo, shift);
            lo = hi;
        }
        w4[TWQ - 1] &= last;
        uint32_t e = 0;
#pragma unroll
        for (int q = 0; q < TWQ; q++) e = __builtin_amdgcn_udot4(w4[q], w4[q], e, false);
        run += e;
        pre[rr * 256 + tid] = run;
        if (rr < 16) sii[rr] = run - e;  // pre[rr]
        // one accumulator per output row: up to 16 independent chains
#pragma unroll
        for (int q = 0; q < TWQ; q++)
#pragma unroll
            for (int r = 0; r < TR; r++) {
                const int j = rr - r;
                if (j >= 0 && j < 16)
                    sit[j] = __builtin_amdgcn_udot4(w4[q], tsw[r][q], sit[j], false);
            }
    }
#pragma unroll
    for (int j = 0; j < 16; j++) sii[j] = pre[(j + sth - 1) * 256 + tid] - sii[j];
    const int X = x0 + tid;
    uint64_t best = 0;
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const int Y = y0 + j;
        if (X < cw && Y < ch) {
            const long long c = sit[j];
            // windows of at most kCoarseMax^2 < 64^2 pixels: finish_score's f32
            // branch, chosen at compile time (n = 0 selects the same branch)
            static_assert(kCoarseMax * kCoarseMax < 64 * 64, "coarse windows stay below 64x64");
            const float v = finish_score<METHOD>(c, (long long)sii[j], stt, 0);
            coarse[(size_t)Y * cw + X] = v;
            best = max_key(best, entry_key(v, Y * cw + X));
        }
    }
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) best = max_key(best, shfl_xor_key(best, o));
    if ((tid & 15) == 0 && (X >> 4) < tiles_x) tkey[blockIdx.y * tiles_x + (X >> 4)] = best;
}

// Network input channels 0-3 of pixel pix of the patch around candidate
// (cx, cy), packed as bf16 pairs: 0 the coarse score minus the candidate's,
// x100; 1 the candidate's score minus the best candidate's, x100 (the same
// everywhere); 2 the small image's pixel at the window centre, / 255 - 0.5;
// 3 one inside the coarse map. Outside the map only channel 1 is set.
__device__ __forceinline__ uint2 feature_pixel(const float *__restrict__ map, int cw, int ch,
                                               const uint8_t *__restrict__ small, int sw, int sh,
                                               int half_t, int cx, int cy, float peak, float best,
                                               int pix) {
    const int x = cx - 8 + (pix & 15), y = cy - 8 + (pix >> 4);
    float f[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (x >= 0 && x < cw && y >= 0 && y < ch) {
        const float v = map[(size_t)y * cw + x];
        f[0] = (v == v && peak == peak) ? (v - peak) * 100.0f : -100.0f;
        const int sx = min(x + half_t, sw - 1), sy = min(y + half_t, sh - 1);
        f[2] = small[(size_t)sy * sw + sx] * (1.0f / 255.0f) - 0.5f;
        f[3] = 1.0f;
    }
    f[1] = (peak == peak && best == best) ? (peak - best) * 100.0f : -100.0f;
    return make_uint2((uint32_t)to_bf16(f[0]) | ((uint32_t)to_bf16(f[1]) << 16),
                      (uint32_t)to_bf16(f[2]) | ((uint32_t)to_bf16(f[3]) << 16));
}

// Feature patch of candidate c (one workgroup), NHWC bf16, channels 4-31 zero.
__global__ __launch_bounds__(256) void features_kernel(const float *__restrict__ map, int cw, int ch,
                                                       const uint8_t *__restrict__ small, int sw,
                                                       int sh, int half_t,
                                                       const int *__restrict__ cand,
                                                       const float *__restrict__ cscore,
                                                       uint16_t *__restrict__ patches) {
    const int c = blockIdx.x, pix = threadIdx.x;
    const uint2 f = feature_pixel(map, cw, ch, small, sw, sh, half_t, cand[2 * c],
                                  cand[2 * c + 1], cscore[c], cscore[0], pix);
    uint16_t *o = patches + ((size_t)c * kPix + pix) * kC;
    reinterpret_cast<uint4 *>(o)[0] = make_uint4(f.x, f.y, 0u, 0u);
    for (int q = 1; q < 4; q++) reinterpret_cast<uint4 *>(o)[q] = make_uint4(0u, 0u, 0u, 0u);
}

// The whole network on candidate c's patch (one workgroup): features built
// in LDS, every layer's activations kept in LDS, logits[c * 256 + pixel] =
// the last layer's channel 0. The same arithmetic as features_kernel and one
// cnn_conv3x3_kernel per layer (ReLU between layers), so the logits are
// identical to theirs. w: `layers` weight blocks of 32 x 288 bf16, bias:
// `layers` x 32 f32.
constexpr int kFwdThreads = 512;  // 8 waves: one 32-pixel tile each
__global__ __launch_bounds__(kFwdThreads) void cnn_proposal_forward_kernel(
    const float *__restrict__ map, int cw, int ch, const uint8_t *__restrict__ small, int sw,
    int sh, int half_t, const int *__restrict__ cand, const float *__restrict__ cscore,
    const uint16_t *__restrict__ w, const float *__restrict__ bias, int layers,
    float *__restrict__ logits) {
    __shared__ __align__(16) uint16_t xs[kXP * kXP * kXPitch];
    __shared__ __align__(16) uint16_t ws[kC * kWPitch];
    constexpr int kWq = kC * (kKK / 8);  // uint4 per layer's weights
    constexpr int kPre = (kWq + kFwdThreads - 1) / kFwdThreads;
    const int c = blockIdx.x, tid = threadIdx.x;
    for (int i = tid; i < kWq; i += kFwdThreads) {
        const int r = i / (kKK / 8), c8 = i - r * (kKK / 8);
        *reinterpret_cast<uint4 *>(ws + r * kWPitch + 8 * c8) = reinterpret_cast<const uint4 *>(w)[i];
    }
    for (int i = tid; i < kXP * kXP * kXPitch / 8; i += kFwdThreads)
        reinterpret_cast<uint4 *>(xs)[i] = make_uint4(0u, 0u, 0u, 0u);
    __syncthreads();
    if (tid < kPix) {
        const uint2 f = feature_pixel(map, cw, ch, small, sw, sh, half_t, cand[2 * c],
                                      cand[2 * c + 1], cscore[c], cscore[0], tid);
        const int px = tid & 15, py = tid >> 4;
        *reinterpret_cast<uint2 *>(xs + ((py + 1) * kXP + px + 1) * kXPitch) = f;
    }
    __syncthreads();
    const int lane = tid & 63, mt = tid >> 6, n = lane & 31, g = lane >> 5;
    for (int l = 0; l < layers; l++) {
        const float b = bias[l * kC + n];
        // next layer's weights, in flight during the MFMAs
        static_assert(kPre == 3, "three uint4 of weights per thread");
        uint4 pre0 = {}, pre1 = {}, pre2 = {};
        if (l + 1 < layers) {
            const uint4 *src = reinterpret_cast<const uint4 *>(w + (size_t)(l + 1) * kC * kKK);
            pre0 = src[tid];
            pre1 = src[tid + kFwdThreads];
            pre2 = src[min(tid + 2 * kFwdThreads, kWq - 1)];
        }
        const v16f acc = conv_tile(xs, ws, mt, n, g);
        __syncthreads();  // every wave has read xs and ws
        if (l + 1 < layers) {
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int pix = tile_pixel(mt, g, r);
                xs[(((pix >> 4) + 1) * kXP + (pix & 15) + 1) * kXPitch + n] =
                    to_bf16(fmaxf(acc[r] + b, 0.0f));
            }
#pragma unroll
            for (int q = 0; q < kPre; q++) {
                const int i = tid + kFwdThreads * q;
                if (i < kWq) {
                    const int r = i / (kKK / 8), c8 = i - r * (kKK / 8);
                    *reinterpret_cast<uint4 *>(ws + r * kWPitch + 8 * c8) =
                        q == 0 ? pre0 : (q == 1 ? pre1 : pre2);
                }
            }
        } else if (n == 0) {
#pragma unroll
            for (int r = 0; r < 16; r++)
                logits[(size_t)c * kPix + tile_pixel(mt, g, r)] = acc[r] + b;
        }
        __syncthreads();
    }
}

__device__ __forceinline__ float block_max256(float v, float *red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}

__device__ __forceinline__ float block_sum256(float v, float *red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return (red[0] + red[1]) + (red[2] + red[3]);
}

// Per-patch half of the choice of the best logits. Entry (c, pix) has the
// coarse place cand[c] - 8 + (pix % 16, pix / 16) and index c * 256 + pix;
// it is selectable when its place is inside the map, its logit is a number,
// and no other patch holds an entry for the same place that comes first
// (entry_key order). Writes patch c's n_top best selectable entry keys, best
// first, to top[c * n_top + r] (0 past the last), and its softmax partials
// part[c] = (max logit, sum of exp(logit - max)) over all 256 logits.
// Logit i is lg[i * stride].
__global__ __launch_bounds__(256) void logit_patch_kernel(const float *__restrict__ lg, int stride,
                                                          int k, const int *__restrict__ cand,
                                                          int cw, int ch, int n_top,
                                                          uint64_t *__restrict__ top,
                                                          float2 *__restrict__ part) {
    __shared__ int cxy[2 * kMaxCand];
    __shared__ uint64_t ent[kPix];
    __shared__ float red[4];
    const int c = blockIdx.x, pix = threadIdx.x;
    for (int i = pix; i < 2 * k; i += 256) cxy[i] = cand[i];
    __syncthreads();
    const int me = c * kPix + pix;
    const float v = lg[(size_t)me * stride];
    const uint64_t mine = entry_key(v, me);
    const int x = cxy[2 * c] - 8 + (pix & 15), y = cxy[2 * c + 1] - 8 + (pix >> 4);
    bool sel = x >= 0 && x < cw && y >= 0 && y < ch && v == v;
    if (sel) {
#pragma unroll 8
        for (int c2 = 0; c2 < k; c2++) {
            const int px = x - (cxy[2 * c2] - 8), py = y - (cxy[2 * c2 + 1] - 8);
            if (c2 != c && (unsigned)px < 16u && (unsigned)py < 16u) {
                const int o = c2 * kPix + py * kPS + px;
                if (entry_key(lg[(size_t)o * stride], o) > mine) sel = false;
            }
        }
    }
    ent[pix] = sel ? mine : 0ull;
    const int nsel = __syncthreads_count(sel);
    if (sel) {
        int rank = 0;
#pragma unroll 8
        for (int j = 0; j < kPix; j++) rank += ent[j] > mine;
        if (rank < n_top) top[(size_t)c * n_top + rank] = mine;
    }
    for (int r = pix; r < n_top; r += 256)
        if (r >= nsel) top[(size_t)c * n_top + r] = 0ull;
    const float mx = block_max256(fmaxf(-INFINITY, v), red);
    const float se = block_sum256(expf(v - mx), red);
    if (pix == 0) part[c] = make_float2(mx, se);
}

// Second half: the n_top best keys of all patches' lists (no place is
// selectable twice). The n_top-th best patch head bounds them from below,
// so only keys at least as large are ranked. centers[r] = coarse place of
// the r-th best ((0, 0) past the last selectable entry); probs[r] =
// exp(logit - max) / sum(exp(logit - max)) over all k * 256 logits (0 past
// the last). One workgroup of 256 threads.
constexpr int kMergeCap = 2048;
__global__ __launch_bounds__(256) void logit_merge_kernel(const uint64_t *__restrict__ top,
                                                          const float2 *__restrict__ part, int k,
                                                          const int *__restrict__ cand, int n_top,
                                                          int *__restrict__ centers,
                                                          float *__restrict__ probs) {
    __shared__ uint64_t head[kMaxCand];
    __shared__ uint64_t sur[kMergeCap];
    __shared__ uint64_t thr;
    __shared__ int cnt;
    __shared__ float red[4];
    const int tid = threadIdx.x, m = k * n_top;
    const float2 pc = tid < k ? part[tid] : make_float2(-INFINITY, 0.0f);
    const float mx = block_max256(pc.x, red);
    const float se = block_sum256(pc.x != -INFINITY ? pc.y * expf(pc.x - mx) : 0.0f, red);
    if (tid == 0) {
        thr = 0ull;
        cnt = 0;
    }
    if (tid < k) head[tid] = top[(size_t)tid * n_top];
    __syncthreads();
    if (tid < k && head[tid] != 0ull) {
        const uint64_t h = head[tid];
        int rank = 0;
#pragma unroll 8
        for (int j = 0; j < k; j++) rank += head[j] > h;
        if (rank == n_top - 1) thr = h;
    }
    __syncthreads();
    const uint64_t T = thr;
    for (int i = tid; i < m; i += 256) {
        const uint64_t key = top[i];
        if (key == 0ull || key < T) continue;
        const int s = atomicAdd(&cnt, 1);
        if (s < kMergeCap) sur[s] = key;
    }
    __syncthreads();
    const int ns = cnt;
    auto emit = [&](int rank, uint64_t key) {
        const int id = key_index(key), c = id / kPix, pix = id - c * kPix;
        centers[2 * rank] = cand[2 * c] - 8 + (pix & 15);
        centers[2 * rank + 1] = cand[2 * c + 1] - 8 + (pix >> 4);
        probs[rank] = expf(key_value(key) - mx) / se;
    };
    if (ns <= kMergeCap) {
        for (int s = tid; s < ns; s += 256) {
            const uint64_t e = sur[s];
            int rank = 0;
#pragma unroll 8
            for (int j = 0; j < ns; j++) rank += sur[j] > e;
            if (rank < n_top) emit(rank, e);
        }
    } else {  // many entries tie with the bound: rank every selectable key against all
        for (int i = tid; i < m; i += 256) {
            const uint64_t e = top[i];
            if (e == 0ull) continue;
            int rank = 0;
            for (int j = 0; j < m && rank < n_top; j++) rank += top[j] > e;
            if (rank < n_top) emit(rank, e);
        }
    }
    for (int r = tid; r < n_top; r += 256)
        if (r >= ns) {
            centers[2 * r] = 0;
            centers[2 * r + 1] = 0;
            probs[r] = 0.0f;
        }
}

// Exact scores of the 17 places of row dy = blockIdx.y of the window around
// centre c = blockIdx.x: scores[c][dy * 17 + dx], NaN where the place is
// outside [0, W - tw] x [0, H - th]. Thread (rs, dx) sums template rows
// rs, rs + 15, ... of place dx with v_dot4 over 4-byte groups (image words
// realigned to the place with v_alignbyte); the workgroup adds the 15
// partial sums and normalises them as the full map does.
constexpr int kRefR = 8;
constexpr int kRefS = 2 * kRefR + 1;
constexpr int kRefN = kRefS * kRefS;
constexpr int kRefSub = 15;
constexpr int kRefMaxSide = 128;
__host__ __device__ constexpr int ref_tpl_words(int tw) { return (tw + 3) >> 2; }
__host__ __device__ constexpr int ref_img_words(int tw) { return ref_tpl_words(tw) + 5; }
// (score, y, x) order of find_extremes' maximum: larger score, then smaller
// y, then smaller x; NaN never wins.
__device__ __forceinline__ bool ref_better(float a, int ay, int ax, float b, int by, int bx) {
    if (a != a) return false;
    if (b != b) return true;
    if (a != b) return a > b;
    return ay != by ? ay < by : ax < bx;
}

// Best place over every refined window and the best score of the places
// more than `sep` pixels (Chebyshev distance) from it. result: x, y, best,
// second; with p_out, *p_out = *p_top (1 without it). One workgroup of up
// to 1024 threads.
__global__ __launch_bounds__(1024) void refine_best_kernel(const float *__restrict__ scores,
                                                           const int *__restrict__ centers,
                                                           int n_c, int scale, int sep,
                                                           const float *__restrict__ p_top,
                                                           float *__restrict__ result,
                                                           float *__restrict__ p_out) {
    __shared__ float sv[16];
    __shared__ int sy[16], sx[16];
    const int n = n_c * kRefN, tid = threadIdx.x, nt = blockDim.x, lane = tid & 63,
              wave = tid >> 6, nw = nt >> 6;
    auto place = [&](int i, int &x, int &y) {
        const int c = i / kRefN, pos = i - c * kRefN, dy = pos / kRefS;
        x = centers[2 * c] * scale - kRefR + (pos - dy * kRefS);
        y = centers[2 * c + 1] * scale - kRefR + dy;
    };
    // each thread's first kRefPer scores stay in registers (one read of
    // them); scores past nt * kRefPer are read again for the second pass
    constexpr int kRefPer = 8;
    float sv_[kRefPer];
    int sx_[kRefPer], sy_[kRefPer];
    float bv = __int_as_float(0x7fc00000);
    int bx = 0, by = 0;
#pragma unroll
    for (int q = 0; q < kRefPer; q++) {
        const int i = tid + q * nt;
        sv_[q] = i < n ? scores[i] : __int_as_float(0x7fc00000);
        sx_[q] = sy_[q] = 0;
        if (i < n) place(i, sx_[q], sy_[q]);
    }
#pragma unroll
    for (int q = 0; q < kRefPer; q++) {
        const float v = sv_[q];
        if (v == v && ref_better(v, sy_[q], sx_[q], bv, by, bx)) {
            bv = v;
            bx = sx_[q];
            by = sy_[q];
        }
    }
    for (int i = tid + kRefPer * nt; i < n; i += nt) {
        const float v = scores[i];
        if (v != v) continue;
        int x, y;
        place(i, x, y);
        if (ref_better(v, y, x, bv, by, bx)) {
            bv = v;
            bx = x;
            by = y;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const float v = __shfl_xor(bv, o);
        const int x = __shfl_xor(bx, o), y = __shfl_xor(by, o);
        if (ref_better(v, y, x, bv, by, bx)) {
            bv = v;
            bx = x;
            by = y;
        }
    }
    if (lane == 0) {
        sv[wave] = bv;
        sx[wave] = bx;
        sy[wave] = by;
    }
    __syncthreads();
    bv = sv[0];
    bx = sx[0];
    by = sy[0];
    for (int w = 1; w < nw; w++)
        if (ref_better(sv[w], sy[w], sx[w], bv, by, bx)) {
            bv = sv[w];
            bx = sx[w];
            by = sy[w];
        }
    float sec = -INFINITY;
#pragma unroll
    for (int q = 0; q < kRefPer; q++)
        if (sv_[q] == sv_[q] && (abs(sx_[q] - bx) > sep || abs(sy_[q] - by) > sep))
            sec = fmaxf(sec, sv_[q]);
    for (int i = tid + kRefPer * nt; i < n; i += nt) {
        const float v = scores[i];
        if (v != v) continue;
        int x, y;
        place(i, x, y);
        if (abs(x - bx) > sep || abs(y - by) > sep) sec = fmaxf(sec, v);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) sec = fmaxf(sec, __shfl_xor(sec, o));
    __syncthreads();
    if (lane == 0) sv[wave] = sec;
    __syncthreads();
    if (tid == 0) {
        for (int w = 0; w < nw; w++) sec = fmaxf(sec, sv[w]);
        result[0] = (float)bx;
        result[1] = (float)by;
        result[2] = bv;
        result[3] = sec;
        if (p_out) *p_out = p_top ? *p_top : 1.0f;
    }
}

template <int METHOD>
__global__ __launch_bounds__(256) void refine_kernel(const uint8_t *__restrict__ img, int W, int H,
                                                     const uint8_t *__restrict__ tpl, int tw,
                                                     int th, const int *__restrict__ centers,
                                                     int scale, float *__restrict__ scores) {
    extern __shared__ __align__(16) uint32_t sm[];
    __shared__ int2 part[kRefSub * kRefS];
    __shared__ long long red[4];
    const int twq = ref_tpl_words(tw), rq = ref_img_words(tw);
    uint32_t *ts = sm;              // th x twq template words
    uint32_t *is = sm + th * twq;   // th x rq image words from column ox
    const int c = blockIdx.x, dy = blockIdx.y, tid = threadIdx.x;
    const int ox = centers[2 * c] * scale - kRefR, oy = centers[2 * c + 1] * scale - kRefR + dy;
    // loads issued kRefBatch words at a time before their LDS stores
    constexpr int kRefBatch = 4;
    long long tsq = 0;
    for (int i0 = tid; i0 < th * twq; i0 += 256 * kRefBatch) {
        uint32_t wv[kRefBatch];
#pragma unroll
        for (int u = 0; u < kRefBatch; u++) {
            const int i = i0 + 256 * u, r = i / twq, q = i - r * twq;
            wv[u] = 0;
            if (i < th * twq)
#pragma unroll
                for (int b = 0; b < 4; b++)
                    if (4 * q + b < tw) wv[u] |= (uint32_t)tpl[r * tw + 4 * q + b] << (8 * b);
        }
#pragma unroll
        for (int u = 0; u < kRefBatch; u++)
            if (i0 + 256 * u < th * twq) {
                ts[i0 + 256 * u] = wv[u];
                tsq += __builtin_amdgcn_udot4(wv[u], wv[u], 0u, false);
            }
    }
    for (int i0 = tid; i0 < th * rq; i0 += 256 * kRefBatch) {
        uint32_t wv[kRefBatch];
#pragma unroll
        for (int u = 0; u < kRefBatch; u++) {
            const int i = i0 + 256 * u, r = i / rq, q = i - r * rq;
            const int gy = oy + r;
            wv[u] = 0;
            if (i < th * rq && gy >= 0 && gy < H) {
                const uint8_t *row = img + (size_t)gy * W;
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    const int gx = ox + 4 * q + b;
                    if (gx >= 0 && gx < W) wv[u] |= (uint32_t)row[gx] << (8 * b);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kRefBatch; u++)
            if (i0 + 256 * u < th * rq) is[i0 + 256 * u] = wv[u];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) tsq += __shfl_xor(tsq, o);
    if ((tid & 63) == 0) red[tid >> 6] = tsq;
    __syncthreads();
    const int rs = tid / kRefS, dx = tid - rs * kRefS;
    if (rs < kRefSub) {
        const int j0 = dx >> 2;
        const uint32_t sh = (uint32_t)(dx & 3);
        const uint32_t last = (tw & 3) ? (1u << (8 * (tw & 3))) - 1u : 0xffffffffu;
        uint32_t sit = 0, sii = 0;
        for (int r = rs; r < th; r += kRefSub) {
            const uint32_t *ir = is + r * rq + j0;
            const uint32_t *tr = ts + r * twq;
            uint32_t lo = ir[0];
            for (int q = 0; q < twq; q++) {
                const uint32_t hi = ir[q + 1];
                uint32_t iw = __builtin_amdgcn_alignbyte(hi, lo, sh);
                if (q == twq - 1) iw &= last;
                sit = __builtin_amdgcn_udot4(iw, tr[q], sit, false);
                sii = __builtin_amdgcn_udot4(iw, iw, sii, false);
                lo = hi;
            }
        }
        part[tid] = make_int2((int)sit, (int)sii);
    }
    __syncthreads();
    if (tid < kRefS) {
        long long sit = 0, sii = 0;
        for (int q = 0; q < kRefSub; q++) {
            sit += (uint32_t)part[q * kRefS + tid].x;
            sii += (uint32_t)part[q * kRefS + tid].y;
        }
        const long long stt = red[0] + red[1] + red[2] + red[3];
        const int x = ox + tid;
        float s = __int_as_float(0x7fc00000);
        if (x >= 0 && x <= W - tw && oy >= 0 && oy <= H - th)
            s = finish_score<METHOD>(sit, sii, stt, (long long)tw * th);
        scores[(size_t)c * kRefN + dy * kRefS + tid] = s;
    }
}

}  // namespace
}  // namespace iph

using namespace iph;

namespace {
inline hipStream_t S(void *s) { return static_cast<hipStream_t>(s); }
inline unsigned blocks(size_t n) { return (unsigned)((n + 255) / 256); }

iph_status launch_downsample(const uint8_t *src, uint32_t w, uint32_t h, uint8_t *dst,
                             hipStream_t s) {
    const uint32_t dw = w / 4, dh = h / 4;
    if (dw == 0 || dh == 0) return IPH_OK;
    if (dh > 65535) {
        set_error("downsampling takes images of up to 262143 rows");
        return IPH_EINVAL;
    }
    const int vec = (w % 16 == 0) && ((uintptr_t)src % 16 == 0) && ((uintptr_t)dst % 4 == 0);
    downsample4_kernel<<<dim3((dw + 1023) / 1024, dh), 256, 0, s>>>(src, (int)w, (int)dw,
                                                                    (int)dh, vec, dst);
    return launch_status("downsample4_kernel");
}

// Candidate search workspace: tile keys and two buffers for the levels of
// chunk-wise top-k.
struct CandPlan {
    int tx, ty, t, nb;
    size_t bytes;
};
CandPlan cand_plan(uint32_t cw, uint32_t ch) {
    CandPlan p;
    p.tx = (int)((cw + 15) / 16);
    p.ty = (int)((ch + 15) / 16);
    p.t = p.tx * p.ty;
    p.nb = (p.t + kChunk - 1) / kChunk;
    WorkspacePlan wp;
    wp.take<uint64_t>(std::max(p.t, 1));
    for (int b = 0; b < 2; b++) wp.take<uint64_t>((size_t)std::max(p.nb, 1) * kMaxCand);
    p.bytes = wp.used;
    return p;
}

iph_status check_candidates(uint32_t cw, uint32_t ch, int k) {
    if (cw == 0 || ch == 0 || (size_t)cw * ch > (size_t)0x7ffffffe || k <= 0 || k > kMaxCand ||
        (ch + 15) / 16 > 65535) {
        set_error("coarse map of %ux%u or %d candidates out of range (1 to %d)", cw, ch, k,
                  kMaxCand);
        return IPH_EINVAL;
    }
    return IPH_OK;
}

// The k-candidate list from t tile keys: one workgroup when the candidate
// bound allows it, else levels of chunk-wise top-k.
iph_status launch_topk(const uint64_t *tk, int t, int k, const float *coarse, uint32_t cw,
                       int *cand, float *cscore, uint64_t *lk0, uint64_t *lk1, hipStream_t s) {
    if (t <= kChunk * kTopKeysPerThread && (long long)k * ((t + kChunk - 1) / kChunk) <= kChunk) {
        topk_single_kernel<<<1, kChunk, 0, s>>>(tk, t, k, coarse, (int)cw, cand, cscore);
        return launch_status("topk_single_kernel");
    }
    uint64_t *lk[2] = {lk0, lk1};
    const uint64_t *src = tk;
    int n = t, b = 0;
    while (n > kChunk) {
        const int nb = (n + kChunk - 1) / kChunk;
        topk_chunk_kernel<<<nb, kChunk, 0, s>>>(src, n, k, lk[b], nullptr, 0, nullptr, nullptr);
        IPH_TRY(launch_status("topk_chunk_kernel"));
        src = lk[b];
        n = nb * k;
        b ^= 1;
    }
    topk_chunk_kernel<<<1, kChunk, 0, s>>>(src, n, k, nullptr, coarse, (int)cw, cand, cscore);
    return launch_status("topk_chunk_kernel");
}

iph_status launch_candidates(const float *coarse, uint32_t cw, uint32_t ch, int k, int *cand,
                             float *cscore, void *ws, size_t ws_bytes, hipStream_t s) {
    IPH_TRY(check_candidates(cw, ch, k));
    const CandPlan p = cand_plan(cw, ch);
    Workspace wsp(ws, ws_bytes);
    uint64_t *tk = wsp.take<uint64_t>(std::max(p.t, 1));
    uint64_t *lk[2];
    for (int b = 0; b < 2; b++) lk[b] = wsp.take<uint64_t>((size_t)std::max(p.nb, 1) * kMaxCand);
    if (!tk || !lk[1]) {
        set_error("workspace too small for the candidate search");
        return IPH_EINVAL;
    }
    tile_max_kernel<<<dim3((cw + 255) / 256, p.ty), 256, 0, s>>>(coarse, (int)cw, (int)ch, p.tx,
                                                                 tk);
    IPH_TRY(launch_status("tile_max_kernel"));
    return launch_topk(tk, p.t, k, coarse, cw, cand, cscore, lk[0], lk[1], s);
}

constexpr int kMaxTop = 64;
size_t top_workspace(int k, int n_top) {
    WorkspacePlan wp;
    wp.take<uint64_t>((size_t)std::max(k, 1) * std::max(n_top, 1));
    wp.take<float2>(std::max(k, 1));
    return wp.used;
}

iph_status launch_top(const float *lg, int stride, int k, const int *cand, uint32_t cw,
                      uint32_t ch, int n_top, int *centers, float *probs, void *ws,
                      size_t ws_bytes, hipStream_t s) {
    if (k <= 0 || k > kMaxCand || n_top <= 0 || n_top > kMaxTop) {
        set_error("1 to %d candidates and 1 to %d top places", kMaxCand, kMaxTop);
        return IPH_EINVAL;
    }
    Workspace wsp(ws, ws_bytes);
    uint64_t *tk = wsp.take<uint64_t>((size_t)k * n_top);
    float2 *part = wsp.take<float2>(k);
    if (!tk || !part) {
        set_error("workspace too small for the top places");
        return IPH_EINVAL;
    }
    logit_patch_kernel<<<k, 256, 0, s>>>(lg, stride, k, cand, (int)cw, (int)ch, n_top, tk, part);
    IPH_TRY(launch_status("logit_patch_kernel"));
    logit_merge_kernel<<<1, 256, 0, s>>>(tk, part, k, cand, n_top, centers, probs);
    return launch_status("logit_merge_kernel");
}

size_t refine_workspace(int n_centers) {
    WorkspacePlan wp;
    wp.take<float>((size_t)std::max(n_centers, 1) * kRefN);
    return wp.used;
}

iph_status launch_refine(const uint8_t *image, uint32_t w, uint32_t h, const uint8_t *tpl,
                         uint32_t tw, uint32_t th, int32_t method, const int32_t *centers,
                         int32_t n_centers, int32_t scale, int32_t sep, const float *p_top,
                         float *result, float *p_out, void *ws, size_t ws_bytes,
                         hipStream_t s) {
    IPH_TRY(check_match(IPH_U8, w, h, tw, th, method, IPH_PATH_AUTO));
    if (tw == 0 || th == 0 || tw > (uint32_t)kRefMaxSide || th > (uint32_t)kRefMaxSide ||
        n_centers <= 0 || n_centers > 1024) {
        set_error("refinement takes templates of 1 to %d pixels a side and 1 to 1024 centres",
                  kRefMaxSide);
        return IPH_EINVAL;
    }
    Workspace wsp(ws, ws_bytes);
    float *scores = wsp.take<float>((size_t)n_centers * kRefN);
    if (!scores) {
        set_error("workspace too small for the refinement");
        return IPH_EINVAL;
    }
    const size_t lds = (size_t)th * (ref_tpl_words((int)tw) + ref_img_words((int)tw)) * 4;
    const dim3 grid(n_centers, kRefS);
#define RF(M)                                                                                \
    refine_kernel<M><<<grid, 256, lds, s>>>(image, (int)w, (int)h, tpl, (int)tw, (int)th,     \
                                            centers, scale, scores)
    switch (method) {
    case IPH_SSE: RF(IPH_SSE); break;
    case IPH_SSE_NORMED: RF(IPH_SSE_NORMED); break;
    case IPH_CCORR: RF(IPH_CCORR); break;
    default: RF(IPH_CCORR_NORMED); break;
    }
#undef RF
    IPH_TRY(launch_status("refine_kernel"));
    refine_best_kernel<<<1, 1024, 0, s>>>(scores, centers, n_centers, scale, sep, p_top, result,
                                          p_out);
    return launch_status("refine_best_kernel");
}

// Sizes of one proposal search: the 4x-downsampled image and template and
// the coarse map between them.
struct SearchPlan {
    uint32_t sw, sh, stw, sth, cw, ch;
};
iph_status search_plan(uint32_t w, uint32_t h, uint32_t tw, uint32_t th, int32_t method,
                       int32_t k, int32_t n_top, SearchPlan &p) {
    if (method != IPH_CCORR && method != IPH_CCORR_NORMED) {
        set_error("the proposal search serves the correlation methods");
        return IPH_EINVAL;
    }
    if (tw < 4 || th < 4 || tw > (uint32_t)kRefMaxSide || th > (uint32_t)kRefMaxSide || w < tw ||
        h < th) {
        set_error("the proposal search takes templates of 4 to %d pixels a side, no larger "
                  "than the image", kRefMaxSide);
        return IPH_EINVAL;
    }
    if (k <= 0 || k > kMaxCand || n_top <= 0 || n_top > kMaxTop || n_top > k) {
        set_error("1 to %d candidates and 1 to min(%d, candidates) exact windows", kMaxCand,
                  kMaxTop);
        return IPH_EINVAL;
    }
    p.sw = w / 4;
    p.sh = h / 4;
    p.stw = tw / 4;
    p.sth = th / 4;
    p.cw = p.sw - p.stw + 1;
    p.ch = p.sh - p.sth + 1;
    return IPH_OK;
}

// The coarse map and tile keys come from one coarse_tiles_kernel pass when
// the downsampled template is small enough, else from the int8 MFMA match
// and tile_max_kernel.
bool coarse_fused(const SearchPlan &p) {
    return p.stw <= (uint32_t)kCoarseMax && p.sth <= (uint32_t)kCoarseMax;
}
}  // namespace

extern "C" {

IPH_API iph_status iph_dev_cnn_conv3x3(const uint16_t *x, const uint16_t *w, const float *bias,
                                       int32_t relu, int32_t n_patches, float *z, uint16_t *a,
                                       void *stream) {
    clear_error();
    if (n_patches <= 0) return IPH_OK;
    cnn_conv3x3_kernel<<<n_patches, 256, 0, S(stream)>>>(x, w, bias, relu, z, a);
    return launch_status("cnn_conv3x3_kernel");
}

IPH_API size_t iph_dev_cnn_conv3x3_dw_workspace(int32_t n_partials) {
    return (size_t)std::max(n_partials, 1) * (kC * kKK + kC) * sizeof(float);
}

IPH_API iph_status iph_dev_cnn_conv3x3_dw(const uint16_t *x, const uint16_t *dz,
                                          int32_t n_patches, int32_t n_partials, void *workspace,
                                          size_t workspace_bytes, float *dw, float *db,
                                          void *stream) {
    clear_error();
    if (n_partials <= 0 || workspace_bytes < iph_dev_cnn_conv3x3_dw_workspace(n_partials)) {
        set_error("workspace too small for %d partial gradients", n_partials);
        return IPH_EINVAL;
    }
    float *part = static_cast<float *>(workspace);
    cnn_conv3x3_dw_kernel<<<n_partials, 256, 0, S(stream)>>>(x, dz, n_patches, part);
    IPH_TRY(launch_status("cnn_conv3x3_dw_kernel"));
    cnn_reduce_partials_kernel<<<blocks(kC * kKK + kC), 256, 0, S(stream)>>>(part, n_partials, dw,
                                                                            db);
    return launch_status("cnn_reduce_partials_kernel");
}

IPH_API iph_status iph_dev_cnn_flip_weights(const uint16_t *w, uint16_t *w_flipped, void *stream) {
    clear_error();
    cnn_flip_kernel<<<blocks(kC * kKK), 256, 0, S(stream)>>>(w, w_flipped);
    return launch_status("cnn_flip_kernel");
}

IPH_API iph_status iph_dev_cnn_relu_backward(const float *dx, const float *z, size_t n,
                                             uint16_t *dz, void *stream) {
    clear_error();
    if (n == 0) return IPH_OK;
    cnn_relu_backward_kernel<<<blocks(n), 256, 0, S(stream)>>>(dx, z, n, dz);
    return launch_status("cnn_relu_backward_kernel");
}

IPH_API iph_status iph_dev_cnn_softmax_xent(const float *z_last, int32_t n_pairs,
                                            int32_t patches_per_pair, const int32_t *target,
                                            float scale, float *loss, uint16_t *dz_last,
                                            void *stream) {
    clear_error();
    if (n_pairs <= 0) return IPH_OK;
    cnn_softmax_xent_kernel<<<n_pairs, 256, 0, S(stream)>>>(z_last, patches_per_pair, target,
                                                            scale, loss, dz_last);
    return launch_status("cnn_softmax_xent_kernel");
}

IPH_API iph_status iph_dev_cnn_adam(float *param, const float *grad, float *m, float *v,
                                    int32_t n, float lr, float beta1, float beta2, float eps,
                                    int32_t step, void *stream) {
    clear_error();
    if (n <= 0) return IPH_OK;
    const float c1 = 1.0f - std::pow(beta1, (float)step), c2 = 1.0f - std::pow(beta2, (float)step);
    cnn_adam_kernel<<<blocks(n), 256, 0, S(stream)>>>(param, grad, m, v, n, lr, beta1, beta2, eps,
                                                      c1, c2);
    return launch_status("cnn_adam_kernel");
}

IPH_API iph_status iph_dev_cnn_to_bf16(const float *src, int32_t n, uint16_t *dst, void *stream) {
    clear_error();
    if (n <= 0) return IPH_OK;
    cnn_to_bf16_kernel<<<blocks(n), 256, 0, S(stream)>>>(src, n, dst);
    return launch_status("cnn_to_bf16_kernel");
}

IPH_API iph_status iph_dev_downsample4_u8(const uint8_t *src, uint32_t w, uint32_t h,
                                          uint8_t *dst, void *stream) {
    clear_error();
    return launch_downsample(src, w, h, dst, S(stream));
}

IPH_API size_t iph_dev_proposal_candidates_workspace(uint32_t cw, uint32_t ch) {
    return cand_plan(cw, ch).bytes;
}

IPH_API iph_status iph_dev_proposal_candidates(const float *coarse, uint32_t cw, uint32_t ch,
                                               int32_t k, int32_t *cand, float *cand_score,
                                               void *workspace, size_t workspace_bytes,
                                               void *stream) {
    clear_error();
    return launch_candidates(coarse, cw, ch, k, cand, cand_score, workspace, workspace_bytes,
                             S(stream));
}

IPH_API iph_status iph_dev_proposal_features(const float *coarse, uint32_t cw, uint32_t ch,
                                             const uint8_t *small, uint32_t sw, uint32_t sh,
                                             uint32_t small_tw, const int32_t *cand,
                                             const float *cand_score, int32_t k,
                                             uint16_t *patches, void *stream) {
    clear_error();
    if (k <= 0) return IPH_OK;
    features_kernel<<<k, 256, 0, S(stream)>>>(coarse, (int)cw, (int)ch, small, (int)sw, (int)sh,
                                              (int)small_tw / 2, cand, cand_score, patches);
    return launch_status("features_kernel");
}

IPH_API size_t iph_dev_proposal_top_workspace(int32_t k, int32_t n_top) {
    return top_workspace(k, n_top);
}

IPH_API iph_status iph_dev_proposal_top(const float *z_last, int32_t k, const int32_t *cand,
                                        uint32_t cw, uint32_t ch, int32_t n_top, int32_t *centers,
                                        float *probs, void *workspace, size_t workspace_bytes,
                                        void *stream) {
    clear_error();
    return launch_top(z_last, kC, k, cand, cw, ch, n_top, centers, probs, workspace,
                      workspace_bytes, S(stream));
}

IPH_API size_t iph_dev_proposal_refine_workspace(int32_t n_centers) {
    return refine_workspace(n_centers);
}

IPH_API iph_status iph_dev_proposal_refine(const uint8_t *image, uint32_t w, uint32_t h,
                                           const uint8_t *tpl, uint32_t tw, uint32_t th,
                                           int32_t method, const int32_t *centers,
                                           int32_t n_centers, int32_t scale, int32_t sep,
                                           float *result, void *workspace, size_t workspace_bytes,
                                           void *stream) {
    clear_error();
    return launch_refine(image, w, h, tpl, tw, th, method, centers, n_centers, scale, sep,
                         nullptr, result, nullptr, workspace, workspace_bytes, S(stream));
}

IPH_API iph_status iph_dev_read_result(const float *result, float *host, size_t n,
                                       void *stream) {
    clear_error();
    IPH_TRY_HIP(hipMemcpyAsync(host, result, n * sizeof(float), hipMemcpyDeviceToHost,
                               S(stream)));
    IPH_TRY_HIP(hipStreamSynchronize(S(stream)));
    return IPH_OK;
}

IPH_API size_t iph_dev_proposal_search_workspace(uint32_t w, uint32_t h, uint32_t tw, uint32_t th,
                                                 int32_t method, int32_t k, int32_t n_top) {
    SearchPlan p;
    if (search_plan(w, h, tw, th, method, k, n_top, p) != IPH_OK) {
        clear_error();
        return 256;
    }
    const CandPlan cp = cand_plan(p.cw, p.ch);
    WorkspacePlan wp;
    wp.take<uint8_t>((size_t)p.sw * p.sh);
    wp.take<float>((size_t)p.cw * p.ch);
    if (!coarse_fused(p)) {
        wp.take<uint8_t>((size_t)p.stw * p.sth);
        wp.take<uint8_t>(mfma_match_workspace(IPH_U8, p.sw, p.sh, p.stw, p.sth, method));
    }
    wp.take<uint64_t>(std::max(cp.t, 1));
    for (int b = 0; b < 2; b++) wp.take<uint64_t>((size_t)std::max(cp.nb, 1) * kMaxCand);
    wp.take<int>(2 * (size_t)k);
    wp.take<float>(k);
    wp.take<float>((size_t)k * kPix);
    wp.take<uint8_t>(top_workspace(k, n_top));
    wp.take<int>(2 * (size_t)n_top);
    wp.take<float>(n_top);
    wp.take<uint8_t>(refine_workspace(n_top));
    return wp.used;
}

IPH_API iph_status iph_dev_proposal_search(const uint8_t *image, uint32_t w, uint32_t h,
                                           const uint8_t *tpl, uint32_t tw, uint32_t th,
                                           int32_t method, const uint16_t *weights,
                                           const float *bias, int32_t k, int32_t n_top,
                                           int32_t sep, float *result, void *workspace,
                                           size_t workspace_bytes, void *stream) {
    clear_error();
    SearchPlan p;
    IPH_TRY(search_plan(w, h, tw, th, method, k, n_top, p));
    IPH_TRY(check_candidates(p.cw, p.ch, k));
    if ((weights == nullptr) != (bias == nullptr)) {
        set_error("weights and bias go together (both NULL: rank by coarse score)");
        return IPH_EINVAL;
    }
    if (workspace_bytes < iph_dev_proposal_search_workspace(w, h, tw, th, method, k, n_top)) {
        set_error("workspace too small for the proposal search");
        return IPH_EINVAL;
    }
    const CandPlan cp = cand_plan(p.cw, p.ch);
    const bool fused = coarse_fused(p);
    Workspace wsp(workspace, workspace_bytes);
    uint8_t *small = wsp.take<uint8_t>((size_t)p.sw * p.sh);
    float *coarse = wsp.take<float>((size_t)p.cw * p.ch);
    uint8_t *stpl = nullptr, *mws = nullptr;
    size_t mws_bytes = 0;
    if (!fused) {
        stpl = wsp.take<uint8_t>((size_t)p.stw * p.sth);
        mws_bytes = mfma_match_workspace(IPH_U8, p.sw, p.sh, p.stw, p.sth, method);
        mws = wsp.take<uint8_t>(mws_bytes);
    }
    uint64_t *tk = wsp.take<uint64_t>(std::max(cp.t, 1));
    uint64_t *lk[2];
    for (int b = 0; b < 2; b++) lk[b] = wsp.take<uint64_t>((size_t)std::max(cp.nb, 1) * kMaxCand);
    int *cand = wsp.take<int>(2 * (size_t)k);
    float *cscore = wsp.take<float>(k);
    float *logits = wsp.take<float>((size_t)k * kPix);
    const size_t tws_bytes = top_workspace(k, n_top);
    uint8_t *tws = wsp.take<uint8_t>(tws_bytes);
    int *ctr = wsp.take<int>(2 * (size_t)n_top);
    float *probs = wsp.take<float>(n_top);
    const size_t rws_bytes = refine_workspace(n_top);
    uint8_t *rws = wsp.take<uint8_t>(rws_bytes);
    if (!rws || image == nullptr || tpl == nullptr || result == nullptr) {
        set_error("null image, template, result or workspace");
        return IPH_EINVAL;
    }
    hipStream_t s = S(stream);
    IPH_TRY(launch_downsample(image, w, h, small, s));
    if (fused) {
        const dim3 grid((p.cw + 255) / 256, cp.ty);
        const int vec = (p.sw % 4 == 0) && ((uintptr_t)small % 4 == 0);
        const int twq = (int)(p.stw + 3) / 4;
#define CTR(M, Q, R)                                                                          \
    coarse_tiles_kernel<M, Q, R><<<grid, 256, 0, s>>>(small, (int)p.sw, (int)p.sh, tpl,         \
                                                      (int)tw, (int)th, (int)p.cw, (int)p.ch,  \
                                                      cp.tx, vec, coarse, tk)
        // templates no taller than wide sum 4 * Q rows, others all kCoarseMax
#define CT(M, Q)                                                                             \
    if ((int)p.sth <= 4 * (Q)) CTR(M, Q, 4 * (Q)); else CTR(M, Q, kCoarseMax)
#define CTM(M)                                                                               \
    switch (twq) {                                                                           \
    case 1: CT(M, 1); break;                                                                 \
    case 2: CT(M, 2); break;                                                                 \
    case 3: CT(M, 3); break;                                                                 \
    default: CT(M, 4); break;                                                                \
    }
        if (method == IPH_CCORR) {
            CTM(IPH_CCORR)
        } else {
            CTM(IPH_CCORR_NORMED)
        }
#undef CTM
#undef CT
#undef CTR
        IPH_TRY(launch_status("coarse_tiles_kernel"));
    } else {
        IPH_TRY(launch_downsample(tpl, tw, th, stpl, s));
        IPH_TRY(launch_mfma_match(small, IPH_U8, p.sw, p.sh, stpl, p.stw, p.sth, method, coarse,
                                  mws, mws_bytes, s));
        tile_max_kernel<<<dim3((p.cw + 255) / 256, cp.ty), 256, 0, s>>>(
            coarse, (int)p.cw, (int)p.ch, cp.tx, tk);
        IPH_TRY(launch_status("tile_max_kernel"));
    }
    IPH_TRY(launch_topk(tk, cp.t, k, coarse, p.cw, cand, cscore, lk[0], lk[1], s));
    const int *centers = cand;
    const float *p_top = nullptr;
    if (weights) {
        cnn_proposal_forward_kernel<<<k, kFwdThreads, 0, s>>>(coarse, (int)p.cw, (int)p.ch, small,
                                                      (int)p.sw, (int)p.sh, (int)p.stw / 2, cand,
                                                      cscore, weights, bias, IPH_CNN_LAYERS,
                                                      logits);
        IPH_TRY(launch_status("cnn_proposal_forward_kernel"));
        IPH_TRY(launch_top(logits, 1, k, cand, p.cw, p.ch, n_top, ctr, probs, tws, tws_bytes,
                           s));
        centers = ctr;
        p_top = probs;
    }
    return launch_refine(image, w, h, tpl, tw, th, method, centers, n_top, 4, sep, p_top, result,
                         result + 4, rws, rws_bytes, s);
}

}  // extern "C"
