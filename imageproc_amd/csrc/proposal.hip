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
    for (int i = tid; i < kC * (kKK / 8); i += 256) {
        const int r = i / (kKK / 8), q = i - r * (kKK / 8);
        *reinterpret_cast<uint4 *>(ws + r * kWPitch + 8 * q) =
            *reinterpret_cast<const uint4 *>(w + r * kKK + 8 * q);
    }
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
        // A row m = n: pixel 32*mt + n; B column n: output channel n
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
        // D: column n = output channel, rows 8*(r/4) + 4*g + r%4 = pixels of the tile
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int pix = 32 * mt + 8 * (r >> 2) + 4 * g + (r & 3);
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
__global__ __launch_bounds__(256) void downsample4_kernel(const uint8_t *__restrict__ src, int W,
                                                          int dw, int dh,
                                                          uint8_t *__restrict__ dst) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= dw || y >= dh) return;
    uint32_t s = 0;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const uint8_t *row = src + (size_t)(4 * y + r) * W + 4 * x;
        s += row[0] + row[1] + row[2] + row[3];
    }
    dst[(size_t)y * dw + x] = (uint8_t)((s + 8) >> 4);
}

// Larger score first; NaN counts as the lowest; equal scores: lower index.
__device__ __forceinline__ bool better(float a, int ia, float b, int ib) {
    const bool an = a != a, bn = b != b;
    if (an != bn) return bn;
    if (!an && a != b) return a > b;
    return ia < ib;
}

// One 16x16 tile of the coarse map per workgroup: its best value and index.
__global__ __launch_bounds__(256) void tile_max_kernel(const float *__restrict__ map, int cw, int ch,
                                                       int tiles_x, float *__restrict__ tscore,
                                                       int *__restrict__ tidx) {
    __shared__ float sv[256];
    __shared__ int si[256];
    const int t = blockIdx.x, tx = t % tiles_x, ty = t / tiles_x;
    const int x = tx * 16 + (threadIdx.x & 15), y = ty * 16 + (threadIdx.x >> 4);
    float v = -INFINITY;
    int idx = 0x7fffffff;
    if (x < cw && y < ch) {
        v = map[(size_t)y * cw + x];
        idx = y * cw + x;
    }
    sv[threadIdx.x] = v;
    si[threadIdx.x] = idx;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o && better(sv[threadIdx.x + o], si[threadIdx.x + o], sv[threadIdx.x],
                                      si[threadIdx.x])) {
            sv[threadIdx.x] = sv[threadIdx.x + o];
            si[threadIdx.x] = si[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        tscore[t] = sv[0];
        tidx[t] = si[0];
    }
}

// The k best tile maxima (bitonic sort of up to 16384 entries in LDS).
constexpr int kMaxTiles = 16384;
__global__ __launch_bounds__(1024) void topk_tiles_kernel(const float *__restrict__ tscore,
                                                          const int *__restrict__ tidx, int n,
                                                          int k, int cw,
                                                          int *__restrict__ cand,
                                                          float *__restrict__ cscore) {
    extern __shared__ __align__(16) unsigned char smem[];
    float *sv = reinterpret_cast<float *>(smem);
    int *si = reinterpret_cast<int *>(sv + kMaxTiles);
    int np = 1;
    while (np < n) np <<= 1;
    for (int i = threadIdx.x; i < np; i += 1024) {
        sv[i] = i < n ? tscore[i] : -INFINITY;
        si[i] = i < n ? tidx[i] : 0x7fffffff;
    }
    __syncthreads();
    for (int size = 2; size <= np; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < np; i += 1024) {
                const int j = i ^ stride;
                if (j > i) {
                    const bool desc = (i & size) == 0;
                    const bool swap = desc ? better(sv[j], si[j], sv[i], si[i])
                                           : better(sv[i], si[i], sv[j], si[j]);
                    if (swap) {
                        float tv = sv[i];
                        sv[i] = sv[j];
                        sv[j] = tv;
                        int ti = si[i];
                        si[i] = si[j];
                        si[j] = ti;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < k; i += 1024) {
        const int id = i < n ? si[i] : 0;
        cand[2 * i] = id % cw;
        cand[2 * i + 1] = id / cw;
        cscore[i] = i < n ? sv[i] : -INFINITY;
    }
}

// Feature patch of candidate c (one workgroup): 16x16 around its coarse
// place, channels: 0 the coarse score minus the candidate's, x100; 1 the
// candidate's score minus the best candidate's, x100 (the same everywhere);
// 2 the small image's pixel at the window centre, / 255 - 0.5; 3 one inside
// the coarse map; the rest zero.
__global__ __launch_bounds__(256) void features_kernel(const float *__restrict__ map, int cw, int ch,
                                                       const uint8_t *__restrict__ small, int sw,
                                                       int sh, int half_t,
                                                       const int *__restrict__ cand,
                                                       const float *__restrict__ cscore,
                                                       uint16_t *__restrict__ patches) {
    const int c = blockIdx.x, pix = threadIdx.x;
    const int px = pix & 15, py = pix >> 4;
    const int x = cand[2 * c] - 8 + px, y = cand[2 * c + 1] - 8 + py;
    const float peak = cscore[c], best = cscore[0];
    const bool in = x >= 0 && x < cw && y >= 0 && y < ch;
    float f[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (in) {
        const float v = map[(size_t)y * cw + x];
        f[0] = (v == v && peak == peak) ? (v - peak) * 100.0f : -100.0f;
        const int cx = min(x + half_t, sw - 1), cy = min(y + half_t, sh - 1);
        f[2] = small[(size_t)cy * sw + cx] * (1.0f / 255.0f) - 0.5f;
        f[3] = 1.0f;
    }
    f[1] = (peak == peak && best == best) ? (peak - best) * 100.0f : -100.0f;
    uint16_t *o = patches + ((size_t)c * kPix + pix) * kC;
    uint32_t w0 = (uint32_t)to_bf16(f[0]) | ((uint32_t)to_bf16(f[1]) << 16);
    uint32_t w1 = (uint32_t)to_bf16(f[2]) | ((uint32_t)to_bf16(f[3]) << 16);
    reinterpret_cast<uint4 *>(o)[0] = make_uint4(w0, w1, 0u, 0u);
    for (int q = 1; q < 4; q++) reinterpret_cast<uint4 *>(o)[q] = make_uint4(0u, 0u, 0u, 0u);
}

// The n_top best logits (channel 0 of the last layer) at distinct coarse
// places inside the map; probs get their softmax probabilities over all k*256.
__global__ __launch_bounds__(256) void top_logits_kernel(const float *__restrict__ zl, int k,
                                                         const int *__restrict__ cand, int cw,
                                                         int ch, int n_top,
                                                         int *__restrict__ centers,
                                                         float *__restrict__ probs) {
    __shared__ float sv[256];
    __shared__ int si[256];
    __shared__ int taken[2 * 64];
    __shared__ float norm[2];
    const int n = k * kPix, tid = threadIdx.x;
    auto place = [&](int i, int &x, int &y) {
        const int c = i / kPix, pix = i - c * kPix;
        x = cand[2 * c] - 8 + (pix & 15);
        y = cand[2 * c + 1] - 8 + (pix >> 4);
    };
    // softmax normaliser over every logit
    float mx = -INFINITY;
    for (int i = tid; i < n; i += 256) mx = fmaxf(mx, zl[(size_t)i * kC]);
    sv[tid] = mx;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o) sv[tid] = fmaxf(sv[tid], sv[tid + o]);
        __syncthreads();
    }
    mx = sv[0];
    __syncthreads();
    float se = 0.0f;
    for (int i = tid; i < n; i += 256) se += expf(zl[(size_t)i * kC] - mx);
    sv[tid] = se;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o) sv[tid] += sv[tid + o];
        __syncthreads();
    }
    if (tid == 0) {
        norm[0] = mx;
        norm[1] = sv[0];
    }
    __syncthreads();
    for (int r = 0; r < n_top; r++) {
        float bv = -INFINITY;
        int bi = 0x7fffffff;
        for (int i = tid; i < n; i += 256) {
            int x, y;
            place(i, x, y);
            if (x < 0 || x >= cw || y < 0 || y >= ch) continue;
            bool dup = false;
            for (int q = 0; q < r; q++) dup |= (taken[2 * q] == x && taken[2 * q + 1] == y);
            const float v = zl[(size_t)i * kC];
            if (!dup && better(v, i, bv, bi)) {
                bv = v;
                bi = i;
            }
        }
        sv[tid] = bv;
        si[tid] = bi;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if (tid < o && better(sv[tid + o], si[tid + o], sv[tid], si[tid])) {
                sv[tid] = sv[tid + o];
                si[tid] = si[tid + o];
            }
            __syncthreads();
        }
        if (tid == 0) {
            int x = 0, y = 0;
            if (si[0] != 0x7fffffff) place(si[0], x, y);
            taken[2 * r] = x;
            taken[2 * r + 1] = y;
            centers[2 * r] = x;
            centers[2 * r + 1] = y;
            probs[r] = si[0] != 0x7fffffff ? expf(sv[0] - norm[0]) / norm[1] : 0.0f;
        }
        __syncthreads();
    }
}

// Exact scores of the (2R+1)^2 window places around one centre per
// workgroup (places outside [0, W-tw] x [0, H-th] are skipped):
// scores[c][(dy+R)*(2R+1) + dx+R], with NaN for skipped places.
constexpr int kRefR = 8;
constexpr int kRefN = (2 * kRefR + 1) * (2 * kRefR + 1);
template <int METHOD>
__global__ __launch_bounds__(256) void refine_kernel(const uint8_t *__restrict__ img, int W, int H,
                                                     const uint8_t *__restrict__ tpl, int tw,
                                                     int th, const int *__restrict__ centers,
                                                     int scale,
                                                     const long long *__restrict__ stt_ptr,
                                                     float *__restrict__ scores) {
    extern __shared__ __align__(16) unsigned char smem[];
    uint8_t *ts = smem;                         // th x tw
    uint8_t *is = smem + ((tw * th + 15) & ~15);  // (th + 2R) x (tw + 2R)
    const int c = blockIdx.x;
    const int cx = centers[2 * c] * scale, cy = centers[2 * c + 1] * scale;
    const int ox = cx - kRefR, oy = cy - kRefR;
    const int iw = tw + 2 * kRefR, ih = th + 2 * kRefR;
    for (int i = threadIdx.x; i < tw * th; i += 256) ts[i] = tpl[i];
    for (int i = threadIdx.x; i < iw * ih; i += 256) {
        const int r = i / iw, q = i - r * iw;
        const int gx = ox + q, gy = oy + r;
        is[i] = (gx >= 0 && gx < W && gy >= 0 && gy < H) ? img[(size_t)gy * W + gx] : 0;
    }
    __syncthreads();
    const long long stt = *stt_ptr;
    for (int pos = threadIdx.x; pos < kRefN; pos += 256) {
        const int dy = pos / (2 * kRefR + 1), dx = pos - dy * (2 * kRefR + 1);
        const int x = ox + dx, y = oy + dy;
        float s = __int_as_float(0x7fc00000);
        if (x >= 0 && x <= W - tw && y >= 0 && y <= H - th) {
            long long sit = 0, sii = 0;
            for (int r = 0; r < th; r++) {
                const uint8_t *ir = is + (dy + r) * iw + dx;
                const uint8_t *tr = ts + r * tw;
                int a = 0, b = 0;
                for (int q = 0; q < tw; q++) {
                    const int iv = ir[q];
                    a += iv * tr[q];
                    b += iv * iv;
                }
                sit += a;
                sii += b;
            }
            s = finish_score<METHOD>(sit, sii, stt, (long long)tw * th);
        }
        scores[(size_t)c * kRefN + pos] = s;
    }
}

// Best place over every refined window (larger score, then smaller y, then
// smaller x: find_extremes' first occurrence) and the best score of the
// places more than `sep` pixels from it (Chebyshev distance). result: x, y,
// best, second.
__global__ __launch_bounds__(256) void refine_best_kernel(const float *__restrict__ scores,
                                                          const int *__restrict__ centers, int n_c,
                                                          int scale, int W, int sep,
                                                          float *__restrict__ result) {
    __shared__ float sv[256];
    __shared__ long long si[256];
    const int n = n_c * kRefN, tid = threadIdx.x;
    auto key = [&](int i) -> long long {
        const int c = i / kRefN, pos = i - c * kRefN;
        const int dy = pos / (2 * kRefR + 1), dx = pos - dy * (2 * kRefR + 1);
        const long long x = centers[2 * c] * scale - kRefR + dx;
        const long long y = centers[2 * c + 1] * scale - kRefR + dy;
        return y * W + x;
    };
    auto bt = [](float a, long long ia, float b, long long ib) {
        const bool an = a != a, bn = b != b;
        if (an != bn) return bn;
        if (!an && a != b) return a > b;
        return ia < ib;
    };
    float bv = __int_as_float(0x7fc00000);
    long long bk = 0x7fffffffffffffffLL;
    for (int i = tid; i < n; i += 256) {
        const float v = scores[i];
        if (v != v) continue;
        const long long k = key(i);
        if (bt(v, k, bv, bk)) {
            bv = v;
            bk = k;
        }
    }
    sv[tid] = bv;
    si[tid] = bk;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o && bt(sv[tid + o], si[tid + o], sv[tid], si[tid])) {
            sv[tid] = sv[tid + o];
            si[tid] = si[tid + o];
        }
        __syncthreads();
    }
    const float best = sv[0];
    const long long bkey = si[0];
    __syncthreads();
    const long long bx = bkey % W, by = bkey / W;
    float sec = -INFINITY;
    for (int i = tid; i < n; i += 256) {
        const float v = scores[i];
        if (v != v) continue;
        const long long k = key(i);
        const long long x = k % W, y = k / W;
        if (llabs(x - bx) > sep || llabs(y - by) > sep) sec = fmaxf(sec, v);
    }
    sv[tid] = sec;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o) sv[tid] = fmaxf(sv[tid], sv[tid + o]);
        __syncthreads();
    }
    if (tid == 0) {
        result[0] = (float)bx;
        result[1] = (float)by;
        result[2] = best;
        result[3] = sv[0];
    }
}

__global__ void template_sq_u8_kernel(const uint8_t *__restrict__ tpl, int n,
                                      long long *__restrict__ out) {
    __shared__ long long red[256];
    long long s = 0;
    for (int i = threadIdx.x; i < n; i += 256) s += (long long)tpl[i] * tpl[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

}  // namespace
}  // namespace iph

using namespace iph;

namespace {
inline hipStream_t S(void *s) { return static_cast<hipStream_t>(s); }
inline unsigned blocks(size_t n) { return (unsigned)((n + 255) / 256); }
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
    const int dw = (int)(w / 4), dh = (int)(h / 4);
    if (dw == 0 || dh == 0) return IPH_OK;
    dim3 grid((dw + 63) / 64, (dh + 3) / 4);
    downsample4_kernel<<<grid, 256, 0, S(stream)>>>(src, (int)w, dw, dh, dst);
    return launch_status("downsample4_kernel");
}

IPH_API size_t iph_dev_proposal_candidates_workspace(uint32_t cw, uint32_t ch) {
    const size_t t = (size_t)((cw + 15) / 16) * ((ch + 15) / 16);
    return std::max<size_t>(t, 1) * (sizeof(float) + sizeof(int)) + 256;
}

IPH_API iph_status iph_dev_proposal_candidates(const float *coarse, uint32_t cw, uint32_t ch,
                                               int32_t k, int32_t *cand, float *cand_score,
                                               void *workspace, size_t workspace_bytes,
                                               void *stream) {
    clear_error();
    const int tx = (int)((cw + 15) / 16), ty = (int)((ch + 15) / 16), t = tx * ty;
    if (t <= 0 || t > kMaxTiles || k <= 0 || k > 1024) {
        set_error("coarse map of %ux%u or %d candidates out of range", cw, ch, k);
        return IPH_EINVAL;
    }
    if (workspace_bytes < iph_dev_proposal_candidates_workspace(cw, ch)) {
        set_error("workspace too small for the candidate search");
        return IPH_EINVAL;
    }
    float *ts = static_cast<float *>(workspace);
    int *ti = reinterpret_cast<int *>(ts + t);
    tile_max_kernel<<<t, 256, 0, S(stream)>>>(coarse, (int)cw, (int)ch, tx, ts, ti);
    IPH_TRY(launch_status("tile_max_kernel"));
    topk_tiles_kernel<<<1, 1024, kMaxTiles * (sizeof(float) + sizeof(int)), S(stream)>>>(
        ts, ti, t, k, (int)cw, cand, cand_score);
    return launch_status("topk_tiles_kernel");
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

IPH_API iph_status iph_dev_proposal_top(const float *z_last, int32_t k, const int32_t *cand,
                                        uint32_t cw, uint32_t ch, int32_t n_top, int32_t *centers,
                                        float *probs, void *stream) {
    clear_error();
    if (n_top <= 0 || n_top > 64) {
        set_error("between 1 and 64 top candidates");
        return IPH_EINVAL;
    }
    top_logits_kernel<<<1, 256, 0, S(stream)>>>(z_last, k, cand, (int)cw, (int)ch, n_top, centers,
                                                probs);
    return launch_status("top_logits_kernel");
}

IPH_API size_t iph_dev_proposal_refine_workspace(int32_t n_centers) {
    return 256 + (size_t)std::max(n_centers, 1) * kRefN * sizeof(float);
}

IPH_API iph_status iph_dev_proposal_refine(const uint8_t *image, uint32_t w, uint32_t h,
                                           const uint8_t *tpl, uint32_t tw, uint32_t th,
                                           int32_t method, const int32_t *centers,
                                           int32_t n_centers, int32_t scale, int32_t sep,
                                           float *result, void *workspace, size_t workspace_bytes,
                                           void *stream) {
    clear_error();
    IPH_TRY(check_match(IPH_U8, w, h, tw, th, method, IPH_PATH_AUTO));
    if (tw == 0 || th == 0 || tw > 128 || th > 128 || n_centers <= 0 || n_centers > 1024) {
        set_error("refinement takes templates of 1 to 128 pixels a side and 1 to 1024 centres");
        return IPH_EINVAL;
    }
    if (workspace_bytes < iph_dev_proposal_refine_workspace(n_centers)) {
        set_error("workspace too small for the refinement");
        return IPH_EINVAL;
    }
    long long *stt = static_cast<long long *>(workspace);
    float *scores = reinterpret_cast<float *>(static_cast<char *>(workspace) + 256);
    template_sq_u8_kernel<<<1, 256, 0, S(stream)>>>(tpl, (int)(tw * th), stt);
    IPH_TRY(launch_status("template_sq_u8_kernel"));
    const size_t lds = ((tw * th + 15) & ~15u) + (size_t)(tw + 2 * kRefR) * (th + 2 * kRefR);
#define RF(M)                                                                                \
    refine_kernel<M><<<n_centers, 256, lds, S(stream)>>>(image, (int)w, (int)h, tpl, (int)tw, \
                                                         (int)th, centers, scale, stt, scores)
    switch (method) {
    case IPH_SSE: RF(IPH_SSE); break;
    case IPH_SSE_NORMED: RF(IPH_SSE_NORMED); break;
    case IPH_CCORR: RF(IPH_CCORR); break;
    default: RF(IPH_CCORR_NORMED); break;
    }
#undef RF
    IPH_TRY(launch_status("refine_kernel"));
    // SSE methods pick the smallest score: negate the search by refining
    // the best of -score is not needed for the correlation methods this path
    // serves; the caller reads `result` as (x, y, best, second).
    refine_best_kernel<<<1, 256, 0, S(stream)>>>(scores, centers, n_centers, scale, (int)w, sep,
                                                 result);
    return launch_status("refine_best_kernel");
}

}  // extern "C"
