// MFMA correlation engine for match_template.
//
// u8 images: exact integer correlation on v_mfma_i32_32x32x32_i8.
//   C[y][x] = sum_{dy,dx} I[y+dy][x+dx] T[dy][dx]
// is computed per 32x32 output tile as
//   C_tile = sum_dy sum_kb A(dy,kb) x B(dy,kb)
// where A(dy,kb) is a 32x32 block of the image (rows y0+dy.., columns
// x0+32kb..), read straight from an LDS copy of the image block, and B(dy,kb)
// is the Toeplitz matrix of template row dy: B[k][n] = T[dy][32kb + k - n]
// (zero outside the row). Both operands are offset by -128 to fit i8:
//   sum I T = sum I'T' + 128 sum_window I + 128 sum T - 128^2 N,
// with I' = I - 128, T' = T - 128, every term an exact integer. The window
// sums sum I and sum I^2 come from a separate exact integer pass, so
//   SumOfSquaredErrors  = sum T^2 - 2 C + sum I^2      (exact)
//   CrossCorrelation    = C                             (exact)
// each rounded once to f32, and the normalised methods divide by
// sqrt(sum I^2 * sum T^2): in f32 like the reference for windows under 64x64
// pixels, in f64 for windows of at least 64x64.
//
// f32 images: BF16x3 on v_mfma_f32_32x32x16_bf16 with FP32 accumulation.
// Each operand is split x = hi + lo (hi = bf16(x), lo = bf16(x - hi)) and
//   sum I T ~= sum (I_hi T_hi + I_hi T_lo + I_lo T_hi),
// with the same image-block x Toeplitz decomposition at K = 16. Template
// rows are processed in chunks of 32, restaging the image rows of each chunk
// in LDS, so any template height fits. The window energy sum I^2 is summed
// in f64.
#include "match_common.hpp"

#include <algorithm>
#include <type_traits>

namespace iph {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int kMaxSide = (int)kMfmaMaxSide;  // largest template side on the int8 path
static_assert(kMaxSide * 255 < 65536, "a template row's pixel sum fits 16 bits");
constexpr int kWinRows = 64;   // output rows per window-sum workgroup
constexpr int kWinCols = 64;   // output columns per window-sum workgroup
constexpr int kHsPitch = kWinCols + 1;  // LDS row pitch of the horizontal sums (odd: no bank clash)
constexpr int kWinThreads = 512;  // threads per window-sum workgroup (8 rows per column segment)

__device__ __forceinline__ int align_up_dev(int v) { return (v + 15) & ~15; }
// Staged bytes per row of the window-sum tile, a multiple of 4.
__host__ __device__ __forceinline__ int win_cols(int tw) { return (kWinCols + tw - 1 + 3) & ~3; }

// Exact window sums (sum I, sum I^2) for every output position of a
// kWinCols x kWinRows output tile. The staged window (rows 4-byte aligned,
// zeros past the image: those bytes only reach discarded outputs) gets its
// horizontal tw-window sums per 16-column segment of a row: the first by
// v_dot4 over whole words, then 15 sliding steps. Each output column then
// slides its vertical th-window sum down a 16-row segment. All arithmetic
// is exact u32; written coalesced.
__global__ __launch_bounds__(kWinThreads) void window_sums_u8_kernel(const uint8_t *__restrict__ img,
                                                             int W, int tw, int th, int ow,
                                                             int oh, uint2 *__restrict__ ws) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int x0 = blockIdx.x * kWinCols, y0 = blockIdx.y * kWinRows;
    const int cols = win_cols(tw), cw4 = cols >> 2, rows = kWinRows + th - 1;
    const int H = oh + th - 1;
    const bool aligned = (W & 3) == 0 && (reinterpret_cast<uintptr_t>(img) & 3) == 0;
    uint32_t *blk = reinterpret_cast<uint32_t *>(smem);                      // rows x cw4 words
    // horizontal sums: sum I (below 2^15 for tw <= 128) and sum I^2, rows x kHsPitch each
    uint16_t *hsa = reinterpret_cast<uint16_t *>(smem + align_up_dev(rows * cols));
    uint32_t *hsb = reinterpret_cast<uint32_t *>(smem + align_up_dev(rows * cols) +
                                                 align_up_dev(rows * kHsPitch * 2));
    // 16 loads per thread in flight
    constexpr int kStage = 16;
    for (int base = threadIdx.x; base < rows * cw4; base += kStage * kWinThreads) {
        uint32_t v[kStage];
#pragma unroll
        for (int b = 0; b < kStage; b++) {
            const int i = min(base + kWinThreads * b, rows * cw4 - 1);
            const int r = i / cw4, gx = x0 + 4 * (i - r * cw4), gy = y0 + r;
            uint32_t wv = 0;
            if (gy < H) {
                const uint8_t *row = img + (size_t)gy * W;
                if (aligned && gx + 4 <= W) {
                    wv = *reinterpret_cast<const uint32_t *>(row + gx);
                } else {
#pragma unroll
                    for (int k = 0; k < 4; k++)
                        if (gx + k < W) wv |= (uint32_t)row[gx + k] << (8 * k);
                }
            }
            v[b] = wv;
        }
#pragma unroll
        for (int b = 0; b < kStage; b++)
            if (base + kWinThreads * b < rows * cw4) blk[base + kWinThreads * b] = v[b];
    }
    __syncthreads();
    constexpr int kSegs = kWinCols / 16;
    const int nw = tw >> 2;
    const uint32_t tail = (tw & 3) ? (1u << (8 * (tw & 3))) - 1u : 0u;
    for (int it = threadIdx.x; it < rows * kSegs; it += kWinThreads) {
        const int r = it / kSegs, q = it - r * kSegs;
        const uint32_t *pw = blk + r * cw4 + 4 * q;
        uint32_t a = 0, b = 0;
#pragma unroll 4
        for (int k = 0; k < nw; k++) {
            const uint32_t wv = pw[k];
            a = __builtin_amdgcn_udot4(wv, 0x01010101u, a, false);
            b = __builtin_amdgcn_udot4(wv, wv, b, false);
        }
        if (tail) {
            const uint32_t wv = pw[nw] & tail;
            a = __builtin_amdgcn_udot4(wv, 0x01010101u, a, false);
            b = __builtin_amdgcn_udot4(wv, wv, b, false);
        }
        const uint8_t *p = reinterpret_cast<const uint8_t *>(pw);
        uint16_t *oa = hsa + r * kHsPitch + 16 * q;
        uint32_t *ob = hsb + r * kHsPitch + 16 * q;
        oa[0] = (uint16_t)a;
        ob[0] = b;
#pragma unroll
        for (int x = 1; x < 16; x++) {
            uint32_t vin = p[x + tw - 1], vout = p[x - 1];
            a += vin - vout;
            b += vin * vin - vout * vout;
            oa[x] = (uint16_t)a;
            ob[x] = b;
        }
    }
    __syncthreads();
    const int x = threadIdx.x & (kWinCols - 1), seg = threadIdx.x / kWinCols;
    constexpr int kSeg = kWinRows / (kWinThreads / kWinCols);
    const int ya = seg * kSeg;
    uint32_t a = 0, b = 0;
#pragma unroll 8
    for (int dy = 0; dy < th; dy++) {
        a += hsa[(ya + dy) * kHsPitch + x];
        b += hsb[(ya + dy) * kHsPitch + x];
    }
#pragma unroll
    for (int y = ya; y < ya + kSeg; y++) {
        const int gy = y0 + y, gx = x0 + x;
        if (gy < oh && gx < ow) ws[(size_t)gy * ow + gx] = make_uint2(a, b);
        const int in = y + th < rows ? (y + th) * kHsPitch + x : 0, out = y * kHsPitch + x;
        a += (uint32_t)hsa[in] - (uint32_t)hsa[out];
        b += hsb[in] - hsb[out];
    }
}

// Bytes per template row staged in LDS: 32 zeros, the row, zeros up to
// 32 * (KB + 1) (every Toeplitz slice of a lane reads inside it).
__host__ __device__ constexpr int tpl_pitch(int KB) { return 32 * KB + 32; }

// One wave computes a 32 x (32*XT) output strip; RT waves stack vertically.
// The workgroup stages the template (as i8, zero-padded rows) next to its
// image rows; each lane cuts its Toeplitz B operand for template row dy and
// K block kb, B[k][n] = T[dy][32kb + k - n], out of the padded row with
// five dword reads and v_alignbyte, and the workgroup sums T and T^2 itself.
template <int KB, int METHOD, int XT>
__global__ __launch_bounds__(256) void match_i8_mfma_kernel(
    const uint8_t *__restrict__ img, int W, int H, const uint8_t *__restrict__ tpl,
    const uint2 *__restrict__ wsum, float *__restrict__ out, int ow, int oh, int tw, int th,
    int pitch) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ long long tred[2][4];
    int8_t *blk = reinterpret_cast<int8_t *>(smem);
    constexpr int BW = 32 * (XT + KB - 1);  // bytes per staged row
    const int RT = blockDim.x >> 6;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int x0 = blockIdx.x * (32 * XT), y0 = blockIdx.y * 32 * RT;
    const int rows = 32 * RT + th - 1;

    // Stage image rows y0.. and columns x0.. as i8 = u8 - 128. Bytes outside
    // the image only meet zero Toeplitz entries or produce discarded outputs.
    const bool aligned = (W & 3) == 0 && (reinterpret_cast<uintptr_t>(img) & 3) == 0 && W >= 4;
    const int total = rows * (BW / 4);
    // loads are issued in batches of 8 before their LDS stores, so a
    // workgroup waits for global latency once per batch, not per dword
    for (int base = threadIdx.x; base < total; base += 8 * (int)blockDim.x) {
        uint32_t v[8];
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const int i = min(base + b * (int)blockDim.x, total - 1);
            const int r = i / (BW / 4), c = (i - r * (BW / 4)) * 4;
            const uint8_t *row = img + (size_t)min(y0 + r, H - 1) * W;
            const int gx = x0 + c;
            if (aligned) {
                v[b] = *reinterpret_cast<const uint32_t *>(row + min(gx, W - 4));
            } else {
                v[b] = (uint32_t)row[min(gx, W - 1)] | ((uint32_t)row[min(gx + 1, W - 1)] << 8) |
                       ((uint32_t)row[min(gx + 2, W - 1)] << 16) |
                       ((uint32_t)row[min(gx + 3, W - 1)] << 24);
            }
        }
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const int i = base + b * (int)blockDim.x;
            if (i < total) {
                const int r = i / (BW / 4), c = (i - r * (BW / 4)) * 4;
                *reinterpret_cast<uint32_t *>(blk + r * pitch + c) = v[b] ^ 0x80808080u;
            }
        }
    }
    // template rows as i8 = u8 - 128 at [32, 32 + tw) of kTP-byte rows
    constexpr int kTP = tpl_pitch(KB);
    uint32_t *tp = reinterpret_cast<uint32_t *>(smem + ((rows * pitch + 15) & ~15));
    uint32_t s1 = 0, s2 = 0;
    for (int i = threadIdx.x; i < th * (kTP / 4) + 4; i += blockDim.x) {
        const int r = i / (kTP / 4), c = 4 * (i - r * (kTP / 4)) - 32;
        uint32_t wv = 0;
#pragma unroll
        for (int b = 0; b < 4; b++)
            if (r < th && c + b >= 0 && c + b < tw) {
                const uint32_t t = tpl[r * tw + c + b];
                s1 += t;
                s2 += t * t;
                wv |= (t ^ 0x80u) << (8 * b);
            }
        tp[i] = wv;
    }
    long long l1 = s1, l2 = s2;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        l1 += __shfl_xor(l1, o);
        l2 += __shfl_xor(l2, o);
    }
    if ((threadIdx.x & 63) == 0) {
        tred[0][threadIdx.x >> 6] = l1;
        tred[1][threadIdx.x >> 6] = l2;
    }
    __syncthreads();

    const int n = lane & 31, g = lane >> 5;
    v16i acc[XT];
#pragma unroll
    for (int t = 0; t < XT; t++)
#pragma unroll
        for (int r = 0; r < 16; r++) acc[t][r] = 0;

    const int8_t *arow = blk + (wave * 32 + n) * pitch + 16 * g;
    // lane (n, g) feeds B rows k = 16g + j of column n: bytes
    // 32 + 32kb + 16g - n .. +15 of the padded row, the same alignment every dy
    const int u0 = 32 + 16 * g - n;
    const uint32_t sh = (uint32_t)(u0 & 3);
    const uint32_t *trow = tp + (u0 >> 2);
    for (int dy = 0; dy < th; dy++) {
        v4i b[KB];
#pragma unroll
        for (int kb = 0; kb < KB; kb++) {
            const uint32_t *q = trow + 8 * kb;
            const uint32_t d0 = q[0], d1 = q[1], d2 = q[2], d3 = q[3], d4 = q[4];
            b[kb][0] = (int)__builtin_amdgcn_alignbyte(d1, d0, sh);
            b[kb][1] = (int)__builtin_amdgcn_alignbyte(d2, d1, sh);
            b[kb][2] = (int)__builtin_amdgcn_alignbyte(d3, d2, sh);
            b[kb][3] = (int)__builtin_amdgcn_alignbyte(d4, d3, sh);
        }
#pragma unroll
        for (int s = 0; s < XT + KB - 1; s++) {
            v4i a = *reinterpret_cast<const v4i *>(arow + 32 * s);
#pragma unroll
            for (int t = 0; t < XT; t++) {
                if (s - t >= 0 && s - t < KB)
                    acc[t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b[s - t], acc[t], 0, 0, 0);
            }
        }
        arow += pitch;
        trow += kTP / 4;
    }

    long long st = 0, stt = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
        st += tred[0][w];
        stt += tred[1][w];
    }
    const long long nn = (long long)tw * th;
    const long long base = 128LL * st - 16384LL * nn;
#pragma unroll
    for (int t = 0; t < XT; t++) {
        const int x = x0 + 32 * t + n;
        // all 16 window-sum loads of this column are issued before any use
        // (clamped addresses, so none is conditional)
        uint2 sums[16];
#pragma unroll
        for (int r = 0; r < 16; r++) {
            // 32x32 accumulator layout: register r of lane l holds row
            // 8*(r/4) + 4*(l/32) + r%4, column l%32.
            const int y = y0 + wave * 32 + 8 * (r >> 2) + 4 * g + (r & 3);
            sums[r] = wsum[(size_t)min(y, oh - 1) * ow + min(x, ow - 1)];
        }
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int y = y0 + wave * 32 + 8 * (r >> 2) + 4 * g + (r & 3);
            long long c = (long long)acc[t][r] + 128LL * sums[r].x + base;
            float v = finish_score<METHOD>(c, (long long)sums[r].y, stt, nn);
            if (x < ow && y < oh) out[(size_t)y * ow + x] = v;
        }
    }
}


// ---- BF16x3 path for f32 images --------------------------------------------------

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

constexpr int kXT16 = 2;     // 32-column output tiles per wave (f32 images)
constexpr int kDyChunk = 32; // template rows per LDS staging round

// Round-to-nearest-even f32 -> bf16 bits (finite inputs).
__device__ __forceinline__ uint16_t bf16_bits(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_value(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// hi/lo split of x: hi = bf16(x), lo = bf16(x - hi).
__device__ __forceinline__ void split_bf16(float x, uint16_t &hi, uint16_t &lo) {
    hi = bf16_bits(x);
    lo = bf16_bits(x - bf16_value(hi));
}

// sum T^2 of the template in f64 (single block tree reduction).
__global__ __launch_bounds__(256) void template_sq_f32_kernel(const float *__restrict__ tpl, int n,
                                                              double *__restrict__ out) {
    __shared__ double s[256];
    double a = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) a += (double)tpl[i] * (double)tpl[i];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = s[0];
}

// Toeplitz operands at K = 16 in MFMA lane order, hi and lo parts:
// entry [dy][kb][part][lane] holds 8 bf16: column n = lane % 32,
// rows k = 8*(lane/32) + j, value T[dy][16kb + k - n].
__global__ __launch_bounds__(64) void toeplitz_bf16_kernel(const float *__restrict__ tpl, int tw,
                                                           int KB, uint16_t *__restrict__ bsw) {
    const int dy = blockIdx.y, kb = blockIdx.x, lane = threadIdx.x;
    const int n = lane & 31, g = lane >> 5;
    uint16_t *hi = bsw + (((size_t)(dy * KB + kb) * 2 + 0) * 64 + lane) * 8;
    uint16_t *lo = bsw + (((size_t)(dy * KB + kb) * 2 + 1) * 64 + lane) * 8;
    for (int j = 0; j < 8; j++) {
        int dx = 16 * kb + 8 * g + j - n;
        uint16_t h = 0, l = 0;
        if (dx >= 0 && dx < tw) split_bf16(tpl[dy * tw + dx], h, l);
        hi[j] = h;
        lo[j] = l;
    }
}

// Window energy sum I^2 in f64 for every output: vertical sums over th rows,
// then horizontal sums over tw, each evaluated directly (no running
// differences, so no cancellation).
// Window energies sum(i^2) of an f32 image in f64, as two sliding-window
// passes. Horizontal: one workgroup per (row, 1024 outputs); p^2 of the
// needed columns is staged in LDS, each thread sums its first window and
// slides over 3 more (4 consecutive outputs).
constexpr int kHsOut = 1024;  // outputs per horizontal workgroup
__global__ __launch_bounds__(256) void window_sq_rows_f32_kernel(const float *__restrict__ img,
                                                                 int W, int tw, int ow,
                                                                 double *__restrict__ hs) {
    extern __shared__ __align__(16) unsigned char smem[];
    double *v = reinterpret_cast<double *>(smem);  // kHsOut + tw - 1
    const int y = blockIdx.y, x0 = blockIdx.x * kHsOut;
    const int ncols = min(kHsOut + tw - 1, W - x0);
    const float *row = img + (size_t)y * W + x0;
    for (int c = threadIdx.x; c < ncols; c += 256) {
        const double p = row[c];
        v[c] = p * p;
    }
    __syncthreads();
    const int xl = threadIdx.x * 4;
    if (x0 + xl >= ow) return;
    double a = 0.0;
    for (int d = 0; d < tw; d++) a += v[xl + d];
    double *o = hs + (size_t)y * ow + x0 + xl;
    o[0] = a;
#pragma unroll
    for (int q = 1; q < 4; q++) {
        if (x0 + xl + q >= ow) break;
        a += v[xl + q - 1 + tw] - v[xl + q - 1];
        o[q] = a;
    }
}

// Vertical: one workgroup per 32 columns x 128 output rows; each thread sums
// its first window of th horizontal sums and slides over 15 more rows.
constexpr int kVsCols = 32, kVsRows = 128;
__global__ __launch_bounds__(256) void window_sq_cols_f32_kernel(const double *__restrict__ hs,
                                                                 int th, int ow, int oh,
                                                                 double *__restrict__ ws) {
    const int x = blockIdx.x * kVsCols + (threadIdx.x & (kVsCols - 1));
    const int y0 = blockIdx.y * kVsRows + (threadIdx.x / kVsCols) * (kVsRows / 8);
    if (x >= ow || y0 >= oh) return;
    const double *col = hs + x;
    double a = 0.0;
    for (int r0 = 0; r0 < th; r0 += 8) {
        double t[8];
#pragma unroll
        for (int q = 0; q < 8; q++) t[q] = (r0 + q < th) ? col[(size_t)(y0 + r0 + q) * ow] : 0.0;
#pragma unroll
        for (int q = 0; q < 8; q++) a += t[q];
    }
    ws[(size_t)y0 * ow + x] = a;
    const int yend = min(y0 + kVsRows / 8, oh);
    for (int y = y0 + 1; y < yend; y++) {
        a += col[(size_t)(y - 1 + th) * ow] - col[(size_t)(y - 1) * ow];
        ws[(size_t)y * ow + x] = a;
    }
}

template <int METHOD>
__device__ __forceinline__ float finish_score_f32(float c, double sii, double stt, long long n) {
    const bool sse = (METHOD == IPH_SSE || METHOD == IPH_SSE_NORMED);
    if (METHOD == IPH_CCORR) return c;
    double num = sse ? (stt - 2.0 * (double)c + sii) : (double)c;
    if (METHOD == IPH_SSE) return (float)num;
    if (n < 64 * 64) {
        float norm = sqrtf((float)sii * (float)stt);
        float s = (float)num;
        return norm > 0.0f ? s / norm : s;
    }
    double norm = sqrt(sii * stt);
    return norm > 0.0 ? (float)(num / norm) : (float)num;
}

// One wave computes a 32 x (32*kXT16) output strip; RT waves stack vertically.
// LDS holds the (32*RT + kDyChunk - 1) image rows of one template-row chunk,
// split into bf16 hi and lo planes.
template <int KB, int METHOD>
__global__ __launch_bounds__(128) void match_bf16x3_mfma_kernel(
    const float *__restrict__ img, int W, int H, const v8bf *__restrict__ bsw,
    const double *__restrict__ wsq, const double *__restrict__ tsq, float *__restrict__ out,
    int ow, int oh, int tw, int th, int pitch) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int BW = 16 * (2 * kXT16 + KB - 1);  // staged columns
    const int RT = blockDim.x >> 6;
    const int rows = 32 * RT + kDyChunk - 1;
    uint16_t *hi_s = reinterpret_cast<uint16_t *>(smem);
    uint16_t *lo_s = hi_s + (size_t)rows * pitch;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int n = lane & 31, g = lane >> 5;
    const int x0 = blockIdx.x * (32 * kXT16), y0 = blockIdx.y * 32 * RT;

    v16f acc[kXT16];
#pragma unroll
    for (int t = 0; t < kXT16; t++)
#pragma unroll
        for (int r = 0; r < 16; r++) acc[t][r] = 0.0f;

    const v8bf *bp = bsw + lane;
    for (int c0 = 0; c0 < th; c0 += kDyChunk) {
        __syncthreads();
        for (int base = threadIdx.x; base < rows * BW; base += 8 * (int)blockDim.x) {
            float v[8];
#pragma unroll
            for (int b = 0; b < 8; b++) {
                const int i = base + b * (int)blockDim.x;
                const int r = i / BW, c = i - r * BW;
                const int gy = min(y0 + c0 + r, H - 1), gx = min(x0 + c, W - 1);
                v[b] = img[(size_t)gy * W + gx];
            }
#pragma unroll
            for (int b = 0; b < 8; b++) {
                const int i = base + b * (int)blockDim.x;
                if (i < rows * BW) {
                    const int r = i / BW, c = i - r * BW;
                    uint16_t h, l;
                    split_bf16(v[b], h, l);
                    hi_s[r * pitch + c] = h;
                    lo_s[r * pitch + c] = l;
                }
            }
        }
        __syncthreads();
        const int dy_end = min(kDyChunk, th - c0);
        const uint16_t *ah = hi_s + (wave * 32 + n) * pitch + 8 * g;
        const uint16_t *al = lo_s + (wave * 32 + n) * pitch + 8 * g;
        for (int d = 0; d < dy_end; d++) {
            const int dy = c0 + d;
            v8bf bh[KB], bl[KB];
#pragma unroll
            for (int kb = 0; kb < KB; kb++) {
                bh[kb] = bp[((size_t)(dy * KB + kb) * 2 + 0) * 64];
                bl[kb] = bp[((size_t)(dy * KB + kb) * 2 + 1) * 64];
            }
#pragma unroll
            for (int s = 0; s < 2 * kXT16 + KB - 1; s++) {
                v8bf a_h = *reinterpret_cast<const v8bf *>(ah + d * pitch + 16 * s);
                v8bf a_l = *reinterpret_cast<const v8bf *>(al + d * pitch + 16 * s);
#pragma unroll
                for (int t = 0; t < kXT16; t++) {
                    const int kb = s - 2 * t;
                    if (kb >= 0 && kb < KB) {
                        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a_l, bh[kb], acc[t], 0, 0, 0);
                        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a_h, bl[kb], acc[t], 0, 0, 0);
                        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a_h, bh[kb], acc[t], 0, 0, 0);
                    }
                }
            }
        }
    }

    const double stt = tsq[0];
    const long long nn = (long long)tw * th;
#pragma unroll
    for (int t = 0; t < kXT16; t++) {
        const int x = x0 + 32 * t + n;
        double sq[16];
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int y = y0 + wave * 32 + 8 * (r >> 2) + 4 * g + (r & 3);
            sq[r] = (METHOD == IPH_CCORR) ? 0.0 : wsq[(size_t)min(y, oh - 1) * ow + min(x, ow - 1)];
        }
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int y = y0 + wave * 32 + 8 * (r >> 2) + 4 * g + (r & 3);
            float v = finish_score_f32<METHOD>(acc[t][r], sq[r], stt, nn);
            if (x < ow && y < oh) out[(size_t)y * ow + x] = v;
        }
    }
}

template <int KB>
iph_status launch_bf16_kb(const float *img, int W, int H, const uint16_t *bsw, const double *wsq,
                          const double *tsq, float *out, int ow, int oh, int tw, int th,
                          int method, hipStream_t s) {
    constexpr int BW = 16 * (2 * kXT16 + KB - 1);
    int pitch = ((BW / 8) % 2 == 0) ? BW + 8 : BW;  // odd multiple of 16 bytes
    int RT = 2;
    while (RT > 1 && (size_t)(32 * RT + kDyChunk - 1) * pitch * 4 > 64 * 1024) RT /= 2;
    size_t lds = (size_t)(32 * RT + kDyChunk - 1) * pitch * 4;
    dim3 grid((ow + 32 * kXT16 - 1) / (32 * kXT16), (oh + 32 * RT - 1) / (32 * RT));
    const v8bf *b = reinterpret_cast<const v8bf *>(bsw);
#define L(M)                                                                                 \
    match_bf16x3_mfma_kernel<KB, M><<<grid, 64 * RT, lds, s>>>(img, W, H, b, wsq, tsq, out,   \
                                                               ow, oh, tw, th, pitch)
    switch (method) {
    case IPH_SSE: L(IPH_SSE); break;
    case IPH_SSE_NORMED: L(IPH_SSE_NORMED); break;
    case IPH_CCORR: L(IPH_CCORR); break;
    default: L(IPH_CCORR_NORMED); break;
    }
#undef L
    return launch_status("match_bf16x3_mfma_kernel");
}

constexpr int kMaxSideF32 = (int)kMfmaMaxSide;

int blocks16_for(int tw) { return (31 + tw + 15) / 16; }

int blocks_for(int tw) { return (31 + tw + 31) / 32; }

size_t win_lds(int tw, int th) {
    int cols = win_cols(tw), rows = kWinRows + th - 1;
    return align_up((size_t)rows * cols, 16) + align_up((size_t)rows * kHsPitch * 2, 16) +
           (size_t)rows * kHsPitch * sizeof(uint32_t);
}

thread_local int g_match_rows_hint = 0;

template <int KB, int XT>
iph_status launch_i8_kb_xt(const uint8_t *img, int W, int H, const uint8_t *tpl,
                           const uint2 *wsum, float *out, int ow, int oh, int tw, int th,
                           int method, hipStream_t s) {
    constexpr int BW = 32 * (XT + KB - 1);
    int pitch = ((BW / 16) % 2 == 0) ? BW + 16 : BW;  // odd multiple of 16: no bank conflicts
    int RT = g_match_rows_hint > 0 ? std::min(g_match_rows_hint, 4) : 4;
    const size_t tbytes = (size_t)th * tpl_pitch(KB) + 16;
    auto lds_for = [&](int rt) { return align_up((size_t)(32 * rt + th - 1) * pitch, 16) + tbytes; };
    while (RT > 1 && lds_for(RT) > 64 * 1024) RT /= 2;
    const size_t lds = lds_for(RT);
    dim3 grid((ow + 32 * XT - 1) / (32 * XT), (oh + 32 * RT - 1) / (32 * RT));
#define L(M)                                                                                 \
    match_i8_mfma_kernel<KB, M, XT><<<grid, 64 * RT, lds, s>>>(img, W, H, tpl, wsum, out, ow,  \
                                                               oh, tw, th, pitch)
    switch (method) {
    case IPH_SSE: L(IPH_SSE); break;
    case IPH_SSE_NORMED: L(IPH_SSE_NORMED); break;
    case IPH_CCORR: L(IPH_CCORR); break;
    default: L(IPH_CCORR_NORMED); break;
    }
#undef L
    return launch_status("match_i8_mfma_kernel");
}

// 32-column output tiles per wave: two (each A operand feeds more MFMAs)
// unless the map has enough 32x32 tiles to give every SIMD (1024 on the
// chip) 16 waves of one tile each; the kernel is latency-bound, so waves
// beat reuse there (8192x8192 with 48x48: 1.18 ms vs 1.23 ms vs 1.35 ms for
// 1, 2, 4 tiles).
int match_xt(int ow, int oh) {
    const long long tiles = (long long)((oh + 31) / 32) * ((ow + 31) / 32);
    return tiles >= 16 * 1024 ? 1 : 2;
}

template <int KB>
iph_status launch_i8_kb(const uint8_t *img, int W, int H, const uint8_t *tpl, const uint2 *wsum,
                        float *out, int ow, int oh, int tw, int th, int method, hipStream_t s) {
    switch (match_xt(ow, oh)) {
    case 2: return launch_i8_kb_xt<KB, 2>(img, W, H, tpl, wsum, out, ow, oh, tw, th, method, s);
    default: return launch_i8_kb_xt<KB, 1>(img, W, H, tpl, wsum, out, ow, oh, tw, th, method, s);
    }
}


// ---- fused int8 correlation + window sums (u8 images) -------------------------------
//
// One workgroup computes a (32 RT) x 32 output tile, wave w its rows
// 32w .. 32w + 31, from one LDS-staged image block (rows y0 .. y0 + 32 RT +
// th - 2, columns x0 .. x0 + 32 KB - 1, as i8 = u8 - 128) and the staged
// template:
//   * C' = sum I'T' on MFMA against the template's Toeplitz slices (as in
//     match_i8_mfma_kernel);
//   * S' = sum I' over each window on MFMA too, with the same A operands
//     against an all-ones Toeplitz band held in registers: exact in i32;
//   * sum I^2 by exact u32 sliding sums: each staged row's horizontal
//     tw-window sums of squares go to LDS (16 outputs per task), and each lane
//     slides the vertical th-window sum down its own output column, so the
//     sums land in the accumulators' layout without an exchange.
// Then sum I T = C' + 128 S' + 128 sum T and every method is finished from
// exact integers (finish_score). Nothing but the map leaves the workgroup.
constexpr int kFusedHP = 33;  // u32 per row of the horizontal sums (odd: column reads hit distinct banks)

__host__ __device__ constexpr int fused_pitch(int KB) {
    return ((KB * 32 / 16) % 2 == 0) ? KB * 32 + 16 : KB * 32;  // odd multiple of 16 bytes
}

// finish_score over 32-bit exact sums: for templates up to 128 x 128, sum I T
// and sum I^2, sum T^2 stay below 255^2 * 128^2 < 2^31, and SSE = sum T^2 -
// 2 sum I T + sum I^2 is exact in u32. F64: normalise in f64 (windows of at
// least 64 x 64 pixels), else in f32 like the reference.
template <int METHOD, bool F64>
__device__ __forceinline__ float finish_u32(uint32_t c, uint32_t sii, uint32_t stt) {
    const bool sse = (METHOD == IPH_SSE || METHOD == IPH_SSE_NORMED);
    const uint32_t num = sse ? stt + sii - 2u * c : c;
    if (METHOD == IPH_SSE || METHOD == IPH_CCORR) return (float)num;
    if constexpr (!F64) {
        const float norm = sqrtf((float)sii * (float)stt);
        const float sn = (float)num;
        return norm > 0.0f ? sn / norm : sn;
    } else {
        const double norm = sqrt((double)sii * (double)stt);
        return norm > 0.0 ? (float)((double)num / norm) : (float)num;
    }
}

// LDS of a fused workgroup that takes template rows at most tc at a time.
size_t fused_lds(int KB, int tw, int th, int RT, bool sums, int tc) {
    tc = std::min(th, tc);
    const int rows = 32 * RT + tc - 1;
    size_t b = align_up((size_t)rows * fused_pitch(KB), 16) + (size_t)tc * tpl_pitch(KB) + 16;
    if (sums) b = align_up(b, 16) + (size_t)rows * kFusedHP * sizeof(uint32_t);
    return b;
}

// The output tile at (x0, y0) of one match (the body of the single-image and
// the job-table kernels). Template rows are taken in ceil(th / tcm) chunks of
// near-equal height (at most tcm rows): per chunk the workgroup stages the
// image rows those template rows meet and the rows themselves, correlates,
// and adds the chunk's rows of every window's sum of squares, so the LDS
// footprint is bounded by tcm rather than by the template height.

template <int KB, int METHOD>
__device__ __forceinline__ void match_i8_fused_tile(const uint8_t *__restrict__ img, int W,
                                                    int H, const uint8_t *__restrict__ tpl,
                                                    float *__restrict__ out, int ow, int oh,
                                                    int tw, int th, int tcm, int x0,
                                                    int y0) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ long long tred[2][4];
    constexpr bool kSums = METHOD != IPH_CCORR;  // sum I^2 needed
    constexpr int BW = 32 * KB, PITCH = fused_pitch(KB), kTP = tpl_pitch(KB);
    const int RT = blockDim.x >> 6;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int n = lane & 31, g = lane >> 5;
    const int nchunk = (th + tcm - 1) / tcm;
    const int tc = (th + nchunk - 1) / nchunk;  // rows per chunk (the last may have fewer)
    const int rmax = 32 * RT + min(th, tcm) - 1;  // staged rows of a chunk, at most
    int8_t *blk = reinterpret_cast<int8_t *>(smem);
    const int tp_off = (rmax * PITCH + 15) & ~15;
    uint32_t *tp = reinterpret_cast<uint32_t *>(smem + tp_off);
    uint32_t *h2 = reinterpret_cast<uint32_t *>(
        smem + ((tp_off + min(th, tcm) * kTP + 16 + 15) & ~15));
    const bool aligned = (W & 3) == 0 && (reinterpret_cast<uintptr_t>(img) & 3) == 0 && W >= 4;
    const int nw = tw >> 2;
    const uint32_t tail = (tw & 3) ? (1u << (8 * (tw & 3))) - 1u : 0u;

    // all-ones band: B[k][n] = 1 for 0 <= 32 kb + k - n < tw
    v4i ones[KB];
#pragma unroll
    for (int kb = 0; kb < KB; kb++)
#pragma unroll
        for (int d = 0; d < 4; d++) {
            uint32_t wv = 0;
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const int dx = 32 * kb + 16 * g + 4 * d + b - n;
                if (dx >= 0 && dx < tw) wv |= 1u << (8 * b);
            }
            ones[kb][d] = (int)wv;
        }
    const int u0 = 32 + 16 * g - n;
    const uint32_t sh = (uint32_t)(u0 & 3);
    v16i acc, acs;
#pragma unroll
    for (int r = 0; r < 16; r++) acc[r] = acs[r] = 0;
    uint32_t sqv[16];
#pragma unroll
    for (int r = 0; r < 16; r++) sqv[r] = 0;
    uint32_t s1 = 0, s2 = 0;  // this thread's share of sum T, sum T^2

    for (int c0 = 0; c0 < th; c0 += tc) {
        const int dyn = min(tc, th - c0);
        const int rows = 32 * RT + dyn - 1;
        if (c0 > 0) __syncthreads();  // the previous chunk is done with blk, tp and h2
        // image rows y0 + c0.. (clamped) and columns x0.. (clamped), as i8; bytes
        // outside the image meet zero Toeplitz entries or discarded outputs
        const int total = rows * (BW / 4);
        for (int base = threadIdx.x; base < total; base += 8 * (int)blockDim.x) {
            uint32_t v[8];
#pragma unroll
            for (int b = 0; b < 8; b++) {
                const int i = min(base + b * (int)blockDim.x, total - 1);
                const int r = i / (BW / 4), c = (i - r * (BW / 4)) * 4;
                const uint8_t *row = img + (size_t)min(y0 + c0 + r, H - 1) * W;
                const int gx = x0 + c;
                if (aligned) {
                    v[b] = *reinterpret_cast<const uint32_t *>(row + min(gx, W - 4));
                } else {
                    v[b] = (uint32_t)row[min(gx, W - 1)] | ((uint32_t)row[min(gx + 1, W - 1)] << 8) |
                           ((uint32_t)row[min(gx + 2, W - 1)] << 16) |
                           ((uint32_t)row[min(gx + 3, W - 1)] << 24);
                }
            }
#pragma unroll
            for (int b = 0; b < 8; b++) {
                const int i = base + b * (int)blockDim.x;
                if (i < total) {
                    const int r = i / (BW / 4), c = (i - r * (BW / 4)) * 4;
                    *reinterpret_cast<uint32_t *>(blk + r * PITCH + c) = v[b] ^ 0x80808080u;
                }
            }
        }
        // template rows c0.. as i8 at [32, 32 + tw) of kTP-byte rows (zeros around)
        for (int i = threadIdx.x; i < dyn * (kTP / 4) + 4; i += blockDim.x) {
            const int r = i / (kTP / 4), c = 4 * (i - r * (kTP / 4)) - 32;
            uint32_t wv = 0;
#pragma unroll
            for (int b = 0; b < 4; b++)
                if (r < dyn && c + b >= 0 && c + b < tw) {
                    const uint32_t t = tpl[(c0 + r) * tw + c + b];
                    s1 += t;
                    s2 += t * t;
                    wv |= (t ^ 0x80u) << (8 * b);
                }
            tp[i] = wv;
        }
        __syncthreads();

        // horizontal tw-window sums of squares of the chunk's rows, 16 outputs
        // per task: columns 16q .. 16q + 15 (the tile's 32 output columns)
        if constexpr (kSums) {
            for (int it = threadIdx.x; it < rows * 2; it += blockDim.x) {
                const int r = it >> 1, q = it & 1;
                const uint32_t *pw = reinterpret_cast<const uint32_t *>(blk + r * PITCH + 16 * q);
                uint32_t b = 0;
                for (int k = 0; k < nw; k++) {
                    const uint32_t wv = pw[k] ^ 0x80808080u;
                    b = __builtin_amdgcn_udot4(wv, wv, b, false);
                }
                if (tail) {
                    const uint32_t wv = (pw[nw] ^ 0x80808080u) & tail;
                    b = __builtin_amdgcn_udot4(wv, wv, b, false);
                }
                const uint8_t *p = reinterpret_cast<const uint8_t *>(pw);
                uint32_t *o = h2 + r * kFusedHP + 16 * q;
                o[0] = b;
#pragma unroll
                for (int x = 1; x < 16; x++) {
                    const uint32_t vin = p[x + tw - 1] ^ 0x80u, vout = p[x - 1] ^ 0x80u;
                    b += vin * vin - vout * vout;
                    o[x] = b;
                }
            }
        }

        // correlation and window sums of I' on MFMA over the chunk's template rows
        const int8_t *arow = blk + (wave * 32 + n) * PITCH + 16 * g;
        const uint32_t *trow = tp + (u0 >> 2);
        for (int dy = 0; dy < dyn; dy++) {
#pragma unroll
            for (int kb = 0; kb < KB; kb++) {
                const uint32_t *q = trow + 8 * kb;
                const uint32_t d0 = q[0], d1 = q[1], d2 = q[2], d3 = q[3], d4 = q[4];
                v4i bt;
                bt[0] = (int)__builtin_amdgcn_alignbyte(d1, d0, sh);
                bt[1] = (int)__builtin_amdgcn_alignbyte(d2, d1, sh);
                bt[2] = (int)__builtin_amdgcn_alignbyte(d3, d2, sh);
                bt[3] = (int)__builtin_amdgcn_alignbyte(d4, d3, sh);
                const v4i a = *reinterpret_cast<const v4i *>(arow + 32 * kb);
                acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, bt, acc, 0, 0, 0);
                acs = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, ones[kb], acs, 0, 0, 0);
            }
            arow += PITCH;
            trow += kTP / 4;
        }

        if constexpr (kSums) {
            __syncthreads();  // every row's horizontal sums are in h2
            // lane (n, g) holds rows 32w + 8j + 4g + i (register 4j + i) of column
            // n: slide the chunk's dyn-row sum of h2 down rows 32w + 4g .. + 27
            const uint32_t *hc = h2 + (wave * 32 + 4 * g) * kFusedHP + n;
            uint32_t sq = 0;
            for (int d = 0; d < dyn; d++) sq += hc[d * kFusedHP];
#pragma unroll
            for (int j = 0; j < 4; j++) {
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int yl = 8 * j + i;  // row offset from 32w + 4g
                    sqv[4 * j + i] += sq;
                    // to row yl + 1 (past the other half's four rows after i = 3)
                    if (j < 3 || i < 3) {
#pragma unroll
                        for (int k = 0; k < ((i == 3) ? 5 : 1); k++) {
                            const int r = yl + k;
                            sq += hc[(r + dyn) * kFusedHP] - hc[r * kFusedHP];
                        }
                    }
                }
            }
        }
    }

    long long l1 = s1, l2 = s2;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        l1 += __shfl_xor(l1, o);
        l2 += __shfl_xor(l2, o);
    }
    if (lane == 0) {
        tred[0][wave] = l1;
        tred[1][wave] = l2;
    }
    __syncthreads();
    uint32_t st = 0, stt = 0;
    for (int w = 0; w < RT; w++) {
        st += (uint32_t)tred[0][w];
        stt += (uint32_t)tred[1][w];
    }
    const int x = x0 + n;
    const int ybase = y0 + wave * 32 + 4 * g;
    auto emit = [&](auto f64) {
        constexpr bool F64 = decltype(f64)::value;
        float v[16];
#pragma unroll
        for (int r = 0; r < 16; r++) {
            // sum I T = C' + 128 (S' + sum T), exact in 32 bits
            const uint32_t c = (uint32_t)(acc[r] + 128 * (acs[r] + (int)st));
            v[r] = finish_u32<METHOD, F64>(c, sqv[r], stt);
        }
        if (x < ow) {
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int y = ybase + 8 * (r >> 2) + (r & 3);
                if (y < oh) out[(uint32_t)y * (uint32_t)ow + (uint32_t)x] = v[r];
            }
        }
    };
    const bool f64 = (METHOD == IPH_SSE_NORMED || METHOD == IPH_CCORR_NORMED) && tw * th >= 64 * 64;
    if (f64) emit(std::true_type{});
    else emit(std::false_type{});
}

template <int KB, int METHOD>
__global__ __launch_bounds__(256) void match_i8_fused_kernel(
    const uint8_t *__restrict__ img, int W, int H, const uint8_t *__restrict__ tpl,
    float *__restrict__ out, int ow, int oh, int tw, int th, int tcm) {
    match_i8_fused_tile<KB, METHOD>(img, W, H, tpl, out, ow, oh, tw, th, tcm, blockIdx.x * 32,
                                    blockIdx.y * 32 * (int)(blockDim.x >> 6));
}

// The same over a device table of matches (one KB and method): job j owns
// tiles [tile0, tile0 + tiles_x * tiles_y) of the launch's tile list, which
// is dealt to the XCDs in eight contiguous ranges (a job's neighbouring tiles,
// which share image rows, meet in one L2).
template <int KB, int METHOD>
__global__ __launch_bounds__(256) void match_i8_fused_jobs_kernel(const MatchJob *__restrict__ jobs,
                                                                  int njobs, int ntiles,
                                                                  int tcm) {
    const int chunk = (ntiles + 7) >> 3;
    const int t = (int)(blockIdx.x & 7) * chunk + (int)(blockIdx.x >> 3);
    if ((int)(blockIdx.x >> 3) >= chunk || t >= ntiles) return;
    int lo = 0, hi = njobs - 1;  // the last job with tile0 <= t
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].tile0 <= t) lo = mid;
        else hi = mid - 1;
    }
    const MatchJob &jb = jobs[lo];
    const int lt = t - jb.tile0, ty = lt / jb.tiles_x, tx = lt - ty * jb.tiles_x;
    match_i8_fused_tile<KB, METHOD>(jb.img, jb.W, jb.H, jb.tpl, jb.out, jb.ow, jb.oh, jb.tw,
                                    jb.th, tcm, tx * 32, ty * 32 * (int)(blockDim.x >> 6));
}

constexpr size_t kFusedMaxLds = 64 * 1024;
// Footprint that still lets three 4-wave workgroups share a CU's 160 KiB.
constexpr size_t kFusedOccLds = 53 * 1024;
constexpr int kFusedMinChunk = 32;  // chunks are never forced below this

// Output rows per fused workgroup (32 * RT): the batch hint or 4 waves,
// halved while the footprint of 32-row chunks is above kFusedMaxLds; 0 when
// even one wave does not fit.
int fused_rt(int tw, int th, int method) {
    const int KB = (31 + tw + 31) / 32;
    int RT = g_match_rows_hint > 0 ? std::min(g_match_rows_hint, 4) : 4;
    while (RT > 0 &&
           fused_lds(KB, tw, th, RT, method != IPH_CCORR, kFusedMinChunk) > kFusedMaxLds)
        RT /= 2;
    return RT;
}

// Largest chunk height (near-equal chunks of th) whose footprint is at most
// kFusedOccLds, else kFusedMinChunk: every chunk restages the image rows it
// meets, so fewer chunks save staging and horizontal sums.
int fused_tc(int KB, int tw, int th, int RT, bool sums) {
    for (int nc = 1; nc <= th; nc++) {
        const int tc = (th + nc - 1) / nc;
        if (tc < kFusedMinChunk && nc > 1) break;
        if (fused_lds(KB, tw, th, RT, sums, tc) <= kFusedOccLds) return tc;
    }
    return std::min(th, kFusedMinChunk);
}

template <int KB>
iph_status launch_i8_fused_kb(const uint8_t *img, int W, int H, const uint8_t *tpl, float *out,
                              int ow, int oh, int tw, int th, int method, int RT,
                              hipStream_t s) {
    const bool sums = method != IPH_CCORR;
    const int tc = fused_tc(KB, tw, th, RT, sums);
    const size_t lds = fused_lds(KB, tw, th, RT, sums, tc);
    dim3 grid((ow + 31) / 32, (oh + 32 * RT - 1) / (32 * RT));
#define L(M)                                                                                 \
    match_i8_fused_kernel<KB, M><<<grid, 64 * RT, lds, s>>>(img, W, H, tpl, out, ow, oh, tw, th, tc)
    switch (method) {
    case IPH_SSE: L(IPH_SSE); break;
    case IPH_SSE_NORMED: L(IPH_SSE_NORMED); break;
    case IPH_CCORR: L(IPH_CCORR); break;
    default: L(IPH_CCORR_NORMED); break;
    }
#undef L
    return launch_status("match_i8_fused_kernel");
}


}  // namespace

iph_status launch_bf16x3_match(const float *img, uint32_t w, uint32_t h, const float *tpl,
                               uint32_t tw, uint32_t th, int32_t method, float *out, void *ws,
                               size_t ws_bytes, hipStream_t s) {
    if (tw > (uint32_t)kMaxSideF32 || th > (uint32_t)kMaxSideF32) {
        set_error("the BF16x3 MFMA path takes templates up to %dx%d; use the direct path",
                  kMaxSideF32, kMaxSideF32);
        return IPH_EUNSUPPORTED;
    }
    const int KB = blocks16_for((int)tw);
    Workspace wsp(ws, ws_bytes);
    double *tsq = wsp.take<double>(1);
    uint16_t *bsw = wsp.take<uint16_t>((size_t)th * KB * 2 * 64 * 8);
    double *wsq = wsp.take<double>((size_t)(w - tw + 1) * (h - th + 1));
    double *hsq = method == IPH_CCORR ? nullptr : wsp.take<double>((size_t)(w - tw + 1) * h);
    if (!tsq || !bsw || !wsq || (method != IPH_CCORR && !hsq)) {
        set_error("workspace too small for the BF16x3 MFMA match_template");
        return IPH_EINVAL;
    }
    int ow = (int)(w - tw + 1), oh = (int)(h - th + 1);
    template_sq_f32_kernel<<<1, 256, 0, s>>>(tpl, (int)(tw * th), tsq);
    IPH_TRY(launch_status("template_sq_f32_kernel"));
    toeplitz_bf16_kernel<<<dim3(KB, th), 64, 0, s>>>(tpl, (int)tw, KB, bsw);
    IPH_TRY(launch_status("toeplitz_bf16_kernel"));
    if (method != IPH_CCORR) {
        dim3 hg((ow + kHsOut - 1) / kHsOut, h);
        size_t lds = (size_t)(kHsOut + tw - 1) * sizeof(double);
        window_sq_rows_f32_kernel<<<hg, 256, lds, s>>>(img, (int)w, (int)tw, ow, hsq);
        IPH_TRY(launch_status("window_sq_rows_f32_kernel"));
        dim3 vg((ow + kVsCols - 1) / kVsCols, (oh + kVsRows - 1) / kVsRows);
        window_sq_cols_f32_kernel<<<vg, 256, 0, s>>>(hsq, (int)th, ow, oh, wsq);
        IPH_TRY(launch_status("window_sq_cols_f32_kernel"));
    }
    switch (KB) {
#define C(K)                                                                                 \
    case K:                                                                                  \
        return launch_bf16_kb<K>(img, w, h, bsw, wsq, tsq, out, ow, oh, tw, th, method, s);
        C(2) C(3) C(4) C(5) C(6) C(7) C(8) C(9) C(10)
#undef C
    default:
        set_error("unsupported template width %u", tw);
        return IPH_EUNSUPPORTED;
    }
}

size_t mfma_match_workspace(int32_t p_type, uint32_t w, uint32_t h, uint32_t tw, uint32_t th,
                            int32_t method) {
    if (tw == 0 || th == 0 || w < tw || h < th) return 256;
    if (p_type == IPH_F32) {
        WorkspacePlan wp;
        wp.take<double>(1);
        wp.take<uint16_t>((size_t)th * blocks16_for((int)tw) * 2 * 64 * 8);
        wp.take<double>((size_t)(w - tw + 1) * (h - th + 1));
        if (method != IPH_CCORR) wp.take<double>((size_t)(w - tw + 1) * h);
        return wp.used;
    }
    if (tw <= (uint32_t)kMaxSide && th <= (uint32_t)kMaxSide && fused_rt((int)tw, (int)th, method) > 0)
        return 256;  // the fused kernel needs no workspace
    WorkspacePlan wp;
    wp.take<uint2>((size_t)(w - tw + 1) * (h - th + 1));
    return wp.used;
}

iph_status launch_mfma_match(const void *img, int32_t p_type, uint32_t w, uint32_t h,
                             const void *tpl, uint32_t tw, uint32_t th, int32_t method,
                             float *out, void *ws, size_t ws_bytes, hipStream_t s) {
    if (p_type == IPH_F32)
        return launch_bf16x3_match(static_cast<const float *>(img), w, h,
                                   static_cast<const float *>(tpl), tw, th, method, out, ws,
                                   ws_bytes, s);
    if (tw > (uint32_t)kMaxSide || th > (uint32_t)kMaxSide) {
        set_error("the int8 MFMA path takes templates up to %dx%d; use the direct path",
                  kMaxSide, kMaxSide);
        return IPH_EUNSUPPORTED;
    }
    const int KB = blocks_for((int)tw);
    const uint8_t *im8 = static_cast<const uint8_t *>(img);
    const uint8_t *tp8 = static_cast<const uint8_t *>(tpl);
    const int RT = fused_rt((int)tw, (int)th, method);
    if (RT > 0) {
        const int ow = (int)(w - tw + 1), oh = (int)(h - th + 1);
        switch (KB) {
        case 1: return launch_i8_fused_kb<1>(im8, w, h, tp8, out, ow, oh, tw, th, method, RT, s);
        case 2: return launch_i8_fused_kb<2>(im8, w, h, tp8, out, ow, oh, tw, th, method, RT, s);
        case 3: return launch_i8_fused_kb<3>(im8, w, h, tp8, out, ow, oh, tw, th, method, RT, s);
        case 4: return launch_i8_fused_kb<4>(im8, w, h, tp8, out, ow, oh, tw, th, method, RT, s);
        default: return launch_i8_fused_kb<5>(im8, w, h, tp8, out, ow, oh, tw, th, method, RT, s);
        }
    }
    Workspace wsp(ws, ws_bytes);
    uint2 *wsum = wsp.take<uint2>((size_t)(w - tw + 1) * (h - th + 1));
    if (!wsum) {
        set_error("workspace too small for the int8 MFMA match_template");
        return IPH_EINVAL;
    }
    const uint8_t *im = static_cast<const uint8_t *>(img);
    const uint8_t *tp = static_cast<const uint8_t *>(tpl);
    int ow = (int)(w - tw + 1), oh = (int)(h - th + 1);
    dim3 wg((ow + kWinCols - 1) / kWinCols, (oh + kWinRows - 1) / kWinRows);
    window_sums_u8_kernel<<<wg, kWinThreads, win_lds((int)tw, (int)th), s>>>(im, (int)w, (int)tw,
                                                                    (int)th, ow, oh, wsum);
    IPH_TRY(launch_status("window_sums_u8_kernel"));
    switch (KB) {
    case 1: return launch_i8_kb<1>(im, w, h, tp, wsum, out, ow, oh, tw, th, method, s);
    case 2: return launch_i8_kb<2>(im, w, h, tp, wsum, out, ow, oh, tw, th, method, s);
    case 3: return launch_i8_kb<3>(im, w, h, tp, wsum, out, ow, oh, tw, th, method, s);
    case 4: return launch_i8_kb<4>(im, w, h, tp, wsum, out, ow, oh, tw, th, method, s);
    default: return launch_i8_kb<5>(im, w, h, tp, wsum, out, ow, oh, tw, th, method, s);
    }
}

void set_match_rows_hint(int rows) { g_match_rows_hint = rows; }

int fused_match_kb(uint32_t tw) { return blocks_for((int)tw); }

int fused_match_rt_max(uint32_t tw, uint32_t th, int32_t method) {
    if (tw == 0 || th == 0 || tw > (uint32_t)kMaxSide || th > (uint32_t)kMaxSide) return 0;
    const int KB = blocks_for((int)tw);
    for (int rt = 4; rt > 0; rt /= 2)
        if (fused_lds(KB, (int)tw, (int)th, rt, method != IPH_CCORR, kFusedMinChunk) <=
            kFusedMaxLds)
            return rt;
    return 0;
}

int fused_match_tc(int kb, uint32_t th, int rt, int32_t method) {
    return fused_tc(kb, 0, (int)th, rt, method != IPH_CCORR);
}

size_t fused_match_lds(int kb, uint32_t tw, uint32_t th, int rt, int32_t method, int tc) {
    return fused_lds(kb, (int)tw, (int)th, rt, method != IPH_CCORR, tc);
}

long long match_jobs_layout(MatchJob *jobs, int n, int rt) {
    long long t = 0;
    for (int i = 0; i < n; i++) {
        MatchJob &j = jobs[i];
        j.tiles_x = (j.ow + 31) / 32;
        j.tile0 = (int)t;
        t += (long long)j.tiles_x * ((j.oh + 32 * rt - 1) / (32 * rt));
    }
    return t;
}

iph_status launch_match_jobs(const MatchJob *jobs, int n, long long ntiles, int kb,
                             int32_t method, int rt, int tc, size_t lds, hipStream_t s) {
    if (n <= 0 || ntiles <= 0) return IPH_OK;
    if (ntiles > (1LL << 30) || kb < 1 || kb > 5 || (rt != 1 && rt != 2 && rt != 4) || tc < 1 ||
        lds > kFusedMaxLds) {
        set_error("match job table out of range (%lld tiles, kb %d, rt %d, chunk %d, lds %zu)",
                  ntiles, kb, rt, tc, lds);
        return IPH_EINVAL;
    }
    const unsigned grid = (unsigned)((ntiles + 7) / 8 * 8);
#define J(K, M)                                                                              \
    match_i8_fused_jobs_kernel<K, M><<<grid, 64 * rt, lds, s>>>(jobs, n, (int)ntiles, tc)
#define JK(K)                                                                                \
    switch (method) {                                                                        \
    case IPH_SSE: J(K, IPH_SSE); break;                                                      \
    case IPH_SSE_NORMED: J(K, IPH_SSE_NORMED); break;                                        \
    case IPH_CCORR: J(K, IPH_CCORR); break;                                                  \
    default: J(K, IPH_CCORR_NORMED); break;                                                  \
    }                                                                                        \
    break
    switch (kb) {
    case 1: JK(1);
    case 2: JK(2);
    case 3: JK(3);
    case 4: JK(4);
    default: JK(5);
    }
#undef JK
#undef J
    return launch_status("match_i8_fused_jobs_kernel");
}

}  // namespace iph
