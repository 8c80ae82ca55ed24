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
// sums sum I and sum I^2 come from the same workgroup (an all-ones Toeplitz
// band on MFMA, and exact u32 sliding sums), so
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
// rows are processed in chunks, restaging the image rows of each chunk and
// the chunk's template rows in LDS (the Toeplitz slices are cut from the
// staged template rows), so any template height fits. The window energy
// sum I^2 is summed in f64 by one separate kernel.
#include "match_common.hpp"

#include <algorithm>
#include <type_traits>

namespace iph {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int kMaxSide = (int)kMfmaMaxSide;  // largest template side on the int8 path
// Bytes per template row staged in LDS: 32 zeros, the row, zeros up to
// 32 * (KB + 1) (every Toeplitz slice of a lane reads inside it).
__host__ __device__ constexpr int tpl_pitch(int KB) { return 32 * KB + 32; }

// ---- BF16x3 path for f32 images --------------------------------------------------

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));


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

// The same split for two values on the conversion unit (v_cvt_pk_bf16_f32
// rounds to nearest even like bf16_bits); x - hi is exact in f32. Returns
// the packed hi pair and writes the packed lo pair.
__device__ __forceinline__ uint32_t split2_bf16(float x0, float x1, uint32_t &lo) {
    typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
    const bf2 h = {(__bf16)x0, (__bf16)x1};
    const uint32_t hu = __builtin_bit_cast(uint32_t, h);
    const bf2 l = {(__bf16)(x0 - __uint_as_float(hu << 16)),
                   (__bf16)(x1 - __uint_as_float(hu & 0xffff0000u))};
    lo = __builtin_bit_cast(uint32_t, l);
    return hu;
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

// Window energies sum(i^2) of an f32 image in f64, one kernel: a workgroup
// owns kWsqCols output columns by kWsqRows output rows. Each thread holds the
// vertical sums over th rows of p^2 for its image columns (summed directly
// once, then slid down a row at a time: the row entering added, the row
// leaving subtracted, all f64); per output row they go to LDS, and each
// thread sums the horizontal windows of kWsqRun consecutive outputs (the
// first directly, the rest slid). p^2 is exact in f64; only the f64
// energies are written.
constexpr int kWsqRun = 4;                    // outputs per thread per row
constexpr int kWsqCols = 256 * kWsqRun;       // output columns per workgroup
constexpr int kWsqRows = 16;                  // output rows per workgroup
constexpr int kWsqOwn = (kWsqCols + 128 + 255) / 256;  // image columns per thread (tw <= 129)
__global__ __launch_bounds__(256) void window_sq_f32_kernel(const float *__restrict__ img, int W,
                                                            int tw, int th, int ow, int oh,
                                                            double *__restrict__ wsq) {
    extern __shared__ __align__(16) double usm[];  // kWsqCols + tw - 1
    const int x0 = blockIdx.x * kWsqCols, y0 = blockIdx.y * kWsqRows;
    const int ncols = min(kWsqCols + tw - 1, W - x0);
    const int yend = min(y0 + kWsqRows, oh);
    const float *base = img + x0;
    double u[kWsqOwn];
#pragma unroll
    for (int k = 0; k < kWsqOwn; k++) {
        const int c = (int)threadIdx.x + 256 * k;
        u[k] = 0.0;
        if (c < ncols) {
            // eight rows' loads in flight, summed in row order
            const float *col = base + (size_t)y0 * W + c;
            int dy = 0;
            for (; dy + 8 <= th; dy += 8) {
                float v[8];
#pragma unroll
                for (int q = 0; q < 8; q++) v[q] = col[(size_t)(dy + q) * W];
#pragma unroll
                for (int q = 0; q < 8; q++) u[k] += (double)v[q] * (double)v[q];
            }
            for (; dy < th; dy++) {
                const double p = col[(size_t)dy * W];
                u[k] += p * p;
            }
        }
    }
    for (int y = y0; y < yend; y++) {
#pragma unroll
        for (int k = 0; k < kWsqOwn; k++) {
            const int c = (int)threadIdx.x + 256 * k;
            if (c < ncols) usm[c] = u[k];
        }
        __syncthreads();
        const int xl = (int)threadIdx.x * kWsqRun;
        if (x0 + xl < ow) {
            double a = 0.0;
            for (int d = 0; d < tw; d++) a += usm[xl + d];
            double *o = wsq + (size_t)y * ow + x0 + xl;
            o[0] = a;
#pragma unroll
            for (int q = 1; q < kWsqRun; q++) {
                if (x0 + xl + q >= ow) break;
                a += usm[xl + q - 1 + tw] - usm[xl + q - 1];
                o[q] = a;
            }
        }
        __syncthreads();
        if (y + 1 < yend) {
#pragma unroll
            for (int k = 0; k < kWsqOwn; k++) {
                const int c = (int)threadIdx.x + 256 * k;
                if (c < ncols) {
                    const double pin = base[(size_t)(y + th) * W + c];
                    const double pout = base[(size_t)y * W + c];
                    u[k] += pin * pin - pout * pout;
                }
            }
        }
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

// BF16x3 correlation. A workgroup of kBfWaves waves side by side computes 32
// output rows by kBfWaves * 64 columns (each wave two 32x32 tiles). Template
// rows are taken in chunks of C. The image rows live in a ring of R >= 32 +
// C - 1 rows of bf16 hi and lo planes (image row r in slot r mod R):
// the first chunk stages all the rows it meets, every later chunk only its C
// new ones, over the slots of the rows the previous chunk was the last to
// use. Per chunk the template rows are staged as bf16 hi and lo rows padded
// with 32 zeros on the left (and zeros to the right). Per template row dy and K block kb, each lane cuts its
// Toeplitz slice B[k][n] = T[dy][16 kb + k - n], k = 8 g + 0..7, from the
// padded row: 8 consecutive bf16 starting at element 32 + 16 kb + 8 g - n,
// read as 5 dwords and realigned by v_alignbyte (the shift, 0 or 2 bytes, is
// the lane's column parity). Each slice pair feeds both tiles (6 MFMAs).
constexpr int kBfWaves = 4;
constexpr int kBfWaveCols = 64;
constexpr int kBfTiles = kBfWaveCols / 32;  // 32x32 output tiles per wave
constexpr int kBfThreads = 64 * kBfWaves;

__host__ __device__ constexpr int bf_img_cols(int KB) { return kBfWaves * kBfWaveCols + 16 * (KB - 2); }
// staged image row pitch in bf16: an odd multiple of 8 (16 bytes), so the 16-byte
// A-operand reads of 8 consecutive rows fall in distinct bank groups
__host__ __device__ constexpr int bf_img_pitch(int KB) {
    return ((bf_img_cols(KB) / 8) % 2 == 0) ? bf_img_cols(KB) + 8 : bf_img_cols(KB);
}
__host__ __device__ constexpr int bf_tpl_pitch(int KB) { return 16 * KB + 64; }
// ring rows for chunks of C template rows: at least 32 + C - 1, a multiple
// of 8 so that the 8 rows one A-operand read meets keep distinct bank groups
// across the wrap
__host__ __device__ constexpr int bf_ring_rows(int C) { return (32 + C - 1 + 7) & ~7; }
__host__ __device__ constexpr size_t bf_lds(int KB, int C) {
    return (size_t)bf_ring_rows(C) * bf_img_pitch(KB) * 4 + (size_t)C * bf_tpl_pitch(KB) * 4;
}

// 8 consecutive bf16 of a padded template row from element e (e >= 0):
// 5 dwords from e / 2, realigned by (e & 1) * 2 bytes.
__device__ __forceinline__ v8bf bf_slice(const uint16_t *row, int e) {
    const uint32_t *d = reinterpret_cast<const uint32_t *>(row) + (e >> 1);
    const uint32_t sh = (uint32_t)(e & 1) * 2u;
    uint32_t w[5];
#pragma unroll
    for (int i = 0; i < 5; i++) w[i] = d[i];
    typedef unsigned int v4u __attribute__((ext_vector_type(4)));
    v4u r;
#pragma unroll
    for (int i = 0; i < 4; i++) r[i] = __builtin_amdgcn_alignbyte(w[i + 1], w[i], sh);
    return __builtin_bit_cast(v8bf, r);
}

template <int KB, int METHOD>
__global__ __launch_bounds__(kBfThreads) void match_bf16x3_mfma_kernel(
    const float *__restrict__ img, int W, int H, const float *__restrict__ tpl,
    const double *__restrict__ wsq, const double *__restrict__ tsq, float *__restrict__ out,
    int ow, int oh, int tw, int th, int C) {
    constexpr int BW = bf_img_cols(KB), PI = bf_img_pitch(KB), PT = bf_tpl_pitch(KB);
    extern __shared__ __align__(16) unsigned char smem[];
    const int R = bf_ring_rows(C);
    uint16_t *ih = reinterpret_cast<uint16_t *>(smem);
    uint16_t *il = ih + (size_t)R * PI;
    uint16_t *th_s = il + (size_t)R * PI;
    uint16_t *tl_s = th_s + (size_t)C * PT;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int n = lane & 31, g = lane >> 5;
    const int x0 = blockIdx.x * (kBfWaves * kBfWaveCols), y0 = blockIdx.y * 32;
    const bool vec = (W & 3) == 0;

    v16f acc[kBfTiles];
#pragma unroll
    for (int t = 0; t < kBfTiles; t++)
#pragma unroll
        for (int r = 0; r < 16; r++) acc[t][r] = 0.0f;

    int base = 0;  // ring slot of image row c0 (relative to y0)
    for (int c0 = 0; c0 < th; c0 += C) {
        const int dend = min(C, th - c0);
        __syncthreads();  // the previous chunk's operands are consumed
        // image rows y0 + [ra, c0 + dend + 31), columns x0 + [0, BW), 4 per
        // thread: all of them for the first chunk, else the rows past the
        // previous chunk's, into slots (ra mod R) + q, wrapped
        const int ra = c0 == 0 ? 0 : c0 + 31;
        const int nr = c0 + dend + 31 - ra;
        const int sa = c0 == 0 ? 0 : (base + 31 < R ? base + 31 : base + 31 - R);
        for (int i = threadIdx.x; i < nr * (BW / 4); i += kBfThreads) {
            const int q = i / (BW / 4), c = 4 * (i - q * (BW / 4));
            const int gy = min(y0 + ra + q, H - 1), gx = x0 + c;
            const int slot = sa + q < R ? sa + q : sa + q - R;
            const float *src = img + (size_t)gy * W;
            float v[4];
            if (vec && gx + 3 < W) {
                const float4 f = *reinterpret_cast<const float4 *>(src + gx);
                v[0] = f.x; v[1] = f.y; v[2] = f.z; v[3] = f.w;
            } else {
#pragma unroll
                for (int b = 0; b < 4; b++) v[b] = src[min(gx + b, W - 1)];
            }
            uint32_t l0, l1;
            const uint32_t h0 = split2_bf16(v[0], v[1], l0), h1 = split2_bf16(v[2], v[3], l1);
            *reinterpret_cast<uint2 *>(ih + slot * PI + c) = make_uint2(h0, h1);
            *reinterpret_cast<uint2 *>(il + slot * PI + c) = make_uint2(l0, l1);
        }
        // template rows c0 + [0, C): 32 zeros, the row, zeros to PT
        for (int i = threadIdx.x; i < C * PT; i += kBfThreads) {
            const int r = i / PT, e = i - r * PT, dx = e - 32;
            uint16_t h = 0, l = 0;
            if (c0 + r < th && dx >= 0 && dx < tw) split_bf16(tpl[(size_t)(c0 + r) * tw + dx], h, l);
            th_s[i] = h;
            tl_s[i] = l;
        }
        __syncthreads();
        const int e0 = 32 + 8 * g - n;
        const int s0 = base + n;  // ring slot of this lane's row at d = 0, before wrapping (< 2R)
        for (int d = 0; d < dend; d++) {
            const int sl = s0 + d < R ? s0 + d : s0 + d - R;
            const uint16_t *ah = ih + sl * PI + kBfWaveCols * wave + 8 * g;
            const uint16_t *al = il + sl * PI + kBfWaveCols * wave + 8 * g;
            v8bf bh[KB], bl[KB];
#pragma unroll
            for (int kb = 0; kb < KB; kb++) {
                bh[kb] = bf_slice(th_s + d * PT, e0 + 16 * kb);
                bl[kb] = bf_slice(tl_s + d * PT, e0 + 16 * kb);
            }
#pragma unroll
            for (int sblk = 0; sblk < KB + 2 * (kBfTiles - 1); sblk++) {
                const v8bf a_h = *reinterpret_cast<const v8bf *>(ah + 16 * sblk);
                const v8bf a_l = *reinterpret_cast<const v8bf *>(al + 16 * sblk);
#pragma unroll
                for (int t = 0; t < kBfTiles; t++) {
                    const int kb = sblk - 2 * t;
                    if (kb >= 0 && kb < KB) {
                        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a_l, bh[kb], acc[t], 0, 0, 0);
                        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a_h, bl[kb], acc[t], 0, 0, 0);
                        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a_h, bh[kb], acc[t], 0, 0, 0);
                    }
                }
            }
        }
        base = base + C < R ? base + C : base + C - R;
    }

    const double stt = METHOD == IPH_CCORR ? 0.0 : tsq[0];  // CCORR needs no sum T^2
    const long long nn = (long long)tw * th;
#pragma unroll
    for (int t = 0; t < kBfTiles; t++) {
        const int x = x0 + kBfWaveCols * wave + 32 * t + n;
        double sq[16];
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int y = y0 + 8 * (r >> 2) + 4 * g + (r & 3);
            sq[r] = (METHOD == IPH_CCORR) ? 0.0 : wsq[(size_t)min(y, oh - 1) * ow + min(x, ow - 1)];
        }
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int y = y0 + 8 * (r >> 2) + 4 * g + (r & 3);
            float v = finish_score_f32<METHOD>(acc[t][r], sq[r], stt, nn);
            if (x < ow && y < oh) out[(size_t)y * ow + x] = v;
        }
    }
}

// Template rows per chunk: the largest of 32, 16, 8 whose LDS lets two
// workgroups share a CU.
int bf_chunk(int KB) {
    for (int C = 32; C >= 8; C /= 2)
        if (bf_lds(KB, C) <= 80 * 1024) return C;
    return 4;
}

template <int KB>
iph_status launch_bf16_kb(const float *img, int W, int H, const float *tpl, const double *wsq,
                          const double *tsq, float *out, int ow, int oh, int tw, int th,
                          int method, hipStream_t s) {
    const int C = bf_chunk(KB);
    const size_t lds = bf_lds(KB, C);
    dim3 grid((ow + kBfWaves * kBfWaveCols - 1) / (kBfWaves * kBfWaveCols), (oh + 31) / 32);
#define L(M)                                                                                 \
    match_bf16x3_mfma_kernel<KB, M><<<grid, kBfThreads, lds, s>>>(img, W, H, tpl, wsq, tsq, \
                                                                      out, ow, oh, tw, th, C)
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

// ---- fused int8 correlation + window sums (u8 images) -------------------------------
//
// One workgroup computes a (32 RT) x 32 output tile, wave w its rows
// 32w .. 32w + 31, from one LDS-staged image block (rows y0 .. y0 + 32 RT +
// th - 2, columns x0 .. x0 + 32 KB - 1, as i8 = u8 - 128) and the staged
// template:
//   * C' = sum I'T' on MFMA against the template's Toeplitz slices, which
//     each lane cuts from the LDS-staged template rows (zero padded) with
//     five dword reads and v_alignbyte: B[k][n] = T'[dy][32 kb + k - n];
//   * S' = sum I' over each window: every staged row's horizontal tw-window
//     sums on MFMA (A = 32 staged rows, B = an all-ones Toeplitz band held
//     in registers), written over the image block as i16 once the
//     correlation is done, then slid vertically like sum I^2 (for KB = 1,
//     whose image block is too narrow, the band MFMA runs per template row);
//   * sum I^2 by exact u32 sliding sums: each staged row's horizontal
//     tw-window sums of squares go to LDS (16 outputs per task), and each lane
//     slides the vertical th-window sum down its own output column, so the
//     sums land in the accumulators' layout without an exchange.
// Then sum I T = C' + 128 S' + 128 sum T and every method is finished from
// exact integers (finish_score). Nothing but the map leaves the workgroup.
constexpr int kFusedHP = 33;  // u32 per row of the horizontal sums (odd: column reads hit distinct banks)
// i16 per row of the horizontal sums of I' (rows 4 apart 16 banks apart); 80
// bytes per row fit the image block's row pitch for KB >= 2.
constexpr int kH1P = 40;

__host__ __device__ constexpr int fused_pitch(int KB) {
    return ((KB * 32 / 16) % 2 == 0) ? KB * 32 + 16 : KB * 32;  // odd multiple of 16 bytes
}
static_assert(fused_pitch(2) >= 2 * kH1P, "the i16 row sums fit over the image block");

// LDS of a fused workgroup that takes template rows at most tc at a time.
size_t fused_lds(int KB, int tw, int th, int RT, bool sums, int tc) {
    tc = std::min(th, tc);
    const int rows = 32 * RT + tc - 1;
    // image rows, template rows + the all-ones row + 16 zero bytes
    size_t b = align_up((size_t)rows * fused_pitch(KB), 16) + (size_t)(tc + 1) * tpl_pitch(KB) + 16;
    if (sums) b = align_up(b, 16) + (size_t)rows * kFusedHP * sizeof(uint32_t);
    return b;
}

// Raw buffer over `bytes` bytes from `base`: dword loads past the end read
// zero and stores past it are dropped (the gfx9 descriptor word 3 for raw
// 32-bit access).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t raw_rsrc(const void *base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, (int)bytes, 0x00020000);
}

// sqrt(x) correctly rounded, for x = 0 or normal x: v_sqrt_f32 is within
// 1 ulp, and the sign of x - s' s for the neighbours s' of s picks the
// rounded root (Tuckerman's test; x and those products are multiples of the
// squared ulp, so the test is exact).
__device__ __forceinline__ float sqrt_rn(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sm = __uint_as_float(__float_as_uint(s) - 1u);
    const float sp = __uint_as_float(__float_as_uint(s) + 1u);
    const float rm = __builtin_fmaf(-sm, s, x), rp = __builtin_fmaf(-sp, s, x);
    const float t = rm <= 0.0f ? sm : s;
    return rp > 0.0f ? sp : t;
}

// a / b correctly rounded for 0 <= a < 2^32 and 1 <= b < 2^32, where no
// intermediate leaves the normal range: the reciprocal refined once by
// Newton's step, then two residual corrections of the quotient (Markstein).
__device__ __forceinline__ float div_rn(float a, float b) {
    float r = __builtin_amdgcn_rcpf(b);
    r = __builtin_fmaf(__builtin_fmaf(-b, r, 1.0f), r, r);
    float q = a * r;
    q = __builtin_fmaf(__builtin_fmaf(-b, q, a), r, q);
    return __builtin_fmaf(__builtin_fmaf(-b, q, a), r, q);
}

// sqrt(x) and a / b correctly rounded in f64 for 1 <= x < 2^64, 0 <= a <
// 2^32, 1 <= b < 2^32 (no scaling: every intermediate is normal). sqrt: the
// v_rsq_f64 seed, one coupled Newton step for sqrt(x) and 1 / (2 sqrt(x))
// (Goldschmidt), then two residual corrections; division: the v_rcp_f64
// seed refined by two Newton steps, then one residual correction
// (Markstein). Both match IEEE sqrt and division on every one of 2e7 random
// operands in these ranges with seeds perturbed by up to 2^-22.
__device__ __forceinline__ double sqrt_rn64(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    g = __builtin_fma(__builtin_fma(-g, g, x), h, g);
    return __builtin_fma(__builtin_fma(-g, g, x), h, g);
}
__device__ __forceinline__ double div_rn64(double a, double b) {
    double r = __builtin_amdgcn_rcp(b);
    r = __builtin_fma(__builtin_fma(-b, r, 1.0), r, r);
    r = __builtin_fma(__builtin_fma(-b, r, 1.0), r, r);
    const double q = a * r;
    return __builtin_fma(__builtin_fma(-b, q, a), r, q);
}

// finish_score over 32-bit exact sums: for templates up to 128 x 128, sum I T
// and sum I^2, sum T^2 stay below 255^2 * 128^2 < 2^31, and SSE = sum T^2 -
// 2 sum I T + sum I^2 is exact in u32. F64: normalise in f64 (windows of at
// least 64 x 64 pixels), else in f32 like the reference. The sqrt and the
// division are IEEE (sqrt_rn, div_rn and their f64 forms in the ranges here:
// sum I^2 * sum T^2 is 0 or at least 1, the quotient below 2^32).
template <int METHOD, bool F64>
__device__ __forceinline__ float finish_u32(uint32_t c, uint32_t sii, uint32_t stt) {
    const bool sse = (METHOD == IPH_SSE || METHOD == IPH_SSE_NORMED);
    const uint32_t num = sse ? stt + sii - 2u * c : c;
    if (METHOD == IPH_SSE || METHOD == IPH_CCORR) return (float)num;
    if constexpr (F64) {
        const double p = (double)sii * (double)stt;  // 0 or at least 1
        const bool pos = p > 0.0;
        const double q = div_rn64((double)num, sqrt_rn64(pos ? p : 1.0));
        return pos ? (float)q : (float)num;
    } else {
        const float norm = sqrt_rn((float)sii * (float)stt);
        const float sn = (float)num;
        return norm > 0.0f ? div_rn(sn, norm) : sn;
    }
}

// The output tile at (x0, y0) of one match (the body of the single-image and
// the job-table kernels). Template rows are taken in ceil(th / tcm) chunks of
// near-equal height (at most tcm rows): per chunk the workgroup stages the
// image rows those template rows meet and the rows themselves, correlates,
// and adds the chunk's rows of every window's sum of squares, so the LDS
// footprint is bounded by tcm rather than by the template height.
// Loads and stores go through raw buffers with 32-bit offsets (the host
// keeps images within mfma_match_fits): the image and the template from
// their 4-byte aligned starts, two dwords and a byte funnel per unaligned
// word, and the score map from the tile's first row.
template <int KB, int METHOD>
__device__ __forceinline__ void match_i8_fused_tile(const uint8_t *__restrict__ img, int W,
                                                    int H, const uint8_t *__restrict__ tpl,
                                                    float *__restrict__ out, int ow, int oh,
                                                    int tw, int th, int tcm, int x0,
                                                    int y0) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ uint32_t tred[2][8];
    constexpr bool kSums = METHOD != IPH_CCORR;  // sum I^2 needed
    constexpr bool kH1 = KB >= 2;  // window sums of I' from per-row sums (fits the image block)
    constexpr int PITCH = fused_pitch(KB), kTP = tpl_pitch(KB);
    constexpr int kTW = kTP / 4;       // u32 words per staged template row
    constexpr int kLPR = 2 * KB;       // lanes per staged image row, 16 bytes each
    const int nthr = (int)blockDim.x, RT = nthr >> 6;
    const int tid = (int)threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int n = lane & 31, g = lane >> 5;
    const int nchunk = (th + tcm - 1) / tcm;
    const int tc = (th + nchunk - 1) / nchunk;  // rows per chunk (the last may have fewer)
    const int tprows = min(th, tcm);
    const int rmax = 32 * RT + tprows - 1;  // staged image rows of a chunk, at most
    int8_t *blk = reinterpret_cast<int8_t *>(smem);
    const int tp_off = (rmax * PITCH + 15) & ~15;
    uint32_t *tp = reinterpret_cast<uint32_t *>(smem + tp_off);
    // one more staged row after the template rows: 0x01 where a template
    // column is (the all-ones band), then 16 zero bytes
    uint32_t *onesrow = tp + tprows * kTW;
    uint32_t *h2 = reinterpret_cast<uint32_t *>(
        smem + ((tp_off + (tprows + 1) * kTP + 16 + 15) & ~15));

    const uintptr_t ia = reinterpret_cast<uintptr_t>(img), ta = reinterpret_cast<uintptr_t>(tpl);
    const uint32_t ib = (uint32_t)ia & 3u, tb = (uint32_t)ta & 3u;
    const auto irs = raw_rsrc(reinterpret_cast<const void *>(ia - ib),
                              ((uint32_t)W * (uint32_t)H + ib + 3u) & ~3u);
    const auto trs = raw_rsrc(reinterpret_cast<const void *>(ta - tb),
                              ((uint32_t)(tw * th) + tb + 3u) & ~3u);

    // image staging: lane (row r0 + k RPP, 16 bytes at column 16 cl)
    const int rpp = nthr / kLPR;
    const int cl = tid % kLPR, r0 = tid / kLPR;
    const uint32_t gxc = (uint32_t)min(x0 + 16 * cl, W - 1);
    // template staging: lane (row rt0 + k TPP, word ct); bytes 4 ct - 32 ..
    // hold template columns, the rest of the row is zero
    const int tpp = nthr / kTW;
    const int ct = tid % kTW, rt0 = tid / kTW;
    const int jcol = 4 * ct - 32;
    uint32_t vm = 0;
#pragma unroll
    for (int b = 0; b < 4; b++)
        if (jcol + b >= 0 && jcol + b < tw) vm |= 0xFFu << (8 * b);
    const uint32_t tcol = (uint32_t)max(jcol, 0) + tb;
    if (rt0 == 0) onesrow[ct] = vm & 0x01010101u;
    if (tid < 4) onesrow[kTW + tid] = 0u;

    v16i acc, acs;
#pragma unroll
    for (int r = 0; r < 16; r++) acc[r] = acs[r] = 0;
    uint32_t sqv[16];
#pragma unroll
    for (int r = 0; r < 16; r++) sqv[r] = 0;
    uint32_t s1 = 0, s2 = 0;  // this thread's share of sum T, sum T^2
    v4i ones[KB];
    const int u0 = 32 + 16 * g - n;
    const uint32_t sh = (uint32_t)(u0 & 3);

    for (int c0 = 0; c0 < th; c0 += tc) {
        const int dyn = min(tc, th - c0);
        const int rows = 32 * RT + dyn - 1;
        if (c0 > 0) __syncthreads();  // the previous chunk is done with blk, tp and h2
        // image rows y0 + c0 .. (clamped), bytes x0 .. x0 + 32 KB - 1 as i8
        // (bytes past a row's end, or the image's, meet zero Toeplitz
        // entries or discarded outputs only), and template rows c0 ..; each
        // round issues all its loads before its LDS writes
        {
            constexpr int kIB = 3, kTB = 2;  // image and template rows per lane and round
            const int ni = (rows - r0 + kIB * rpp - 1) / (kIB * rpp);
            const int nt = max((dyn - rt0 + kTB * tpp - 1) / (kTB * tpp), 0);
            const int rounds = max(r0 < rpp ? ni : 0, rt0 < tpp ? nt : 0);
            for (int k = 0; k < rounds; k++) {
                uint32_t d[kIB][5], e[kTB][2], osh[kIB], tsh8[kTB];
#pragma unroll
                for (int u = 0; u < kIB; u++) {
                    const int rr = min(r0 + (k * kIB + u) * rpp, rows - 1);
                    const uint32_t y = (uint32_t)min(y0 + c0 + rr, H - 1);
                    const uint32_t o = y * (uint32_t)W + gxc + ib;
                    osh[u] = o & 3u;
#pragma unroll
                    for (int q = 0; q < 5; q++)
                        d[u][q] = __builtin_amdgcn_raw_buffer_load_b32(irs, (int)((o & ~3u) + 4 * q),
                                                                       0, 0);
                }
#pragma unroll
                for (int u = 0; u < kTB; u++) {
                    const int r = min(rt0 + (k * kTB + u) * tpp, dyn - 1);
                    const uint32_t o = (uint32_t)((c0 + r) * tw) + tcol;
                    tsh8[u] = o & 3u;
                    e[u][0] = __builtin_amdgcn_raw_buffer_load_b32(trs, (int)(o & ~3u), 0, 0);
                    e[u][1] = __builtin_amdgcn_raw_buffer_load_b32(trs, (int)((o & ~3u) + 4), 0, 0);
                }
#pragma unroll
                for (int u = 0; u < kIB; u++) {
                    const int rr = r0 + (k * kIB + u) * rpp;
                    if (r0 < rpp && rr < rows) {
                        v4i w;
#pragma unroll
                        for (int q = 0; q < 4; q++)
                            w[q] = (int)(__builtin_amdgcn_alignbyte(d[u][q + 1], d[u][q], osh[u]) ^
                                         0x80808080u);
                        *reinterpret_cast<v4i *>(blk + rr * PITCH + 16 * cl) = w;
                    }
                }
#pragma unroll
                for (int u = 0; u < kTB; u++) {
                    const int r = rt0 + (k * kTB + u) * tpp;
                    if (rt0 < tpp && r < dyn) {
                        const uint32_t wv = __builtin_amdgcn_alignbyte(e[u][1], e[u][0], tsh8[u]) & vm;
                        s1 = __builtin_amdgcn_udot4(wv, 0x01010101u, s1, false);
                        s2 = __builtin_amdgcn_udot4(wv, wv, s2, false);
                        tp[r * kTW + ct] = wv ^ (vm & 0x80808080u);
                    }
                }
            }
        }
        __syncthreads();
        if (!kH1 && c0 == 0) {
            // all-ones band: B[k][n] = 1 for 0 <= 32 kb + k - n < tw, cut from
            // the ones row like the template's Toeplitz slices
            const uint32_t *q0 = onesrow + (u0 >> 2);
#pragma unroll
            for (int kb = 0; kb < KB; kb++) {
                const uint32_t *q = q0 + 8 * kb;
#pragma unroll
                for (int d = 0; d < 4; d++)
                    ones[kb][d] = (int)__builtin_amdgcn_alignbyte(q[d + 1], q[d], sh);
            }
        }

        // horizontal tw-window sums of squares of the chunk's rows, 16 outputs
        // per task: columns 16q .. 16q + 15 (the tile's 32 output columns);
        // the first from v_dot4 over whole words, then 15 steps adding the
        // byte entering (16q + tw + x - 1) and dropping the one leaving
        // (16q + x - 1), both taken from words
        if constexpr (kSums) {
            const int nw = tw >> 2;
            const uint32_t tail = (tw & 3) ? (1u << (8 * (tw & 3))) - 1u : 0u;
            const uint32_t tsh = (uint32_t)(tw & 3);
            for (int it = tid; it < rows * 2; it += nthr) {
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
                uint32_t vo[4], vi[4], e[5];
#pragma unroll
                for (int k = 0; k < 4; k++) vo[k] = pw[k] ^ 0x80808080u;
#pragma unroll
                for (int k = 0; k < 5; k++) e[k] = pw[nw + k];
#pragma unroll
                for (int k = 0; k < 4; k++)
                    vi[k] = __builtin_amdgcn_alignbyte(e[k + 1], e[k], tsh) ^ 0x80808080u;
                uint32_t *o = h2 + r * kFusedHP + 16 * q;
                o[0] = b;
#pragma unroll
                for (int x = 1; x < 16; x++) {
                    const uint32_t bi = (vi[(x - 1) >> 2] >> (8 * ((x - 1) & 3))) & 0xFFu;
                    const uint32_t bo = (vo[(x - 1) >> 2] >> (8 * ((x - 1) & 3))) & 0xFFu;
                    b += bi * bi - bo * bo;
                    o[x] = b;
                }
            }
        }

        // correlation on MFMA over the chunk's template rows (with KB = 1, the
        // window sums of I' too, against the all-ones band)
        const int8_t *arow = blk + (wave * 32 + n) * PITCH + 16 * g;
        const uint32_t *trow = tp + (u0 >> 2);
#pragma unroll 1
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
                if constexpr (!kH1) acs = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, ones[kb], acs, 0, 0, 0);
            }
            arow += PITCH;
            trow += kTW;
        }

        // KB >= 2: the horizontal tw-window sums H[r][n] of I' of every staged
        // row, once per row, on MFMA against the all-ones band (row blocks
        // wave and wave + RT: chunks are at most 32 RT + 1 rows high, so there
        // are at most 2 RT blocks), then written over the image block as i16
        // (|H| <= 128 * 128) for the vertical slide
        const int16_t *h1s = reinterpret_cast<const int16_t *>(blk);
        if constexpr (kH1) {
            const int nrb = (rows + 31) >> 5;
            // both blocks on every wave (rows clamped): a block past the last
            // one is computed and dropped, which keeps the MFMAs unconditional
            const int8_t *ar0 = blk + min(32 * wave + n, rows - 1) * PITCH + 16 * g;
            const int8_t *ar1 = blk + min(32 * (wave + RT) + n, rows - 1) * PITCH + 16 * g;
            v16i hb0, hb1;
#pragma unroll
            for (int r = 0; r < 16; r++) hb0[r] = hb1[r] = 0;
#pragma unroll
            for (int kb = 0; kb < KB; kb++) {
                // the all-ones band B[k][n] = 1 for 0 <= 32 kb + k - n < tw,
                // cut from the ones row like the template's Toeplitz slices
                const uint32_t *q = onesrow + (u0 >> 2) + 8 * kb;
                v4i ob;
#pragma unroll
                for (int d = 0; d < 4; d++)
                    ob[d] = (int)__builtin_amdgcn_alignbyte(q[d + 1], q[d], sh);
                hb0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(
                    *reinterpret_cast<const v4i *>(ar0 + 32 * kb), ob, hb0, 0, 0, 0);
                hb1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(
                    *reinterpret_cast<const v4i *>(ar1 + 32 * kb), ob, hb1, 0, 0, 0);
            }
            __syncthreads();  // every wave is done reading the image block
            // register r holds row 32 b + 4 g + 8 (r >> 2) + (r & 3) of column n
            int16_t *hw = reinterpret_cast<int16_t *>(blk) + (32 * wave + 4 * g) * kH1P + n;
            const int lim = rows - (32 * wave + 4 * g);  // rows left from the lane's first
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int k = 8 * (r >> 2) + (r & 3);
                if (k < lim) hw[k * kH1P] = (int16_t)hb0[r];
            }
            if (wave + RT < nrb) {
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    const int k = 32 * RT + 8 * (r >> 2) + (r & 3);
                    if (k < lim) hw[k * kH1P] = (int16_t)hb1[r];
                }
            }
        }

        if constexpr (kSums || kH1) {
            __syncthreads();  // every row's horizontal sums are in LDS
            // lane (n, g) holds rows 32w + 8j + 4g + i (register 4j + i) of column
            // n: slide the chunk's dyn-row sums down rows 32w + 4g .. + 27
            const uint32_t *hc = h2 + (wave * 32 + 4 * g) * kFusedHP + n;
            const int16_t *hc1 = h1s + (wave * 32 + 4 * g) * kH1P + n;
            uint32_t sq = 0;
            int s1w = 0;
            for (int d = 0; d < dyn; d++) {
                if constexpr (kSums) sq += hc[d * kFusedHP];
                if constexpr (kH1) s1w += hc1[d * kH1P];
            }
#pragma unroll
            for (int j = 0; j < 4; j++) {
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int yl = 8 * j + i;  // row offset from 32w + 4g
                    if constexpr (kSums) sqv[4 * j + i] += sq;
                    if constexpr (kH1) acs[4 * j + i] += s1w;
                    // to row yl + 1 (past the other half's four rows after i = 3)
                    if (j < 3 || i < 3) {
#pragma unroll
                        for (int k = 0; k < ((i == 3) ? 5 : 1); k++) {
                            const int r = yl + k;
                            if constexpr (kSums) sq += hc[(r + dyn) * kFusedHP] - hc[r * kFusedHP];
                            if constexpr (kH1) s1w += hc1[(r + dyn) * kH1P] - hc1[r * kH1P];
                        }
                    }
                }
            }
        }
    }

    uint32_t l1 = s1, l2 = s2;  // sum T <= 2^22, sum T^2 < 2^31: no wrap
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
        st += tred[0][w];
        stt += tred[1][w];
    }
    // the score map from the tile's first row: offsets below 32 RT rows
    const int x = x0 + n;
    float *obase = out + (size_t)y0 * (uint32_t)ow;
    const uint32_t orows = (uint32_t)min(oh - y0, 32 * RT);
    const auto ors = raw_rsrc(obase, orows * (uint32_t)ow * 4u);
    const uint32_t voff = ((uint32_t)(wave * 32 + 4 * g) * (uint32_t)ow + (uint32_t)x) * 4u;
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
                const uint32_t soff = (uint32_t)(8 * (r >> 2) + (r & 3)) * (uint32_t)ow * 4u;
                __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v[r]), ors, (int)voff,
                                                      (int)soff, 0);
            }
        }
    };
    const bool f64 = (METHOD == IPH_SSE_NORMED || METHOD == IPH_CCORR_NORMED) && tw * th >= 64 * 64;
    if (f64) emit(std::true_type{});
    else emit(std::false_type{});
}

template <int KB, int METHOD>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void match_i8_fused_kernel(
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
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void match_i8_fused_jobs_kernel(const MatchJob *__restrict__ jobs,
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

// LDS per workgroup and per CU on the current device (64 KiB and 160 KiB if
// the queries fail).
size_t device_lds(hipDeviceAttribute_t what, size_t dflt) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&v, what, dev) != hipSuccess ||
        v <= 0)
        return dflt;
    return (size_t)v;
}
size_t fused_lds_limit() {
    static const size_t v = device_lds(hipDeviceAttributeMaxSharedMemoryPerBlock, 64 * 1024);
    return v;
}
size_t fused_lds_cu() {
    static const size_t v = device_lds(hipDeviceAttributeMaxSharedMemoryPerMultiprocessor,
                                       160 * 1024);
    return v;
}
constexpr int kFusedMinChunk = 16;  // chunks are never cut below this many rows
constexpr int kFusedMaxWaves = 16;  // per CU: the kernel's registers allow 4 per SIMD

// Waves per CU that a workgroup of RT waves and `lds` bytes allows.
int fused_waves(size_t lds, int RT) {
    return std::min<int>(kFusedMaxWaves, (int)(fused_lds_cu() / std::max<size_t>(lds, 1)) * RT);
}

// Template rows per chunk for RT waves: the fewest near-equal chunks whose
// footprint allows `min_waves` (12 or 8) waves per CU (else 8); 0 if no chunk of at
// least kFusedMinChunk rows fits a workgroup. Every chunk restages the image
// rows it meets, so fewer chunks save staging and horizontal sums. Chunks
// are at most 32 RT + 1 rows (the staged rows then form at most 2 RT blocks
// of 32 for the per-row window sums).
int fused_tc(int KB, int th, int RT, bool sums, int min_waves = 12) {
    for (int waves : {12, 8}) {
        if (waves > min_waves) continue;
        for (int nc = 1; nc <= th; nc++) {
            const int tc = (th + nc - 1) / nc;
            if (tc > 32 * RT + 1) continue;
            if (tc < kFusedMinChunk && nc > 1) break;
            const size_t lds = fused_lds(KB, 0, th, RT, sums, tc);
            if (lds <= fused_lds_limit() && fused_waves(lds, RT) >= waves) return tc;
        }
    }
    const int tc = std::min({th, kFusedMinChunk, 32 * RT + 1});
    return fused_lds(KB, 0, th, RT, sums, tc) <= fused_lds_limit() ? tc : 0;
}

// Workgroup shape for one match: waves (32 output rows each) and chunk
// height. Most waves per CU first, counting 12 or more as equal, then the
// fewest staged image rows per output row (taller tiles spread the th - 1
// halo rows over more outputs; fewer chunks restage fewer rows).
void fused_shape(int KB, int th, bool sums, int &RT, int &tc) {
    RT = 0;
    tc = 0;
    int best_w = 0;
    double best_rows = 1e30;
    for (int rt : {8, 4, 2, 1}) {
        const int c = fused_tc(KB, th, rt, sums);
        if (c == 0) continue;
        const int w = fused_waves(fused_lds(KB, 0, th, rt, sums, c), rt);
        const int nc = (th + c - 1) / c;
        const double rows = (double)nc * (32 * rt + c - 1) / (32.0 * rt);
        const int wc = std::min(w, 12);
        if (wc > best_w || (wc == best_w && rows < best_rows - 1e-9)) {
            best_w = wc;
            best_rows = rows;
            RT = rt;
            tc = c;
        }
    }
}

template <int KB>
iph_status launch_i8_fused_kb(const uint8_t *img, int W, int H, const uint8_t *tpl, float *out,
                              int ow, int oh, int tw, int th, int method, int RT, int tc,
                              hipStream_t s) {
    const size_t lds = fused_lds(KB, tw, th, RT, method != IPH_CCORR, tc);
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
    double *wsq = method == IPH_CCORR ? nullptr : wsp.take<double>((size_t)(w - tw + 1) * (h - th + 1));
    if (!tsq || (method != IPH_CCORR && !wsq)) {
        set_error("workspace too small for the BF16x3 MFMA match_template");
        return IPH_EINVAL;
    }
    int ow = (int)(w - tw + 1), oh = (int)(h - th + 1);
    if (method != IPH_CCORR) {
        template_sq_f32_kernel<<<1, 256, 0, s>>>(tpl, (int)(tw * th), tsq);
        IPH_TRY(launch_status("template_sq_f32_kernel"));
    }
    if (method != IPH_CCORR) {
        dim3 eg((ow + kWsqCols - 1) / kWsqCols, (oh + kWsqRows - 1) / kWsqRows);
        size_t lds = (size_t)(kWsqCols + tw - 1) * sizeof(double);
        window_sq_f32_kernel<<<eg, 256, lds, s>>>(img, (int)w, (int)tw, (int)th, ow, oh, wsq);
        IPH_TRY(launch_status("window_sq_f32_kernel"));
    }
    switch (KB) {
#define C(K)                                                                                 \
    case K:                                                                                  \
        return launch_bf16_kb<K>(img, w, h, tpl, wsq, tsq, out, ow, oh, tw, th, method, s);
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
        if (method != IPH_CCORR) wp.take<double>((size_t)(w - tw + 1) * (h - th + 1));
        return wp.used;
    }
    return 256;  // the fused int8 kernel needs no workspace
}

iph_status launch_mfma_match(const void *img, int32_t p_type, uint32_t w, uint32_t h,
                             const void *tpl, uint32_t tw, uint32_t th, int32_t method,
                             float *out, void *ws, size_t ws_bytes, hipStream_t s) {
    if (!mfma_match_fits(w, h)) {
        set_error("the MFMA match paths take images below 2^32 - 64 pixels and 2^23 columns; "
                  "use the direct path");
        return IPH_EUNSUPPORTED;
    }
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
    int RT, tc;
    fused_shape(KB, (int)th, method != IPH_CCORR, RT, tc);
    if (RT == 0) {
        set_error("no fused int8 workgroup shape fits %zu bytes of LDS", fused_lds_limit());
        return IPH_EUNSUPPORTED;
    }
    const int ow = (int)(w - tw + 1), oh = (int)(h - th + 1);
#define C(K)                                                                                 \
    case K: return launch_i8_fused_kb<K>(im8, w, h, tp8, out, ow, oh, tw, th, method, RT, tc, s)
    switch (KB) {
        C(1); C(2); C(3); C(4);
    default: return launch_i8_fused_kb<5>(im8, w, h, tp8, out, ow, oh, tw, th, method, RT, tc, s);
    }
#undef C
}



int fused_match_kb(uint32_t tw) { return blocks_for((int)tw); }

int fused_match_rt_max(uint32_t tw, uint32_t th, int32_t method) {
    if (tw == 0 || th == 0 || tw > (uint32_t)kMaxSide || th > (uint32_t)kMaxSide) return 0;
    const int KB = blocks_for((int)tw);
    for (int rt : {8, 4, 2, 1})
        if (fused_tc(KB, (int)th, rt, method != IPH_CCORR) > 0) return rt;
    return 0;
}

int fused_match_rt_default(uint32_t tw, uint32_t th, int32_t method) {
    int rt, tc;
    fused_shape(blocks_for((int)tw), (int)th, method != IPH_CCORR, rt, tc);
    return rt;
}

int fused_match_tc(int kb, uint32_t th, int rt, int32_t method, int min_waves) {
    return fused_tc(kb, (int)th, rt, method != IPH_CCORR, min_waves);
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
    if (ntiles > (1LL << 30) || kb < 1 || kb > 5 || (rt != 1 && rt != 2 && rt != 4 && rt != 8) ||
        tc < 1 || lds > fused_lds_limit()) {
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
