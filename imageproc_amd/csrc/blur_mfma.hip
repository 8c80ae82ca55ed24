// Separable filter for GrayImage (u8 pixels, f32 taps) on f16 MFMA, with the
// reference's exact result (filter/mod.rs:228-429).
//
// The reference computes, per output, the sequential non-fused f32 chain
//   R = (((0 + p0 w0) + p1 w1) + ...),  out = R < 255 ? (R > 0 ? trunc(R) : 0) : 255
// once along rows into a u8 intermediate and once along columns of it. Only
// trunc(R) matters, so this kernel
//   1. estimates every sum E on v_mfma_f32_32x32x16_f16: pixels (0..255) are
//      exact in f16, each tap is split w*2^S = hi + lo in f16 (22 bits), and
//      the 1-D correlation of a 32x32 block is A(32x(16*NKB)) x B, B the
//      Toeplitz matrix B[k][n] = w[k - n];
//   2. proves trunc(E) = trunc(R) whenever E is farther than eps from an
//      integer, eps >= |E - R| (bound derived in blur_eps below);
//   3. recomputes the few outputs within eps of an integer with the
//      reference's sequential chain (separately rounded __fmul_rn /
//      __fadd_rn) from the same LDS data, and overwrites them.
// Both passes run in one workgroup: the horizontal pass writes the u8
// intermediate (as exact f16 integers, transposed) to LDS, the vertical pass
// reads it back as its A operand.
#include "iph_internal.hpp"

#include <algorithm>
#include <cmath>
#include <type_traits>

namespace iph {
namespace {

typedef _Float16 v8h __attribute__((ext_vector_type(8)));
typedef _Float16 v2h __attribute__((ext_vector_type(2)));
typedef float v16f __attribute__((ext_vector_type(16)));

constexpr int kTW = 128;       // output columns per workgroup (4 column blocks)
constexpr int kTH = 96;        // output rows per workgroup (3 row blocks)
constexpr int kMaxTaps = 48;   // NKB = ceil((31 + L) / 16) <= 5

struct BlurArgs {
    float h[kMaxTaps];
    float v[kMaxTaps];
    float scale;      // 2^S applied to the taps fed to MFMA
    float inv_scale;  // 2^-S
    float eps_h, eps_v;
    int L;
    int m;  // (x0 - L/2) mod 4: the window's staged columns start 4-aligned
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

constexpr int kTapPad = 40;  // zeros on both sides of the taps in LDS

// Toeplitz B operand for this lane: column n = lane % 32, rows
// k = 16*kb + 8*(lane/32) + j, value w[k - n - shift] * 2^S split into hi + lo
// (shift: offset of the first tap's input column inside the staged window).
// `taps_s` holds the scaled taps at [kTapPad, kTapPad + L) and zeros around.
template <int NKB>
__device__ __forceinline__ void build_b(const float *taps_s, int lane, int shift,
                                        v8h (&bhi)[NKB], v8h (&blo)[NKB]) {
    const int n = lane & 31, g = lane >> 5;
#pragma unroll
    for (int kb = 0; kb < NKB; kb++) {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            float w = taps_s[kTapPad + 16 * kb + 8 * g + j - n - shift];
            _Float16 hi = (_Float16)w;
            _Float16 lo = (_Float16)(w - (float)hi);
            bhi[kb][j] = hi;
            blo[kb][j] = lo;
        }
    }
}

// 8 u8 pixels -> 8 exact f16: (0x6400 | p) is 1024 + p in f16.
__device__ __forceinline__ v8h bytes_to_half8(uint2 b) {
    const uint32_t magic = 0x64646464u;
    uint32_t q0 = __builtin_amdgcn_perm(magic, b.x, 0x04010400u);
    uint32_t q1 = __builtin_amdgcn_perm(magic, b.x, 0x04030402u);
    uint32_t q2 = __builtin_amdgcn_perm(magic, b.y, 0x04010400u);
    uint32_t q3 = __builtin_amdgcn_perm(magic, b.y, 0x04030402u);
    const v2h off = {(_Float16)-1024.0f, (_Float16)-1024.0f};
    v2h h0 = __builtin_bit_cast(v2h, q0) + off;
    v2h h1 = __builtin_bit_cast(v2h, q1) + off;
    v2h h2 = __builtin_bit_cast(v2h, q2) + off;
    v2h h3 = __builtin_bit_cast(v2h, q3) + off;
    v8h r;
    r[0] = h0[0]; r[1] = h0[1]; r[2] = h1[0]; r[3] = h1[1];
    r[4] = h2[0]; r[5] = h2[1]; r[6] = h3[0]; r[7] = h3[1];
    return r;
}

// trunc/clamp of the reference (Clamp<f32> for u8) applied to an estimate
// known to be more than eps away from any integer: floor == trunc for t > 0.
__device__ __forceinline__ float clamp_floor(float t) {
    return __builtin_amdgcn_fmed3f(floorf(t), 0.0f, 255.0f);
}

__device__ __forceinline__ bool near_integer(float t, float eps) {
    float d = t - floorf(t);
    return fabsf(d - 0.5f) > 0.5f - eps;
}

__device__ __forceinline__ uint8_t clamp_u8_exact(float x) {
    return x < 255.0f ? (x > 0.0f ? (uint8_t)(uint32_t)x : (uint8_t)0) : (uint8_t)255;
}


// Flagged-output queue entries per wave; drained whenever more than
// kQDrain entries wait (one group of 4 accumulator slots adds at most 256).
constexpr int kQCap = 512;
constexpr int kQDrain = kQCap - 256;

// Appends this lane's entry to the wave's queue when `flag` is set. `qn` is
// the wave-uniform queue length. Costs one compare and a scalar branch when
// no lane is flagged.
__device__ __forceinline__ void enqueue(bool flag, uint32_t entry, uint32_t *q, int &qn) {
    const unsigned long long m = __ballot(flag);
    if (m) {
        const int below = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                         __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        if (flag && qn + below < kQCap) q[qn + below] = entry;
        qn += __popcll(m);
    }
}

// Appends this lane's 16-bit entry when `flag` is set; `m` is the flag's
// wave-wide ballot. The caller guarantees room (qn + popcount(m) <= capacity).
__device__ __forceinline__ void enqueue_ballot(bool flag, unsigned long long m, uint32_t entry,
                                               uint16_t *q, int &qn) {
    if (m) {
        const int below = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                         __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        if (flag) q[qn + below] = (uint16_t)entry;
        qn += __popcll(m);
    }
}

// The reference's sequential chain over taps w*2^S (exact: power-of-two
// scaling commutes with every f32 rounding here), returned unscaled.
// `stride` steps through the pixels; all loads are issued before the chain.
template <int LMAX, typename T>
__device__ __forceinline__ float exact_chain(const T *p, const float *taps_s, int L,
                                             float inv_scale) {
    float v[LMAX];
#pragma unroll
    for (int q = 0; q < LMAX; q++) v[q] = (float)p[q];
    float s = 0.0f;
#pragma unroll
    for (int q = 0; q < LMAX; q++)
        if (q < L) s = __fadd_rn(s, __fmul_rn(v[q], taps_s[kTapPad + q]));
    return s * inv_scale;
}

// exact_chain with its loads issued eight at a time (fewer live registers).
template <int LMAX, typename T>
__device__ __forceinline__ float exact_chain8(const T *p, const float *taps, int L) {
    float s = 0.0f;
#pragma unroll
    for (int q0 = 0; q0 < LMAX; q0 += 8) {
        float v[8];
#pragma unroll
        for (int q = 0; q < 8; q++)
            if (q0 + q < LMAX) v[q] = (float)p[q0 + q];
#pragma unroll
        for (int q = 0; q < 8; q++)
            if (q0 + q < LMAX && q0 + q < L)
                s = __fadd_rn(s, __fmul_rn(v[q], taps[kTapPad + q0 + q]));
    }
    return s;
}

// Staged rows per tile: max(kTH + L - 1, 32*(kTH/32 - 1) + 16*NKB) <= 128
// for L <= 33, rounded up to whole 32-row blocks.
constexpr int kRows = 128;

template <int NKB, int ROWS = kRows> struct BlurShape {
    static constexpr int CIN = kTW - 32 + 16 * NKB;  // staged input columns
    static constexpr int C4 = CIN / 4;                 // dwords per staged row
    static constexpr int NTOT = ROWS * C4;             // staged dwords per tile
    static constexpr int NREG = (NTOT + 255) / 256;    // staged dwords per thread
    // staged input row pitch: an odd multiple of 2 dwords, so the 32 rows one
    // 8-byte A-operand read touches fall in distinct LDS banks
    static constexpr int PIN = ((CIN / 4) % 2 == 0) ? CIN + 8 : CIN + 4;
    static constexpr int PITCH_T = ROWS + 8;           // f16 per intermediate column
    static constexpr int LMAX = 16 * NKB - 31;         // largest tap count for NKB
};

// Issues this thread's share of one tile's input-window loads into
// registers: rows clamp(ys + r), columns xs + c, xs = x0 - L/2 - m a multiple
// of 4. ALIGNED (rows start 4-byte aligned): every staged dword lies wholly
// inside or wholly outside its row and is one dword load from the clamped
// column; the edge byte is replicated later, in store_tile, once the data has
// arrived (so no load waits for its predecessor). Otherwise each byte is
// loaded from its clamped column.
template <typename S, bool ALIGNED>
__device__ __forceinline__ void load_tile(const uint8_t *__restrict__ img, int W, int H, int xs,
                                          int ys, uint32_t (&regs)[S::NREG]) {
#pragma unroll
    for (int k = 0; k < S::NREG; k++) {
        const int i = min((int)threadIdx.x + 256 * k, S::NTOT - 1);
        const int r = i / S::C4, c = (i - r * S::C4) * 4;
        const uint8_t *row = img + (size_t)clampi(ys + r, 0, H - 1) * W;
        const int gx = xs + c;
        if (ALIGNED) {
            regs[k] = *reinterpret_cast<const uint32_t *>(row + clampi(gx, 0, W - 4));
        } else {
            regs[k] = (uint32_t)row[clampi(gx, 0, W - 1)] |
                      ((uint32_t)row[clampi(gx + 1, 0, W - 1)] << 8) |
                      ((uint32_t)row[clampi(gx + 2, 0, W - 1)] << 16) |
                      ((uint32_t)row[clampi(gx + 3, 0, W - 1)] << 24);
        }
    }
}

// Writes the loaded window to LDS; for ALIGNED loads, dwords left of the row
// become its first byte replicated and dwords right of it its last byte.
template <typename S, bool ALIGNED>
__device__ __forceinline__ void store_tile(uint8_t *in_s, int W, int xs,
                                           const uint32_t (&regs)[S::NREG]) {
#pragma unroll
    for (int k = 0; k < S::NREG; k++) {
        const int i = (int)threadIdx.x + 256 * k;
        if (S::NTOT % 256 != 0 && i >= S::NTOT) break;
        const int r = i / S::C4, c = (i - r * S::C4) * 4;
        uint32_t v = regs[k];
        if (ALIGNED) {
            const int gx = xs + c;
            const uint32_t left = (v & 0xffu) * 0x01010101u;
            const uint32_t right = (v >> 24) * 0x01010101u;
            v = gx < 0 ? left : v;
            v = gx >= W ? right : v;
        }
        *reinterpret_cast<uint32_t *>(in_s + r * S::PIN + c) = v;
    }
}

// Persistent kernel: each workgroup walks a strided list of 128x96 output
// tiles. Tiles are split into 8 contiguous ranges, one per XCD (workgroups
// are dispatched round-robin over XCDs), so neighbouring tiles, which share
// halo rows and columns, meet in the same L2. The next tile's input is
// loaded into registers while the current tile is computed; the output tile
// is assembled in LDS and written with 16-byte row-contiguous stores.
// LDS holds two workgroups per CU (2 waves per SIMD), so the register budget
// of 2 waves per SIMD is free: B operands and the prefetch stay resident.
template <int NKB, bool SAME_TAPS, bool ALIGNED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void blur_u8_mfma_kernel(const uint8_t *__restrict__ src,
                                                           uint8_t *__restrict__ dst, int W,
                                                           int H, size_t stride, int n_img,
                                                           int tiles_x, int tiles_y,
                                                           BlurArgs a) {
    using S = BlurShape<NKB>;
    constexpr int PIN = S::PIN, PITCH_T = S::PITCH_T, LMAX = S::LMAX;
    extern __shared__ __align__(16) unsigned char smem[];
    uint8_t *in_s = smem;                                                    // kRows x PIN
    _Float16 *ht = reinterpret_cast<_Float16 *>(smem + (size_t)kRows * PIN);  // kTW x PITCH_T
    uint8_t *out_s = reinterpret_cast<uint8_t *>(ht + (size_t)kTW * PITCH_T);  // kTH x kTW
    float *taps_h = reinterpret_cast<float *>(out_s + (size_t)kTH * kTW);
    float *taps_v = taps_h + (2 * kTapPad + kMaxTaps);
    uint32_t *queue = reinterpret_cast<uint32_t *>(taps_v + (2 * kTapPad + kMaxTaps)) +
                      (threadIdx.x >> 6) * kQCap;
    const int L = a.L;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n = lane & 31, g = lane >> 5;
    const int half = L / 2;

    // scaled taps, zero padded, for branch-free Toeplitz construction
    if (threadIdx.x < 2 * kTapPad + kMaxTaps) {
        int t = (int)threadIdx.x - kTapPad;
        bool in = t >= 0 && t < L;
        taps_h[threadIdx.x] = in ? a.h[t] * a.scale : 0.0f;
        taps_v[threadIdx.x] = in ? a.v[t] * a.scale : 0.0f;
    }
    __syncthreads();
    v8h bh_hi[NKB], bh_lo[NKB], bv_hi[NKB], bv_lo[NKB];
    const int m = a.m;
    build_b<NKB>(taps_h, lane, m, bh_hi, bh_lo);
    build_b<NKB>(taps_v, lane, 0, bv_hi, bv_lo);
    const float *vts = SAME_TAPS ? taps_h : taps_v;

    const int per_img = tiles_x * tiles_y;
    const int ntiles = n_img * per_img;
    const int per_xcd = (int)gridDim.x >> 3;
    const int chunk = (ntiles + 7) >> 3;
    const int t_begin = (int)(blockIdx.x & 7) * chunk;
    const int t_end = min(ntiles, t_begin + chunk);
    int t = t_begin + (int)(blockIdx.x >> 3);

    uint32_t regs[S::NREG];
    if (t < t_end) {
        int im = t / per_img, rem = t - im * per_img, ty = rem / tiles_x;
        load_tile<S, ALIGNED>(src + (size_t)im * stride, W, H,
                                (rem - ty * tiles_x) * kTW - half - m, ty * kTH - half, regs);
    }
    for (; t < t_end; t += per_xcd) {
        const int im = t / per_img, rem = t - im * per_img, ty = rem / tiles_x;
        const int x0 = (rem - ty * tiles_x) * kTW, y0 = ty * kTH;
        uint8_t *out = dst + (size_t)im * stride;

        __syncthreads();  // the previous tile is done with in_s, ht and out_s
        store_tile<S, ALIGNED>(in_s, W, x0 - half - m, regs);
        const int tn = t + per_xcd;
        if (tn < t_end) {
            int im2 = tn / per_img, rem2 = tn - im2 * per_img, ty2 = rem2 / tiles_x;
            load_tile<S, ALIGNED>(src + (size_t)im2 * stride, W, H,
                                    (rem2 - ty2 * tiles_x) * kTW - half - m, ty2 * kTH - half, regs);
        }
        __syncthreads();

        // ---- horizontal pass: (kRows/32) x 4 blocks of 32x32 -> ht (transposed f16)
        int qn = 0;
        for (int blk = wave; blk < (kRows / 32) * (kTW / 32); blk += 4) {
            const int rb = blk >> 2, cb = blk & 3;
            v16f acc;
#pragma unroll
            for (int r = 0; r < 16; r++) acc[r] = 0.0f;
            const uint8_t *arow = in_s + (32 * rb + n) * PIN + 32 * cb + 8 * g;
#pragma unroll
            for (int kb = 0; kb < NKB; kb++) {
                v8h av = bytes_to_half8(*reinterpret_cast<const uint2 *>(arow + 16 * kb));
                acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(av, bh_lo[kb], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(av, bh_hi[kb], acc, 0, 0, 0);
            }
            // lane holds column n of the block, rows 8*(r/4) + 4*g + r%4
            _Float16 *col = ht + (size_t)(32 * cb + n) * PITCH_T + 32 * rb + 4 * g;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                v2h p0, p1;
                float tv[4];
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    tv[i] = acc[4 * j + i] * a.inv_scale;
                    const int row = 32 * rb + 8 * j + 4 * g + i;
                    enqueue(near_integer(tv[i], a.eps_h), (uint32_t)(row << 8) | (32 * cb + n),
                            queue, qn);
                }
                p0[0] = (_Float16)clamp_floor(tv[0]);
                p0[1] = (_Float16)clamp_floor(tv[1]);
                p1[0] = (_Float16)clamp_floor(tv[2]);
                p1[1] = (_Float16)clamp_floor(tv[3]);
                uint2 pk = make_uint2(__builtin_bit_cast(uint32_t, p0),
                                      __builtin_bit_cast(uint32_t, p1));
                *reinterpret_cast<uint2 *>(col + 8 * j) = pk;
                if (qn > kQDrain) {
                    // exact recomputation of the outputs an integer boundary could flip
                    for (int e = lane; e < qn; e += 64) {
                        const uint32_t ent = queue[e];
                        const int row = (int)(ent >> 8), cl = (int)(ent & 255);
                        float r = exact_chain<LMAX>(in_s + row * PIN + m + cl, taps_h, L,
                                                    a.inv_scale);
                        ht[(size_t)cl * PITCH_T + row] = (_Float16)(float)clamp_u8_exact(r);
                    }
                    qn = 0;
                }
            }
        }
        for (int e = lane; e < qn; e += 64) {
            const uint32_t ent = queue[e];
            const int row = (int)(ent >> 8), cl = (int)(ent & 255);
            float r = exact_chain<LMAX>(in_s + row * PIN + m + cl, taps_h, L, a.inv_scale);
            ht[(size_t)cl * PITCH_T + row] = (_Float16)(float)clamp_u8_exact(r);
        }
        __syncthreads();

        // ---- vertical pass: D[m = x][n = y] = sum_k H^T[x][k] B[k][y] -> out_s
        qn = 0;
        for (int blk = wave; blk < (kTH / 32) * (kTW / 32); blk += 4) {
            const int rb = blk >> 2, cb = blk & 3;
            v16f acc;
#pragma unroll
            for (int r = 0; r < 16; r++) acc[r] = 0.0f;
            const _Float16 *acol = ht + (size_t)(32 * cb + n) * PITCH_T + 32 * rb + 8 * g;
#pragma unroll
            for (int kb = 0; kb < NKB; kb++) {
                v8h av = *reinterpret_cast<const v8h *>(acol + 16 * kb);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(av, bv_lo[kb], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(av, bv_hi[kb], acc, 0, 0, 0);
            }
            // lane holds output row y = 32*rb + n, columns x = 32*cb + 8*(r/4) + 4*g + r%4
            const int yl = 32 * rb + n;
            uint8_t *orow = out_s + yl * kTW;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                uint32_t word = 0;
                const int xl = 32 * cb + 8 * j + 4 * g;
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    float tv = acc[4 * j + i] * a.inv_scale;
                    enqueue(near_integer(tv, a.eps_v), (uint32_t)(yl << 8) | (uint32_t)(xl + i),
                            queue, qn);
                    word |= (uint32_t)clamp_floor(tv) << (8 * i);
                }
                *reinterpret_cast<uint32_t *>(orow + xl) = word;
                if (qn > kQDrain) {
                    for (int e = lane; e < qn; e += 64) {
                        const uint32_t ent = queue[e];
                        const int y2 = (int)(ent >> 8), x2 = (int)(ent & 255);
                        float r = exact_chain<LMAX>(ht + (size_t)x2 * PITCH_T + y2, vts, L,
                                                    a.inv_scale);
                        out_s[y2 * kTW + x2] = clamp_u8_exact(r);
                    }
                    qn = 0;
                }
            }
        }
        for (int e = lane; e < qn; e += 64) {
            const uint32_t ent = queue[e];
            const int y2 = (int)(ent >> 8), x2 = (int)(ent & 255);
            float r = exact_chain<LMAX>(ht + (size_t)x2 * PITCH_T + y2, vts, L, a.inv_scale);
            out_s[y2 * kTW + x2] = clamp_u8_exact(r);
        }
        __syncthreads();

        // ---- copy the output tile out: 8 lanes x 16 bytes per row
        const bool vec_ok = ((W & 15) == 0) && ((reinterpret_cast<uintptr_t>(out) & 15) == 0);
        for (int i = threadIdx.x; i < kTH * (kTW / 16); i += 256) {
            const int yl = i >> 3, xl = (i & 7) * 16;
            const int gy = y0 + yl, gx = x0 + xl;
            if (gy >= H || gx >= W) continue;
            const uint4 v = *reinterpret_cast<const uint4 *>(out_s + yl * kTW + xl);
            uint8_t *o = out + (size_t)gy * W + gx;
            if (vec_ok && gx + 15 < W) {
                *reinterpret_cast<uint4 *>(o) = v;
            } else {
                const uint8_t *b = out_s + yl * kTW + xl;
                for (int q = 0; q < 16 && gx + q < W; q++) o[q] = b[q];
            }
        }
    }
}

// ---- fixed-point readout: taps >= 0 with 255 * sum(w) < 256 ------------------
//
// Taps are scaled by 2^15 and split w * 2^15 = hi + lo with hi = rint(w * 2^15)
// (an integer, exact in f16) and lo in f16. The hi products p * hi are
// integers, accumulated from 2^23 + c: every partial sum is an integer below
// 2^24, so that accumulator is exact whatever order the MFMA sums in. The lo
// products go to a second accumulator started at 0, whose partial sums stay
// small enough that its roundings are negligible. Their f32 sum is
// 2^23 + c + E, E the estimate in units of 2^-15 (one rounding, at most half a
// unit), and its bit pattern u carries floor((E + c) / 2^15) in bits 15..22
// and (E + c) mod 2^15 in bits 0..14:
//   * (u & 0x7fff) >= 2c  proves the estimate is at least c units (>= eps) away
//     from an integer, so bits 15..22 are the reference's truncated result (the
//     reference's sum lies in [0, 256), where its clamp is the identity);
//   * otherwise the output is recomputed with the reference's chain.
// Readout is an add, an AND, a compare and 1.25 byte-packing ops per output.

// Flagged-output queue of the fixed-point kernel: 16-bit entries (row << 8 |
// column), room for half a block (512 outputs) above the drain threshold, so
// the queue is drained at most twice per block.
constexpr int kFxQDrain = 64;
constexpr int kFxQCap = kFxQDrain + 512;

constexpr int kTH2 = 128;           // output rows per tile (4 row blocks)
constexpr int kRows2 = 160;         // staged rows: kTH2 + 32 covers L <= 33
constexpr int kPitchB = kRows2 + 8; // bytes per intermediate column (42 dwords)
constexpr int kOPitch = kTW + 16;   // bytes per output-tile row

constexpr int kFxShift = 15;  // estimates in units of 2^-kFxShift
constexpr uint32_t kFxFracMask = (1u << kFxShift) - 1u;

// One blur job of the fixed-point kernel: n_img images of W x H (rows W bytes
// apart, images `stride` bytes apart) and their taps. A launch walks the
// global tile list [0, ntiles); a job owns tiles [tile0, tile0 + n_img *
// tiles_x * tiles_y).
constexpr int kFxMaxL = 33;
struct FxJob {
    const uint8_t *src;
    uint8_t *dst;
    long long stride;
    int W, H, n_img, tiles_x, tiles_y, tile0;
    int L, m;                    // taps, window shift (the staged columns start 4-aligned)
    int ch, cv;                  // readout margins c (units of 2^-kFxShift) of the two passes
    float init_h, init_v;        // hi accumulator start values (integers) of the two passes
    float h[kFxMaxL], v[kFxMaxL];
};

// Job accessors: one job passed by value, or a device table.
struct FxOneJob {
    const FxJob &j;  // the kernel argument itself (no private copy)
    __device__ __forceinline__ int count() const { return 1; }
    __device__ __forceinline__ const FxJob &operator[](int) const { return j; }
};
struct FxJobTable {
    const FxJob *p;
    int n;
    __device__ __forceinline__ int count() const { return n; }
    __device__ __forceinline__ const FxJob &operator[](int i) const { return p[i]; }
};

__device__ __forceinline__ v8h fx_a_operand(uint2 b) { return bytes_to_half8(b); }

// Toeplitz B operand (see build_b) with taps * 2^kFxShift split into an
// integer hi and an f16 lo.
template <int NKB>
__device__ __forceinline__ void build_b_fx(const float *taps, int lane, int shift,
                                           v8h (&bhi)[NKB], v8h (&blo)[NKB]) {
    const int n = lane & 31, g = lane >> 5;
#pragma unroll
    for (int kb = 0; kb < NKB; kb++) {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            float w = taps[kTapPad + 16 * kb + 8 * g + j - n - shift] * (float)(1 << kFxShift);
            _Float16 hi = (_Float16)rintf(w);
            _Float16 lo = (_Float16)(w - (float)hi);
            bhi[kb][j] = hi;
            blo[kb][j] = lo;
        }
    }
}

// Smallest fraction field of four accumulator bit patterns: one compare
// against the margin then tells whether any of the four needs the exact chain.
__device__ __forceinline__ uint32_t min_frac4(const uint32_t (&u)[4]) {
    return min(min(u[0] & kFxFracMask, u[1] & kFxFracMask),
               min(u[2] & kFxFracMask, u[3] & kFxFracMask));
}

// Queues the outputs among four whose fraction lies below the margin (entry
// e0 + i * step for element i). Taken rarely; the compares are redone here
// (the empty asm keeps the compiler from holding their masks across the
// caller's common path).
__device__ __forceinline__ void enqueue_flags4(uint32_t (&u)[4], uint32_t thr, uint32_t e0,
                                               uint32_t step, uint16_t *q, int &qn) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
        asm volatile("" : "+v"(u[i]));
        const bool f = (u[i] & kFxFracMask) < thr;
        enqueue_ballot(f, __ballot(f), e0 + i * step, q, qn);
    }
}

typedef unsigned short v2u16 __attribute__((ext_vector_type(2)));

// Bits kFxShift..kFxShift+7 of four accumulator bit patterns, packed into one
// dword: bytes 1..2 of two values form two 16-bit halves (bits 8..23), a
// packed shift by kFxShift - 8 leaves the byte alone in each half (the bits
// above it up to bit 22 are zero, and bit 23, the lowest exponent bit of a
// value in [2^23, 2^24), is zero too), and a last byte permute gathers the
// four bytes.
__device__ __forceinline__ uint32_t pack_bits15(uint32_t u0, uint32_t u1, uint32_t u2,
                                                uint32_t u3) {
    const uint32_t p01 = __builtin_amdgcn_perm(u1, u0, 0x06050201u);
    const uint32_t p23 = __builtin_amdgcn_perm(u3, u2, 0x06050201u);
    const v2u16 sh = {kFxShift - 8, kFxShift - 8};
    const v2u16 h01 = __builtin_bit_cast(v2u16, p01) >> sh;
    const v2u16 h23 = __builtin_bit_cast(v2u16, p23) >> sh;
    return __builtin_amdgcn_perm(__builtin_bit_cast(uint32_t, h23),
                                 __builtin_bit_cast(uint32_t, h01), 0x06040200u);
}

// Row-wise staging of the fixed-point kernel's input window: each wave loads
// RPW whole staged rows per step, CH lanes per row, 16 bytes per lane. Only
// the row index depends on the step, so a load costs a clamp and a multiply.
template <int NKB, int ROWS> struct RowStage {
    static constexpr int CIN = kTW - 32 + 16 * NKB;  // staged columns (multiple of 16)
    static constexpr int CH = CIN / 16;              // 16-byte chunks per row
    static constexpr int RPW = 64 / CH;              // rows per wave per step
    static constexpr int RPK = 4 * RPW;              // rows per workgroup step
    static constexpr int NLD = (ROWS + RPK - 1) / RPK;
    static constexpr int PIN = ((CIN / 4) % 2 == 0) ? CIN + 8 : CIN + 4;
    static constexpr int LMAX = 16 * NKB - 31;
};

// Issues this thread's loads of one tile's window (rows clamp(ys + r), columns
// xs + c; xs is a multiple of 4). INTERIOR tiles (the window lies inside the
// row) load 16 bytes at once; at the left and right image edges ALIGNED rows
// load whole dwords from the clamped column and store_rows replicates the edge
// byte; unaligned rows load each byte from its clamped column.
template <typename RS, bool ALIGNED>
__device__ __forceinline__ void load_rows(const uint8_t *__restrict__ img, int W, int H, int xs,
                                          int ys, bool interior, uint4 (&regs)[RS::NLD]) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int rr = lane / RS::CH, c0 = (lane - rr * RS::CH) * 16;
    if (rr >= RS::RPW) return;
    if (ALIGNED && interior) {
#pragma unroll
        for (int k = 0; k < RS::NLD; k++) {
            const int r = k * RS::RPK + wave * RS::RPW + rr;
            const uint32_t row = (uint32_t)clampi(ys + min(r, kRows2 - 1), 0, H - 1);
            regs[k] = *reinterpret_cast<const uint4 *>(img + (row * (uint32_t)W + (uint32_t)(xs + c0)));
        }
        return;
    }
#pragma unroll
    for (int k = 0; k < RS::NLD; k++) {
        const int r = k * RS::RPK + wave * RS::RPW + rr;
        const uint8_t *row = img + (size_t)clampi(ys + min(r, kRows2 - 1), 0, H - 1) * W;
        uint32_t d[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int gx = xs + c0 + 4 * q;
            if (ALIGNED) {
                d[q] = *reinterpret_cast<const uint32_t *>(row + clampi(gx, 0, W - 4));
            } else {
                d[q] = (uint32_t)row[clampi(gx, 0, W - 1)] |
                       ((uint32_t)row[clampi(gx + 1, 0, W - 1)] << 8) |
                       ((uint32_t)row[clampi(gx + 2, 0, W - 1)] << 16) |
                       ((uint32_t)row[clampi(gx + 3, 0, W - 1)] << 24);
            }
        }
        regs[k] = make_uint4(d[0], d[1], d[2], d[3]);
    }
}

template <typename RS, bool ALIGNED>
__device__ __forceinline__ void store_rows(uint8_t *in_s, int W, int xs, bool interior,
                                           const uint4 (&regs)[RS::NLD]) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int rr = lane / RS::CH, c0 = (lane - rr * RS::CH) * 16;
    if (rr >= RS::RPW) return;
    uint8_t *base = in_s + (wave * RS::RPW + rr) * RS::PIN + c0;
#pragma unroll
    for (int k = 0; k < RS::NLD; k++) {
        const int r = k * RS::RPK + wave * RS::RPW + rr;
        if (k == RS::NLD - 1 && r >= kRows2) break;
        uint32_t d[4] = {regs[k].x, regs[k].y, regs[k].z, regs[k].w};
        if (ALIGNED && !interior) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int gx = xs + c0 + 4 * q;
                const uint32_t left = (d[q] & 0xffu) * 0x01010101u;
                const uint32_t right = (d[q] >> 24) * 0x01010101u;
                d[q] = gx < 0 ? left : d[q];
                d[q] = gx >= W ? right : d[q];
            }
        }
        uint8_t *dst = base + k * RS::RPK * RS::PIN;
        *reinterpret_cast<uint2 *>(dst) = make_uint2(d[0], d[1]);
        *reinterpret_cast<uint2 *>(dst + 8) = make_uint2(d[2], d[3]);
    }
}

// Waves per SIMD the fixed-point kernel is compiled for: 3 (three workgroups
// per CU) where its registers fit in 168 without spilling, else 2.
constexpr int fx_waves(int nkb, int taps, bool aligned) {
    return (taps == 2 && nkb <= 3 && aligned) ? 3 : 2;
}

// TAPS: 0 = distinct horizontal and vertical taps; 1 = equal taps; 2 = equal
// taps with the rows staged under the same shift m as the columns (L + m <=
// 33 keeps the shifted rows inside the staged window), so both passes use the
// same B operands.
template <int NKB, int TAPS, bool ALIGNED, typename Jobs>
__device__ __forceinline__ void blur_u8_fx_body(const Jobs &jobs, int ntiles) {
    using S = RowStage<NKB, kRows2>;
    constexpr int PIN = S::PIN, LMAX = S::LMAX;
    constexpr int IN_BYTES = (kRows2 * PIN + 15) / 16 * 16;
    static_assert(kTH2 * kOPitch <= IN_BYTES, "output tile reuses the input window");
    extern __shared__ __align__(16) unsigned char smem[];
    uint8_t *in_s = smem;                     // kRows2 x PIN, then the output tile
    uint8_t *out_s = smem;                    // kTH2 x kOPitch (after the horizontal pass)
    uint8_t *ht = smem + IN_BYTES;            // kTW x kPitchB, column-major u8
    float *taps_h = reinterpret_cast<float *>(ht + kTW * kPitchB);
    float *taps_v = taps_h + (2 * kTapPad + kMaxTaps);
    uint16_t *queue = reinterpret_cast<uint16_t *>(taps_v + (2 * kTapPad + kMaxTaps)) +
                      (threadIdx.x >> 6) * kFxQCap;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n = lane & 31, g = lane >> 5;
    constexpr bool SHARE_B = TAPS == 2;

    // per-job state, rebuilt when a workgroup's tile belongs to another job
    int L = 0, half = 0, m = 0;
    uint32_t thr_h = 0, thr_v = 0;
    v8h bh_hi[NKB], bh_lo[NKB], bv_hi_[SHARE_B ? 1 : NKB], bv_lo_[SHARE_B ? 1 : NKB];
    v16f init_hv, init_vv_;
    auto &bv_hi = *reinterpret_cast<v8h(*)[NKB]>(SHARE_B ? (void *)bh_hi : (void *)bv_hi_);
    auto &bv_lo = *reinterpret_cast<v8h(*)[NKB]>(SHARE_B ? (void *)bh_lo : (void *)bv_lo_);
    const float *vts = TAPS != 0 ? taps_h : taps_v;
    const v16f &init_vv = TAPS == 0 ? init_vv_ : init_hv;  // equal taps: equal start values
    // every thread: the job's zero-padded taps into LDS (a barrier follows)
    auto stage_taps = [&](const FxJob &jb) {
        if (threadIdx.x < 2 * kTapPad + kMaxTaps) {
            int t = (int)threadIdx.x - kTapPad;
            bool in = t >= 0 && t < jb.L;
            taps_h[threadIdx.x] = in ? jb.h[t] : 0.0f;
            taps_v[threadIdx.x] = in ? jb.v[t] : 0.0f;
        }
    };
    // after the barrier: this lane's B operands and start vectors
    auto setup_job = [&](const FxJob &jb) {
        L = jb.L;
        half = L / 2;
        m = jb.m;
        thr_h = 2u * (uint32_t)jb.ch;
        thr_v = 2u * (uint32_t)jb.cv;
        build_b_fx<NKB>(taps_h, lane, m, bh_hi, bh_lo);
        if constexpr (!SHARE_B) {
            build_b_fx<NKB>(taps_v, lane, 0, bv_hi_, bv_lo_);
        }
        // hi accumulator start vectors, held in VGPRs (the empty asm keeps
        // the compiler from re-splatting them before every block)
#pragma unroll
        for (int r = 0; r < 16; r++) {
            float th = jb.init_h, tv = jb.init_v;
            asm volatile("" : "+v"(th));
            init_hv[r] = th;
            if constexpr (TAPS == 0) {
                asm volatile("" : "+v"(tv));
                init_vv_[r] = tv;
            }
        }
    };
    // the job owning tile t, scanning forward from job j
    auto job_of = [&](int t, int j) {
        while (j + 1 < jobs.count() && t >= jobs[j + 1].tile0) j++;
        return j;
    };

    const int per_xcd = (int)gridDim.x >> 3;
    const int chunk = (ntiles + 7) >> 3;
    const int t_begin = (int)(blockIdx.x & 7) * chunk;
    const int t_end = min(ntiles, t_begin + chunk);
    int t = t_begin + (int)(blockIdx.x >> 3);

    // a window lies inside its rows when xs >= 0 and xs + CIN <= W
    uint4 regs[S::NLD];
    bool interior = false;
    int jc = 0, jl = -1;  // job of the current tile; job whose state is loaded
    auto issue_loads = [&](int tt, int j) {
        const FxJob &jb = jobs[j];
        const int per_img = jb.tiles_x * jb.tiles_y;
        const int lt = tt - jb.tile0;
        const int im = lt / per_img, rem = lt - im * per_img, ty = rem / jb.tiles_x;
        // shared B operands: the rows are staged with the same shift m as the
        // columns, so both passes use one Toeplitz matrix
        const int hf = jb.L / 2, mm = jb.m;
        const int xs = (rem - ty * jb.tiles_x) * kTW - hf - mm;
        interior = xs >= 0 && xs + S::CIN <= jb.W;
        load_rows<S, ALIGNED>(jb.src + (size_t)im * jb.stride, jb.W, jb.H, xs,
                              ty * kTH2 - hf - (SHARE_B ? mm : 0), interior, regs);
    };
    if (t < t_end) {
        jc = job_of(t, 0);
        issue_loads(t, jc);
    }
    for (; t < t_end; t += per_xcd) {
        const FxJob &jb = jobs[jc];
        const int per_img = jb.tiles_x * jb.tiles_y;
        const int lt = t - jb.tile0;
        const int im = lt / per_img, rem = lt - im * per_img, ty = rem / jb.tiles_x;
        const int x0 = (rem - ty * jb.tiles_x) * kTW, y0 = ty * kTH2;
        const int W = jb.W, H = jb.H;
        uint8_t *out = jb.dst + (size_t)im * jb.stride;
        const bool new_job = jc != jl;

        __syncthreads();  // the previous tile is done with in_s / out_s, ht and the taps
        if (new_job) stage_taps(jb);
        store_rows<S, ALIGNED>(in_s, W, x0 - jb.L / 2 - jb.m, interior, regs);
        const int tn = t + per_xcd;
        int jn = jc;
        if (tn < t_end) {
            jn = job_of(tn, jc);
            issue_loads(tn, jn);
        }
        __syncthreads();
        if (new_job) {
            setup_job(jb);
            jl = jc;
        }

        // ---- horizontal pass: (kRows2 / 32) x 4 blocks of 32x32 -> ht (column-major u8);
        // the A operand of a wave's next block is read while the current one is read out.
        int qn = 0;
        auto h_load = [&](int blk, uint2 (&raw)[NKB]) {
            const int rb = blk >> 2, cb = blk & 3;
            const uint8_t *arow = in_s + (32 * rb + n) * PIN + 32 * cb + 8 * g;
#pragma unroll
            for (int kb = 0; kb < NKB; kb++) raw[kb] = *reinterpret_cast<const uint2 *>(arow + 16 * kb);
        };
        auto h_mma = [&](const uint2 (&raw)[NKB], v16f &acc, v16f &acl) {
            const v16f zero = {};
#pragma unroll
            for (int kb = 0; kb < NKB; kb++) {
                v8h av = fx_a_operand(raw[kb]);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(av, bh_hi[kb], kb == 0 ? init_hv : acc, 0, 0, 0);
                acl = __builtin_amdgcn_mfma_f32_32x32x16_f16(av, bh_lo[kb], kb == 0 ? zero : acl, 0, 0, 0);
            }
        };
        auto h_drain = [&]() {
            for (int e = lane; e < qn; e += 64) {
                const uint32_t ent = queue[e];
                const int row = (int)(ent >> 8), cl = (int)(ent & 255);
                float r = exact_chain8<LMAX>(in_s + row * PIN + m + cl, taps_h, L);
                ht[cl * kPitchB + row] = clamp_u8_exact(r);
            }
            qn = 0;
        };
        // lane holds column 32*cb + n, rows 32*rb + 8*(r/4) + 4*g + r%4
        auto h_out = [&](int blk, const v16f &acc, const v16f &acl) {
            const int rb = blk >> 2, cb = blk & 3;
            uint8_t *col = ht + (32 * cb + n) * kPitchB + 32 * rb + 4 * g;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                uint32_t u[4];
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    // (a bit_cast of a vector element reads element 0 with this compiler)
                    u[i] = __float_as_uint(acc[4 * j + i] + acl[4 * j + i]);
                }
                *reinterpret_cast<uint32_t *>(col + 8 * j) = pack_bits15(u[0], u[1], u[2], u[3]);
                if (__ballot(min_frac4(u) < thr_h)) {
                    const uint32_t e0 = (uint32_t)((32 * rb + 8 * j + 4 * g) << 8) | (32 * cb + n);
                    enqueue_flags4(u, thr_h, e0, 1u << 8, queue, qn);
                }
                if ((j & 1) && qn > kFxQDrain) h_drain();
            }
        };
        {
            constexpr int NB = (kRows2 / 32) * (kTW / 32) / 4;  // blocks per wave
            uint2 raw[NKB];
            h_load(wave, raw);
#pragma unroll 1
            for (int i = 0; i < NB; i++) {
                v16f acc, acl;
                h_mma(raw, acc, acl);
                if (i + 1 < NB) h_load(wave + 4 * (i + 1), raw);
                h_out(wave + 4 * i, acc, acl);
            }
        }
        h_drain();
        __syncthreads();

        // ---- vertical pass: D[m = x][n = y] = sum_k H^T[x][k] B[k][y] -> out_s
        qn = 0;
        auto v_load = [&](int blk, uint2 (&raw)[NKB]) {
            const int rb = blk >> 2, cb = blk & 3;
            const uint8_t *acol = ht + (32 * cb + n) * kPitchB + 32 * rb + 8 * g;
#pragma unroll
            for (int kb = 0; kb < NKB; kb++) raw[kb] = *reinterpret_cast<const uint2 *>(acol + 16 * kb);
        };
        auto v_mma = [&](const uint2 (&raw)[NKB], v16f &acc, v16f &acl) {
            const v16f zero = {};
#pragma unroll
            for (int kb = 0; kb < NKB; kb++) {
                v8h av = fx_a_operand(raw[kb]);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(av, bv_hi[kb], kb == 0 ? init_vv : acc, 0, 0, 0);
                acl = __builtin_amdgcn_mfma_f32_32x32x16_f16(av, bv_lo[kb], kb == 0 ? zero : acl, 0, 0, 0);
            }
        };
        auto v_drain = [&]() {
            for (int e = lane; e < qn; e += 64) {
                const uint32_t ent = queue[e];
                const int y2 = (int)(ent >> 8), x2 = (int)(ent & 255);
                float r = exact_chain8<LMAX>(ht + x2 * kPitchB + y2 + (SHARE_B ? m : 0), vts, L);
                out_s[y2 * kOPitch + x2] = clamp_u8_exact(r);
            }
            qn = 0;
        };
        // lane holds output row 32*rb + n, columns 32*cb + 8*(r/4) + 4*g + r%4
        auto v_out = [&](int blk, const v16f &acc, const v16f &acl) {
            const int rb = blk >> 2, cb = blk & 3;
            const int yl = 32 * rb + n;
            uint8_t *orow = out_s + yl * kOPitch;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int xl = 32 * cb + 8 * j + 4 * g;
                uint32_t u[4];
#pragma unroll
                for (int i = 0; i < 4; i++) u[i] = __float_as_uint(acc[4 * j + i] + acl[4 * j + i]);
                *reinterpret_cast<uint32_t *>(orow + xl) = pack_bits15(u[0], u[1], u[2], u[3]);
                if (__ballot(min_frac4(u) < thr_v)) {
                    const uint32_t e0 = (uint32_t)(yl << 8) | (uint32_t)xl;
                    enqueue_flags4(u, thr_v, e0, 1u, queue, qn);
                }
                if ((j & 1) && qn > kFxQDrain) v_drain();
            }
        };
        {
            constexpr int NB = (kTH2 / 32) * (kTW / 32) / 4;  // blocks per wave
            uint2 raw[NKB];
            v_load(wave, raw);
#pragma unroll 1
            for (int i = 0; i < NB; i++) {
                v16f acc, acl;
                v_mma(raw, acc, acl);
                if (i + 1 < NB) v_load(wave + 4 * (i + 1), raw);
                v_out(wave + 4 * i, acc, acl);
            }
        }
        v_drain();
        __syncthreads();

        // ---- copy the output tile out: 8 lanes x 16 bytes per row
        const bool vec_ok = ((W & 15) == 0) && ((reinterpret_cast<uintptr_t>(out) & 15) == 0);
#pragma unroll 2
        for (int i = threadIdx.x; i < kTH2 * (kTW / 16); i += 256) {
            const int yl = i >> 3, xl = (i & 7) * 16;
            const int gy = y0 + yl, gx = x0 + xl;
            if (gy >= H || gx >= W) continue;
            const uint4 v = *reinterpret_cast<const uint4 *>(out_s + yl * kOPitch + xl);
            uint8_t *o = out + (size_t)gy * W + gx;
            if (vec_ok && gx + 15 < W) {
                *reinterpret_cast<uint4 *>(o) = v;
            } else {
                const uint8_t *b = out_s + yl * kOPitch + xl;
                for (int q = 0; q < 16 && gx + q < W; q++) o[q] = b[q];
            }
        }
        jc = jn;
    }
}

template <int NKB, int TAPS, bool ALIGNED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(fx_waves(NKB, TAPS, ALIGNED), fx_waves(NKB, TAPS, ALIGNED)))) void blur_u8_fx_kernel(
    FxJob job) {
    FxOneJob one{job};
    blur_u8_fx_body<NKB, TAPS, ALIGNED>(one, job.n_img * job.tiles_x * job.tiles_y);
}

// The same kernel over a device table of jobs: images of different sizes and
// taps that share NKB, the tap class and row alignment.
template <int NKB, int TAPS, bool ALIGNED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(fx_waves(NKB, TAPS, ALIGNED), fx_waves(NKB, TAPS, ALIGNED)))) void blur_u8_fx_jobs_kernel(
    const FxJob *__restrict__ jobs, int njobs, int ntiles) {
    FxJobTable tab{jobs, njobs};
    blur_u8_fx_body<NKB, TAPS, ALIGNED>(tab, ntiles);
}

// |E - R| bound for one pass, in output units. E: MFMA estimate, at most
// 2L non-zero products plus 2*NKB accumulations, each rounded (toward zero
// in the worst case, 2^-23 relative) at a magnitude <= 255 * sum|w|, plus
// the 2^-22 relative residual of the hi/lo tap split. R: the reference's
// chain, L products and L additions rounded to nearest (2^-24 relative).
float blur_eps(const float *w, int L, int nkb) {
    double sabs = 0.0;
    for (int i = 0; i < L; i++) sabs += std::fabs((double)w[i]);
    double mag = 255.0 * sabs;
    double e = mag * (std::ldexp(1.0, -23) * (2.0 * L + 2.0 * nkb) + std::ldexp(1.0, -22) +
                      std::ldexp(1.0, -24) * (2.0 * L));
    return (float)(e * 1.25 + 1e-7);
}

// Readout margin c of the fixed-point kernel (units of 2^-kFxShift) and its
// accumulator start values, or c = -1 when the taps do not qualify. c bounds
// |E - R| in units, E the kernel's estimate and R the reference's f32 result:
//   * the hi accumulator is exact; it starts at 2^23 + c; every partial sum
//     of the lo accumulator (started at 0) stays below B_lo = max(A) * sum|lo|, so each
//     of its at most 2L + 2*NKB roundings costs at most B_lo * 2^-23; the
//     final f32 add rounds once (half a unit);
//   * the split leaves |w * 2^kFxShift - hi - lo| per tap, times a pixel
//     <= 255 (an f16 lo below 2^-14 counts as lost entirely, should it be
//     flushed);
//   * the reference's chain: product i is rounded once and then passes
//     through L - max(i, 1) rounded additions (the first addition, 0 + x0,
//     is exact), each off by at most 2^-24 relative, all terms >= 0 and
//     pixels <= 255 (1 % slack for second-order terms); or, where smaller,
//     the sum of the half-ulps of every rounded value's largest possible
//     magnitude: product i is at most 255 w_i and partial sum k at most
//     255 (w_0 + ... + w_k), and a round-to-nearest result is off by at
//     most half an ulp of any bound on its magnitude.
// c exceeds the bound by 25 % plus one unit.
struct FxPlan {
    int c;
    float init;
};

FxPlan blur_fx_plan(const float *w, int L, int nkb) {
    const double unit = std::ldexp(1.0, kFxShift);
    const double amax = 255.0;  // largest A operand
    double sum = 0.0, sum_hi = 0.0, lo_abs = 0.0, split = 0.0, chain = 0.0, chain_ulp = 0.0;
    // half an ulp of every f32 value of magnitude at most b (b > 0)
    auto half_ulp = [](double b) {
        int ex = 0;
        std::frexp(b, &ex);  // b in [2^(ex-1), 2^ex): ulp 2^(ex-24)
        return std::ldexp(1.0, ex - 25);
    };
    for (int i = 0; i < L; i++) {
        if (!(w[i] >= 0.0f) || !std::isfinite(w[i])) return {-1, 0.0f};
        const float ws = w[i] * (float)unit;
        const _Float16 hi = (_Float16)rintf(ws);
        const _Float16 lo = (_Float16)(ws - (float)hi);
        sum += (double)w[i];
        sum_hi += (double)(float)hi;
        lo_abs += std::fabs((double)(float)lo);
        double res = std::fabs((double)ws - (double)(float)hi - (double)(float)lo);
        if (std::fabs((double)(float)lo) < std::ldexp(1.0, -14)) res += std::fabs((double)(float)lo);
        split += res;
        chain += (double)w[i] * (1.0 + (double)(L - std::max(i, 1)));
        const double slack = 1.0 + 1e-5;  // rounded terms exceed exact ones by < 34 * 2^-24
        if (w[i] > 0.0f) chain_ulp += half_ulp(amax * (double)w[i] * slack);
        if (i >= 1 && sum > 0.0) chain_ulp += half_ulp(amax * sum * slack);
    }
    const double b_lo = amax * lo_abs;
    const double chain_err =
        unit * std::min(255.0 * chain * std::ldexp(1.0, -24) * 1.01, chain_ulp);
    double e = 0.5 + (2.0 * L + 2.0 * nkb) * b_lo * std::ldexp(1.0, -23) + 255.0 * split +
               chain_err;
    const double c = std::ceil(1.25 * e + 1.0);
    const double init_hi = 8388608.0 + c;
    // the hi accumulator is exact: every partial sum is an integer of
    // magnitude below 2^24; the final value stays in [2^23, 2^24) with the
    // estimate's integer part in 8 bits
    if (std::fabs(init_hi) + amax * sum_hi >= 16777216.0) return {-1, 0.0f};
    if (255.0 * std::max(sum_hi + lo_abs, sum * unit * (1.0 + 1e-6)) + 2.0 * c >= 255.99 * unit)
        return {-1, 0.0f};
    if (c >= unit / 8.0) return {-1, 0.0f};
    return {(int)c, (float)init_hi};
}

// ---- streaming fixed-point kernel ---------------------------------------------------
//
// Each workgroup owns a 128-column strip of one image over a band of rows and
// walks down it 32 output rows per step. A step's horizontal pass computes the
// 32 intermediate rows that become available (each intermediate row exactly
// once per strip) into a 64-row ring of f16 columns; its vertical pass reads
// the 16 * NKB ring rows that its 32 output rows need. The input window of
// the next step is loaded into registers while the current step is computed,
// and the strips of one image go to one XCD, so a strip's halo columns are
// read from the L2 its neighbours just filled.
//
// Readout: the fixed-point scheme of blur_u8_fx_kernel (blur_fx_plan): an
// exact hi accumulator started at 2^23 + c, a lo accumulator started at 0,
// one f32 add; bits 15..22 of the sum are the reference's byte whenever bits
// 0..14 are at least 2c, else the output is recomputed with the reference's
// chain from the same LDS data.
constexpr int kSR = 32;     // output rows per step
constexpr int kSW = 128;    // strip width (output columns)
constexpr int kRing = 64;   // intermediate rows held (2*half + m + d + 32 <= 64 for L <= 33)
constexpr int kSOP = kSW + 16;  // bytes per output-step row in LDS

template <int NKB> struct StreamShape {
    static constexpr int CIN = kSW - 32 + 16 * NKB;  // staged input columns (multiple of 16)
    static constexpr int CH = CIN / 16;              // 16-byte chunks per staged row
    static constexpr int NCH = kSR * CH;             // chunks per step
    static constexpr int NLD = (NCH + 255) / 256;    // chunk loads per thread
    // f16 row pitch in dwords = 4 * odd: the 8 rows one 16-byte A-operand read
    // of 8 lanes touches fall in distinct bank groups
    static constexpr int PIN = CIN * 2 + 16;         // bytes per staged f16 row
    static constexpr int LMAX = 16 * NKB - 31;
};
constexpr int kRingPitch = kRing * 2 + 16;  // bytes per ring column (36 dwords)

struct StreamJob {
    const uint8_t *src;
    uint8_t *dst;
    long long stride;
    int W, H, n_img, strips, bands, band_rows;
    int L, m, d;                 // taps, window shift, horizontal run-ahead rows
    int ch, cv;                  // readout margins (units of 2^-15)
    float init_h, init_v;        // accumulator start values 2^23 + c
    float h[kFxMaxL], v[kFxMaxL];
};

// 16 u8 pixels -> 16 exact f16, written as 32 bytes.
__device__ __forceinline__ void store_half16(_Float16 *dst, uint4 b) {
    const v8h lo = bytes_to_half8(make_uint2(b.x, b.y));
    const v8h hi = bytes_to_half8(make_uint2(b.z, b.w));
    *reinterpret_cast<v8h *>(dst) = lo;
    *reinterpret_cast<v8h *>(dst + 8) = hi;
}

// Bits 15..22 of two accumulator bit patterns as two exact f16 values.
__device__ __forceinline__ uint32_t half2_bits15(uint32_t u0, uint32_t u1) {
    const uint32_t p = __builtin_amdgcn_perm(u1, u0, 0x06050201u);
    const v2u16 sh = {7, 7};
    const uint32_t b = __builtin_bit_cast(uint32_t, __builtin_bit_cast(v2u16, p) >> sh);
    const v2h f = __builtin_bit_cast(v2h, b | 0x64006400u);
    const v2h off = {(_Float16)-1024.0f, (_Float16)-1024.0f};
    return __builtin_bit_cast(uint32_t, f + off);
}

// The reference's chain down one ring column: rows (base + q) mod kRing.
template <int LMAX>
__device__ __forceinline__ float ring_chain(const _Float16 *col, int base, const float *taps,
                                            int L) {
    float s = 0.0f;
#pragma unroll
    for (int q = 0; q < LMAX; q++)
        if (q < L) s = __fadd_rn(s, __fmul_rn((float)col[(base + q) & (kRing - 1)], taps[kTapPad + q]));
    return s;
}

template <int NKB, bool SHARE_B, bool ALIGNED>
__device__ __forceinline__ void blur_u8_stream_body(const StreamJob &jb) {
    using S = StreamShape<NKB>;
    constexpr int PIN = S::PIN, LMAX = S::LMAX;
    constexpr int PINH = PIN / 2, RPH = kRingPitch / 2;  // pitches in f16 elements
    extern __shared__ __align__(16) unsigned char smem[];
    _Float16 *in_s = reinterpret_cast<_Float16 *>(smem);                       // kSR x PIN
    _Float16 *ring = reinterpret_cast<_Float16 *>(smem + kSR * PIN);           // kSW x kRingPitch
    uint8_t *out_s = smem + kSR * PIN + kSW * kRingPitch;                      // kSR x kSOP
    float *taps_h = reinterpret_cast<float *>(out_s + kSR * kSOP);
    float *taps_v = taps_h + (2 * kTapPad + kMaxTaps);
    uint16_t *queue = reinterpret_cast<uint16_t *>(taps_v + (2 * kTapPad + kMaxTaps)) +
                      (threadIdx.x >> 6) * kFxQCap;

    // unit = (image, band, strip), strip fastest; contiguous eighths of the
    // unit list go to the eight XCDs (workgroups are dealt round-robin)
    const int units = jb.n_img * jb.bands * jb.strips;
    const int chunk = (units + 7) >> 3;
    const int u = (int)(blockIdx.x & 7) * chunk + (int)(blockIdx.x >> 3);
    if ((int)(blockIdx.x >> 3) >= chunk || u >= units) return;
    const int per_img = jb.bands * jb.strips;
    const int im = u / per_img, rem = u - im * per_img;
    const int band = rem / jb.strips, strip = rem - band * jb.strips;
    const int W = jb.W, H = jb.H, L = jb.L, m = jb.m, half = L / 2;
    const int x0 = strip * kSW, yb0 = band * jb.band_rows;
    const int yb1 = min(H, yb0 + jb.band_rows);
    const int nsteps = (yb1 - yb0 + kSR - 1) / kSR;
    const uint8_t *src = jb.src + (size_t)im * jb.stride;
    uint8_t *dst = jb.dst + (size_t)im * jb.stride;
    const int xs = x0 - half - m;               // first staged column (multiple of 4)
    const bool interior = ALIGNED && xs >= 0 && xs + S::CIN <= W;
    const int ahead = half + jb.d;              // H(s) makes intermediate rows y(s) + ahead ..
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n = lane & 31, g = lane >> 5;

    // taps (zero padded), and a zeroed ring: rows the vertical pass reads
    // with zero weight must hold finite values
    if (threadIdx.x < 2 * kTapPad + kMaxTaps) {
        const int t = (int)threadIdx.x - kTapPad;
        const bool in = t >= 0 && t < L;
        taps_h[threadIdx.x] = in ? jb.h[t] : 0.0f;
        taps_v[threadIdx.x] = in ? jb.v[t] : 0.0f;
    }
    for (int i = threadIdx.x; i < kSW * kRingPitch / 16; i += 256)
        reinterpret_cast<uint4 *>(ring)[i] = make_uint4(0, 0, 0, 0);

    // input rows of step s: y(s) + ahead + r, r < kSR (clamped to the image)
    uint4 regs[S::NLD];
    auto issue_loads = [&](int s) {
        const int ybase = yb0 + kSR * s + ahead;
#pragma unroll
        for (int k = 0; k < S::NLD; k++) {
            const int c = min((int)threadIdx.x + 256 * k, S::NCH - 1);
            const int r = c / S::CH, q = c - r * S::CH;
            // 32-bit offsets from the image's (uniform) base: W * H < 2^32
            const uint32_t row = (uint32_t)clampi(ybase + r, 0, H - 1) * (uint32_t)W;
            const int gx = xs + 16 * q;
            if (interior) {
                regs[k] = *reinterpret_cast<const uint4 *>(src + (row + (uint32_t)gx));
            } else {
                uint32_t dw[4];
#pragma unroll
                for (int p = 0; p < 4; p++) {
                    const int x4 = gx + 4 * p;
                    if (ALIGNED) {
                        dw[p] = *reinterpret_cast<const uint32_t *>(
                            src + (row + (uint32_t)clampi(x4, 0, W - 4)));
                    } else {
                        dw[p] = (uint32_t)src[row + (uint32_t)clampi(x4, 0, W - 1)] |
                                ((uint32_t)src[row + (uint32_t)clampi(x4 + 1, 0, W - 1)] << 8) |
                                ((uint32_t)src[row + (uint32_t)clampi(x4 + 2, 0, W - 1)] << 16) |
                                ((uint32_t)src[row + (uint32_t)clampi(x4 + 3, 0, W - 1)] << 24);
                    }
                }
                regs[k] = make_uint4(dw[0], dw[1], dw[2], dw[3]);
            }
        }
    };
    // the loaded window into in_s as f16; at the image edges, aligned
    // dwords left of the row become its first byte, right of it its last
    auto store_rows = [&]() {
#pragma unroll
        for (int k = 0; k < S::NLD; k++) {
            const int c = (int)threadIdx.x + 256 * k;
            if (S::NCH % 256 != 0 && c >= S::NCH) break;
            const int r = c / S::CH, q = c - r * S::CH;
            uint4 v = regs[k];
            if (ALIGNED && !interior) {
                uint32_t dw[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int p = 0; p < 4; p++) {
                    const int x4 = xs + 16 * q + 4 * p;
                    const uint32_t left = (dw[p] & 0xffu) * 0x01010101u;
                    const uint32_t right = (dw[p] >> 24) * 0x01010101u;
                    dw[p] = x4 < 0 ? left : dw[p];
                    dw[p] = x4 >= W ? right : dw[p];
                }
                v = make_uint4(dw[0], dw[1], dw[2], dw[3]);
            }
            store_half16(in_s + r * PINH + 16 * q, v);
        }
    };

    issue_loads(-1);
    __syncthreads();  // taps staged
    v8h bh_hi[NKB], bh_lo[NKB], bv_hi_[SHARE_B ? 1 : NKB], bv_lo_[SHARE_B ? 1 : NKB];
    build_b_fx<NKB>(taps_h, lane, m, bh_hi, bh_lo);
    if constexpr (!SHARE_B) build_b_fx<NKB>(taps_v, lane, m, bv_hi_, bv_lo_);
    auto &bv_hi = *reinterpret_cast<v8h(*)[NKB]>(SHARE_B ? (void *)bh_hi : (void *)bv_hi_);
    auto &bv_lo = *reinterpret_cast<v8h(*)[NKB]>(SHARE_B ? (void *)bh_lo : (void *)bv_lo_);
    const float *vts = SHARE_B ? taps_h : taps_v;
    const uint32_t thr_h = 2u * (uint32_t)jb.ch, thr_v = 2u * (uint32_t)jb.cv;
    v16f init_h, init_v_;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        float a = jb.init_h, b = jb.init_v;
        asm volatile("" : "+v"(a));
        init_h[r] = a;
        if constexpr (!SHARE_B) {
            asm volatile("" : "+v"(b));
            init_v_[r] = b;
        }
    }
    const v16f &init_v = SHARE_B ? init_h : init_v_;
    store_rows();
    if (nsteps > 0) issue_loads(0);

    // the output rows y .. y + 31 of a finished step, from out_s; 32-bit
    // offsets from the image base (W * H < 2^32)
    const bool vec_out = ((W & 15) == 0) && ((reinterpret_cast<uintptr_t>(dst) & 15) == 0);
    auto copy_out = [&](int y) {
        const int yl = threadIdx.x >> 3, xl = (threadIdx.x & 7) * 16;
        const int gy = y + yl, gx = x0 + xl;
        if (gy < yb1 && gx < W) {
            const uint4 v = *reinterpret_cast<const uint4 *>(out_s + yl * kSOP + xl);
            const uint32_t o = (uint32_t)gy * (uint32_t)W + (uint32_t)gx;
            if (vec_out && gx + 16 <= W) {
                *reinterpret_cast<uint4 *>(dst + o) = v;
            } else {
                const uint8_t *b = out_s + yl * kSOP + xl;
                for (int q = 0; q < 16 && gx + q < W; q++) dst[o + q] = b[q];
            }
        }
    };
    // exact recomputation of the queued outputs an integer boundary could flip
    int qn = 0;
    auto h_drain = [&](int posh) {
        for (int e = lane; e < qn; e += 64) {
            const uint32_t ent = queue[e];
            const int row = (int)(ent >> 8), cl = (int)(ent & 255);
            const float r = exact_chain8<LMAX>(in_s + row * PINH + m + cl, taps_h, L);
            ring[cl * RPH + ((posh + row) & (kRing - 1))] = (_Float16)(float)clamp_u8_exact(r);
        }
        qn = 0;
    };
    auto v_drain = [&](int pos0) {
        for (int e = lane; e < qn; e += 64) {
            const uint32_t ent = queue[e];
            const int yl = (int)(ent >> 8), x2 = (int)(ent & 255);
            const float r = ring_chain<LMAX>(ring + x2 * RPH, pos0 + yl + m, vts, L);
            out_s[yl * kSOP + x2] = clamp_u8_exact(r);
        }
        qn = 0;
    };
    for (int s = -1; s < nsteps; s++) {
        const int y = yb0 + kSR * s;
        __syncthreads();  // in_s holds H(s)'s rows; out_s holds V(s-1)

        // ---- copy V(s-1) out: one 16-byte row chunk per thread
        if (s >= 1) copy_out(y - kSR);

        // ---- horizontal pass: rows y + ahead .. + 31 into the ring (wave = column block)
        {
            const int posh = (y + ahead + half + m) & (kRing - 1);  // a multiple of 4
            const _Float16 *arow = in_s + n * PINH + 32 * wave + 8 * g;
            v8h a[NKB];
#pragma unroll
            for (int kb = 0; kb < NKB; kb++) a[kb] = *reinterpret_cast<const v8h *>(arow + 16 * kb);
            v16f acc, acl;
            const v16f zero = {};
#pragma unroll
            for (int kb = 0; kb < NKB; kb++) {
                acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[kb], bh_hi[kb], kb == 0 ? init_h : acc, 0, 0, 0);
                acl = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[kb], bh_lo[kb], kb == 0 ? zero : acl, 0, 0, 0);
            }
            // lane holds column 32*wave + n, step rows 8*(r/4) + 4*g + r%4
            _Float16 *col = ring + (32 * wave + n) * RPH;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                uint32_t uu[4];
#pragma unroll
                for (int i = 0; i < 4; i++) uu[i] = __float_as_uint(acc[4 * j + i] + acl[4 * j + i]);
                const uint2 pk = make_uint2(half2_bits15(uu[0], uu[1]), half2_bits15(uu[2], uu[3]));
                *reinterpret_cast<uint2 *>(col + ((posh + 8 * j + 4 * g) & (kRing - 1))) = pk;
                if (__ballot(min_frac4(uu) < thr_h)) {
                    const uint32_t e0 = (uint32_t)((8 * j + 4 * g) << 8) | (32 * wave + n);
                    enqueue_flags4(uu, thr_h, e0, 1u << 8, queue, qn);
                }
                // a j-group adds at most 256 entries: draining above kFxQDrain after
                // every second group keeps the queue within kFxQCap
                if ((j & 1) && qn > kFxQDrain) h_drain(posh);
            }
            h_drain(posh);
        }
        __syncthreads();  // the ring holds every row V(s) reads; in_s is free

        if (s + 1 < nsteps) {
            store_rows();
            if (s + 2 < nsteps) issue_loads(s + 2);
        }

        // ---- vertical pass: output rows y .. y + 31 into out_s (wave = column block)
        if (s >= 0) {
            const int pos0 = y & (kRing - 1);  // ring row of intermediate row y - half - m
            const _Float16 *acol = ring + (32 * wave + n) * RPH;
            v8h a[NKB];
#pragma unroll
            for (int kb = 0; kb < NKB; kb++)
                a[kb] = *reinterpret_cast<const v8h *>(acol + ((pos0 + 16 * kb + 8 * g) & (kRing - 1)));
            v16f acc, acl;
            const v16f zero = {};
#pragma unroll
            for (int kb = 0; kb < NKB; kb++) {
                acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[kb], bv_hi[kb], kb == 0 ? init_v : acc, 0, 0, 0);
                acl = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[kb], bv_lo[kb], kb == 0 ? zero : acl, 0, 0, 0);
            }
            // lane holds output row n, columns 32*wave + 8*(r/4) + 4*g + r%4
            uint8_t *orow = out_s + n * kSOP;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int xl = 32 * wave + 8 * j + 4 * g;
                uint32_t uu[4];
#pragma unroll
                for (int i = 0; i < 4; i++) uu[i] = __float_as_uint(acc[4 * j + i] + acl[4 * j + i]);
                *reinterpret_cast<uint32_t *>(orow + xl) = pack_bits15(uu[0], uu[1], uu[2], uu[3]);
                if (__ballot(min_frac4(uu) < thr_v)) {
                    const uint32_t e0 = (uint32_t)(n << 8) | (uint32_t)xl;
                    enqueue_flags4(uu, thr_v, e0, 1u, queue, qn);
                }
                if ((j & 1) && qn > kFxQDrain) v_drain(pos0);
            }
            v_drain(pos0);
        }
    }
    __syncthreads();
    if (nsteps > 0) copy_out(yb0 + kSR * (nsteps - 1));
}

// The job is read through a reference to the kernel argument: its taps are
// indexed per thread, which must not copy the argument into registers.
template <int NKB, bool SHARE_B, bool ALIGNED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) void
blur_u8_stream_kernel(StreamJob job) {
    blur_u8_stream_body<NKB, SHARE_B, ALIGNED>(job);
}

constexpr size_t stream_lds(int nkb) {
    return (size_t)kSR * (2 * (kSW - 32 + 16 * nkb) + 16) + (size_t)kSW * kRingPitch +
           (size_t)kSR * kSOP + 2 * (2 * kTapPad + kMaxTaps) * sizeof(float) +
           4 * kFxQCap * sizeof(uint16_t);
}

iph_status launch_blur_stream(const uint8_t *src, uint8_t *dst, uint32_t w, uint32_t h,
                              uint32_t n, size_t stride, const float *hk, const float *vk,
                              int L, int m, int nkb, const FxPlan &ph, const FxPlan &pv,
                              bool same, bool aligned, hipStream_t s) {
    StreamJob jb{};
    jb.src = src;
    jb.dst = dst;
    jb.stride = (long long)stride;
    jb.W = (int)w;
    jb.H = (int)h;
    jb.n_img = (int)n;
    jb.strips = (int)((w + kSW - 1) / kSW);
    // bands of rows: enough units to fill the chip a few times over
    const long long per_band = (long long)n * jb.strips;
    const int steps_total = (int)((h + kSR - 1) / kSR);
    int bands = 1;
    if (per_band < 4096) {
        bands = (int)std::min<long long>(steps_total / 2 > 0 ? steps_total / 2 : 1,
                                         (4096 + per_band - 1) / per_band);
        bands = std::max(bands, 1);
    }
    const int steps_per_band = (steps_total + bands - 1) / bands;
    jb.band_rows = steps_per_band * kSR;
    jb.bands = (int)((h + jb.band_rows - 1) / jb.band_rows);
    const long long units = per_band * jb.bands;
    if (units == 0) return IPH_OK;
    if (units > (1LL << 30)) {
        set_error("batch of %u images of %ux%u is too large for one launch", n, w, h);
        return IPH_EINVAL;
    }
    jb.L = L;
    jb.m = m;
    jb.d = (4 - ((2 * (L / 2) + m) & 3)) & 3;
    jb.ch = ph.c;
    jb.cv = pv.c;
    jb.init_h = ph.init;
    jb.init_v = pv.init;
    for (int i = 0; i < L; i++) {
        jb.h[i] = hk[i];
        jb.v[i] = vk[i];
    }
    const unsigned grid = (unsigned)((units + 7) / 8 * 8);
#define SB(K, SAME, AL)                                                                      \
    blur_u8_stream_kernel<K, SAME, AL><<<grid, 256, stream_lds(K), s>>>(jb)
#define SA(K, SAME)                                                                          \
    if (aligned) SB(K, SAME, true); else SB(K, SAME, false)
#define SK(K)                                                                                \
    if (same) { SA(K, true); } else { SA(K, false); }                                        \
    break
    switch (nkb) {
    case 2: SK(2);
    case 3: SK(3);
    default: SK(4);
    }
#undef SK
#undef SA
#undef SB
    return launch_status("blur_u8_stream_kernel");
}


// ---- per-wave streaming kernel ------------------------------------------------------
//
// Each wave owns a 32-column strip of one image over a band of rows and walks
// down it 32 rows per step with no workgroup barrier: the columns its
// vertical pass reads are exactly the intermediate columns its own horizontal
// pass wrote, so a wave's LDS (its staged input rows, a 64-row ring of
// intermediate columns and its output rows) is private to it. The four waves
// of a workgroup take four neighbouring strips and the workgroups of one
// image go to one XCD, so the 16 halo columns a strip loads beside its own
// come from the L2 (or the L1) its neighbours filled.
//
// Step s of a wave (y = first output row of the band + 32 s):
//   V(s):   output rows y .. y + 31 from ring rows 32 (s & 1) + [0, 16 NKB)
//           (intermediate rows y - half - m + k), into the output rows in LDS,
//           then 16 bytes per lane to the image;
//   H(s+1): intermediate rows y + 64 - half - m + [0, 32) into ring slot
//           s & 1, from input rows loaded two steps ahead straight into
//           registers (converted to f16 in registers for the A operand, kept
//           as bytes in LDS for the exact recomputations).
// Ring offsets depend only on the step's parity: the loop is unrolled by two
// and every LDS address is a per-lane base plus a constant.
//
// Readout: the fixed-point scheme of blur_fx_plan (exact hi accumulator from
// 2^23 + c, lo accumulator from 0, one f32 add; bits 15..22 are the
// reference's byte whenever bits 0..14 are at least 2c). The start value
// enters through the last two K rows, which the taps leave zero (L + m + 33
// <= 16 NKB): A = 2048 and 1 there against hi B rows 4096 and c, so every
// partial sum of the hi accumulator is an integer below 2^24 and exact
// whatever the order of the MFMA's additions. One ballot per pass tells
// whether any output lies inside the margin; if one does, the wave re-reads
// its operands from LDS, recomputes the accumulators, and each lane
// recomputes its flagged outputs with the reference's chain and overwrites
// its estimates in LDS.
constexpr int kWaveOutPitch = 48;  // bytes per output row in LDS (16-byte aligned rows)

// staged input row pitch: 8 bytes past a multiple of 16 (18 dwords for
// NKB = 3), so the 8-byte A-operand reads of 32 rows spread over the banks
template <int NKB> constexpr int wave_in_pitch() { return 16 * NKB + 8 + (NKB == 2 ? 16 : 0); }
template <int NKB> constexpr int wave_lds_bytes() {
    return 32 * wave_in_pitch<NKB>() + 32 * kRingPitch + 32 * kWaveOutPitch;
}
// LDS shared by the workgroup's waves: the zero-padded taps and the
// Toeplitz tables (one, or two for distinct horizontal and vertical taps).
template <int NKB, bool SHARE_B> constexpr int wave_shared_lds() {
    return 2 * (2 * kTapPad + kMaxTaps) * (int)sizeof(float) + (SHARE_B ? 1 : 2) * 2 * NKB * 64 * 16;
}

struct WaveJob {
    const uint8_t *src;
    uint8_t *dst;
    long long stride;
    int W, H, n_img, strips, bands, band_rows;
    int L, m;                    // taps, window shift (the staged columns start 4-aligned)
    int ch, cv;                  // readout margins (units of 2^-15)
    int unit0;                   // a table job's first wave unit (a multiple of 4)
    float h[kFxMaxL], v[kFxMaxL];
};

// The reference's chain over taps t < L in chunks of four: a chunk's pixel
// and tap loads are issued together, then its products and sums run in
// order. Zero taps pad the last chunk: with non-negative taps every partial
// sum is >= +0, and adding +0 leaves it unchanged. `taps` is the
// zero-padded LDS copy (tap t at kTapPad + t); `p(t)` returns input t as
// f32 and must return a finite value for every t below L rounded up to 4.
// The chunk loop stays rolled: unrolled, its tap loads would be hoisted out
// of the caller's loop over flagged outputs into registers the strip loop
// needs.
template <typename F>
__device__ __forceinline__ float wave_chain(const float *taps, int L, F p) {
    float s = 0.0f;
#pragma unroll 1
    for (int t0 = 0; t0 < L; t0 += 4) {
        float x[4], w[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            x[t] = p(t0 + t);
            w[t] = taps[kTapPad + t0 + t];
        }
#pragma unroll
        for (int t = 0; t < 4; t++) s = __fadd_rn(s, __fmul_rn(x[t], w[t]));
    }
    return s;
}

__device__ __forceinline__ void wave_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// The accumulator start value in the last two K rows: B rows 4096 and c
// (exact f16 integers, c < 4096) in the hi operand, zero in the lo operand.
template <int NKB>
__device__ __forceinline__ void wave_b_init(int g, int c, v8h (&bhi)[NKB], v8h (&blo)[NKB]) {
    if (g == 1) {
        bhi[NKB - 1][6] = (_Float16)4096.0f;
        bhi[NKB - 1][7] = (_Float16)(float)c;
        blo[NKB - 1][6] = (_Float16)0.0f;
        blo[NKB - 1][7] = (_Float16)0.0f;
    }
}
// ... and their A values 2048 and 1 (f16 0x6800, 0x3c00) in the lanes of
// the upper K half of the last block.
template <int NKB>
__device__ __forceinline__ void wave_a_init(int g, v8h (&a)[NKB]) {
    typedef unsigned int v4u __attribute__((ext_vector_type(4)));
    v4u w = __builtin_bit_cast(v4u, a[NKB - 1]);
    w[3] = g ? 0x3c006800u : w[3];
    a[NKB - 1] = __builtin_bit_cast(v8h, w);
}

// The 2 NKB MFMAs of one pass. The Toeplitz operands (one 16-byte fragment
// per lane and K block, hi and lo) are built once per workgroup into LDS
// tables laid out [kb][hi, lo][lane] (wave_b_table).
template <int NKB>
__device__ __forceinline__ void wave_mfma_r(const v8h (&a)[NKB], const v8h (&bhi)[NKB],
                                            const v8h (&blo)[NKB], v16f &acc, v16f &acl) {
    const v16f zero = {};
#pragma unroll
    for (int kb = 0; kb < NKB; kb++) {
        acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[kb], bhi[kb], kb == 0 ? zero : acc, 0, 0, 0);
        acl = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[kb], blo[kb], kb == 0 ? zero : acl, 0, 0, 0);
    }
}

// One wave writes the Toeplitz table of one tap set (with the start value
// in its last two K rows) to LDS.
template <int NKB>
__device__ __forceinline__ void wave_b_table(const float *taps, int lane, int m, int c, v8h *bt) {
    v8h bhi[NKB], blo[NKB];
    build_b_fx<NKB>(taps, lane, m, bhi, blo);
    wave_b_init<NKB>(lane >> 5, c, bhi, blo);
#pragma unroll
    for (int kb = 0; kb < NKB; kb++) {
        bt[(2 * kb) * 64 + lane] = bhi[kb];
        bt[(2 * kb + 1) * 64 + lane] = blo[kb];
    }
}

// Strip modes: interior (the staged window lies inside 4-aligned rows),
// edge (4-aligned rows, window crossing the left or right image edge) and
// unaligned rows (every byte loaded from its clamped column). Each mode is
// its own inlined copy of the strip loop, so the rarer modes' registers do
// not weigh on the interior one.
enum { kWaveInterior = 0, kWaveEdge = 1, kWaveUnaligned = 2, kWaveInteriorVec = 3 };

template <int NKB, bool SHARE_B, int MODE>
__device__ __forceinline__ void blur_u8_wave_strip(const WaveJob &jb, const float *taps_h,
                                                   const float *taps_v, const v8h *bt_h,
                                                   const v8h *bt_v, uint8_t *in_w,
                                                   _Float16 *ring, uint8_t *out_w, int im,
                                                   int band, int strip) {
    constexpr int CIN = 16 * NKB, PI = wave_in_pitch<NKB>(), RPH = kRingPitch / 2;
    constexpr int PO = kWaveOutPitch;
    constexpr bool ALIGNED = MODE != kWaveUnaligned;
    constexpr bool interior = MODE == kWaveInterior || MODE == kWaveInteriorVec;
    // interior strips of images whose rows and dst are 16-aligned: 16-byte
    // stores only (the other store forms are not compiled into this copy)
    constexpr bool VEC = MODE == kWaveInteriorVec;
    const int L = jb.L, m = jb.m, half = L / 2;
    const int lane = threadIdx.x & 63;
    const int n = lane & 31, g = lane >> 5;
    const int W = jb.W, H = jb.H;
    const int x0 = strip * 32, yb0 = band * jb.band_rows;
    const int yb1 = min(H, yb0 + jb.band_rows);
    const int nsteps = (yb1 - yb0 + 31) >> 5;
    const uint8_t *src = jb.src + (size_t)im * jb.stride;
    uint8_t *dst = jb.dst + (size_t)im * jb.stride;
    const int xs = x0 - half - m;  // first staged column (a multiple of 4)
    const int yh0 = yb0 + 32 - half - m;  // first intermediate row of H(0)

    // zero ring: rows read with zero weight must hold finite values
    for (int i = lane; i < 32 * kRingPitch / 16; i += 64)
        reinterpret_cast<uint4 *>(ring)[i] = make_uint4(0, 0, 0, 0);

    const uint32_t thr_h = 2u * (uint32_t)jb.ch, thr_v = 2u * (uint32_t)jb.cv;
    // shared taps: the Toeplitz operands, from the workgroup's LDS table into
    // registers for the whole strip (two sets would not fit; different H and
    // V taps read them from LDS in each pass)
    constexpr bool BREG = SHARE_B;
    v8h rbh[BREG ? NKB : 1], rbl[BREG ? NKB : 1];
    if constexpr (BREG) {
#pragma unroll
        for (int kb = 0; kb < NKB; kb++) {
            rbh[kb] = bt_h[(2 * kb) * 64 + lane];
            rbl[kb] = bt_h[(2 * kb + 1) * 64 + lane];
        }
    }
    auto pass_mfma = [&](const v8h(&a)[NKB], const v8h *bt, v16f &acc, v16f &acl) {
        if constexpr (BREG) {
            wave_mfma_r<NKB>(a, rbh, rbl, acc, acl);
        } else {
            v8h bh[NKB], bl[NKB];
#pragma unroll
            for (int kb = 0; kb < NKB; kb++) {
                bh[kb] = bt[(2 * kb) * 64 + lane];
                bl[kb] = bt[(2 * kb + 1) * 64 + lane];
            }
            wave_mfma_r<NKB>(a, bh, bl, acc, acl);
        }
    };

    // input row yh0 + 32 t + n of H(t), staged columns xs + 8 g + 16 kb + [0, 8),
    // through raw buffers with 32-bit offsets from the image base (W * H < 2^31)
    const auto srs = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint8_t *>(src), 0,
                                                       W * H, 0x00020000);
    const auto drs = __builtin_amdgcn_make_buffer_rsrc(dst, 0, W * H, 0x00020000);
    // dword stores where dst or its rows are not 4-aligned: a descriptor from
    // the 4-aligned address at or below dst (stores stay inside the image)
    const uint32_t ibd = (uint32_t)reinterpret_cast<uintptr_t>(dst) & 3u;
    const auto drs4 = __builtin_amdgcn_make_buffer_rsrc(dst - ibd, 0,
                                                        (int)(((uint32_t)(W * H) + ibd + 3u) & ~3u),
                                                        0x00020000);
    // H(t)'s window, input rows yh0 + 32 t + [0, 32) by staged columns
    // xs + [0, CIN), is loaded as 16-byte row chunks (NLD per lane): chunk c
    // is row c / NKB, columns 16 (c % NKB) + [0, 16). Rows then take few
    // cache lines per load; the A operands are cut from the bytes in LDS.
    constexpr int NCH = 32 * NKB, NLD = (NCH + 63) / 64;
    // a chunk as loaded: four dwords (aligned rows), or five dwords from the
    // 4-aligned address at or below its first byte and the byte shift that
    // realigns them (rows that are not 4-aligned; shift 0 where the chunk
    // was assembled byte by byte)
    struct Chunks {
        uint32_t w[NLD][5];
        uint32_t sh[NLD];
    };
    // unaligned rows: dword loads through a descriptor whose base is the
    // 4-aligned address at or below src, over every dword holding an image byte
    const uint32_t ib = (uint32_t)reinterpret_cast<uintptr_t>(src) & 3u;
    const auto srs4 = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint8_t *>(src - ib), 0,
                                                        (int)(((uint32_t)(W * H) + ib + 3u) & ~3u),
                                                        0x00020000);
    auto load_h = [&](int t, Chunks &d) {
#pragma unroll
        for (int k = 0; k < NLD; k++) {
            const int c = min(lane + 64 * k, NCH - 1);
            const int r = c / NKB, q = c - r * NKB;
            const int row = clampi(yh0 + 32 * t + r, 0, H - 1) * W;
            const int x = xs + 16 * q;
            if (interior) {
                // two 8-byte loads: a 16-byte load at an address 12 past a
                // multiple of 16 (as xs is for 9 taps) returned wrong bytes
                // at times on gfx950 here, 8-byte and 4-byte loads never did
                const auto v0 = __builtin_amdgcn_raw_buffer_load_b64(srs, row + x, 0, 0);
                const auto v1 = __builtin_amdgcn_raw_buffer_load_b64(srs, row + x + 8, 0, 0);
                d.w[k][0] = v0[0];
                d.w[k][1] = v0[1];
                d.w[k][2] = v1[0];
                d.w[k][3] = v1[1];
                continue;
            }
            if (ALIGNED) {
#pragma unroll
                for (int p = 0; p < 4; p++)
                    d.w[k][p] = __builtin_amdgcn_raw_buffer_load_b32(
                        srs, row + clampi(x + 4 * p, 0, W - 4), 0, 0);
                continue;
            }
            // (the branches leave values, not stores: stores to the chunk
            // array in both arms would be merged into one at a computed
            // element, which moves the array to scratch memory)
            uint32_t wv[5], shv;
            if (x >= 0 && x + 16 <= W) {
                // inside the row: five aligned dwords, realigned when staged
                const uint32_t o = (uint32_t)(row + x) + ib;
                shv = o & 3u;
#pragma unroll
                for (int p = 0; p < 5; p++)
                    wv[p] = __builtin_amdgcn_raw_buffer_load_b32(srs4, (int)((o & ~3u) + 4 * p), 0, 0);
            } else {
                // across an image edge: byte by byte from the clamped columns
                shv = 0;
                wv[4] = 0;
#pragma unroll
                for (int p = 0; p < 4; p++) {
                    uint32_t v = 0;
#pragma unroll
                    for (int b = 0; b < 4; b++)
                        v |= (uint32_t)__builtin_amdgcn_raw_buffer_load_b8(
                                 srs, row + clampi(x + 4 * p + b, 0, W - 1), 0, 0)
                             << (8 * b);
                    wv[p] = v;
                }
            }
#pragma unroll
            for (int p = 0; p < 5; p++) d.w[k][p] = wv[p];
            d.sh[k] = shv;
        }
    };
    // the loaded chunks into in_w (at the image edges, aligned dwords left of
    // the row become its first byte, right of it its last), then this lane's
    // A operand, row n, columns 8 g + 16 kb + [0, 8), as exact f16
    auto stage_h = [&](Chunks &d, v8h (&a)[NKB]) {
#pragma unroll
        for (int k = 0; k < NLD; k++) {
            const int c = lane + 64 * k;
            if (NCH % 64 != 0 && c >= NCH) break;
            const int r = c / NKB, q = c - r * NKB;
            uint32_t dw[4];
#pragma unroll
            for (int p = 0; p < 4; p++)
                dw[p] = ALIGNED ? d.w[k][p]
                                : __builtin_amdgcn_alignbyte(d.w[k][p + 1], d.w[k][p], d.sh[k]);
            if (ALIGNED && !interior) {
#pragma unroll
                for (int p = 0; p < 4; p++) {
                    const int x4 = xs + 16 * q + 4 * p;
                    const uint32_t left = (dw[p] & 0xffu) * 0x01010101u;
                    const uint32_t right = (dw[p] >> 24) * 0x01010101u;
                    dw[p] = x4 < 0 ? left : dw[p];
                    dw[p] = x4 >= W ? right : dw[p];
                }
            }
            uint8_t *o = in_w + r * PI + 16 * q;
            *reinterpret_cast<uint2 *>(o) = make_uint2(dw[0], dw[1]);
            *reinterpret_cast<uint2 *>(o + 8) = make_uint2(dw[2], dw[3]);
        }
        wave_lds_fence();  // the rows were written by other lanes
#pragma unroll
        for (int kb = 0; kb < NKB; kb++)
            a[kb] = bytes_to_half8(*reinterpret_cast<const uint2 *>(in_w + n * PI + 8 * g + 16 * kb));
        wave_a_init<NKB>(g, a);
    };
    // H readout into ring slot SLOT: the lane holds intermediate column n,
    // rows 8 (r / 4) + 4 g + r % 4 of the step. In a flagged pass only the
    // groups of four whose smallest fraction was below the margin are
    // checked again, from the accumulators (still live).
    // Intermediate columns right of the image and output rows below it are
    // never stored, so their flagged outputs are dropped before the exact
    // chain: such columns repeat the edge byte and such rows the last row,
    // and the blur of a constant run lands on an integer, so nearly all of
    // them are flagged. (Interior strips hold no column right of the image;
    // the V pass of such a column blurs the edge column's values, as near
    // integers as any other.)
    const bool hclip = !interior && x0 + n >= W;
    auto h_readout = [&](auto slot_c, const v16f &acc, const v16f &acl) {
        constexpr int SLOT = decltype(slot_c)::value;
        _Float16 *col = ring + n * RPH + 32 * SLOT;
        uint32_t mg[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            uint32_t q[4];
#pragma unroll
            for (int i = 0; i < 4; i++) q[i] = __float_as_uint(acc[4 * j + i] + acl[4 * j + i]);
            *reinterpret_cast<uint2 *>(col + 8 * j + 4 * g) =
                make_uint2(half2_bits15(q[0], q[1]), half2_bits15(q[2], q[3]));
            mg[j] = min_frac4(q);
        }
        if (__ballot(min(min(mg[0], mg[1]), min(mg[2], mg[3])) < thr_h)) {
            wave_lds_fence();  // in_w rows were written by other lanes
            uint32_t fl = 0;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                if (mg[j] < thr_h) {
#pragma unroll
                    for (int i = 0; i < 4; i++)
                        fl |= ((__float_as_uint(acc[4 * j + i] + acl[4 * j + i]) & kFxFracMask) < thr_h)
                                  ? (1u << (4 * j + i)) : 0u;
                }
            }
            if (hclip) fl = 0;
            while (fl) {
                const int r = __builtin_ctz(fl);
                fl &= fl - 1u;
                const int i = 8 * (r >> 2) + 4 * g + (r & 3);
                const uint8_t *p = in_w + i * PI + n + m;
                const float e = wave_chain(taps_h, L, [&](int t) { return (float)p[t]; });
                col[i] = (_Float16)(float)clamp_u8_exact(e);
            }
        }
    };
    // V pass of step s (parity P): the lane holds output row n, columns
    // 8 (r / 4) + 4 g + r % 4; the rows go to out_w, then to the image
    const bool vec_out = ((W & 15) == 0) && ((reinterpret_cast<uintptr_t>(dst) & 15) == 0);
    const bool dw_out = ALIGNED && ((reinterpret_cast<uintptr_t>(dst) & 3) == 0);
    auto v_mfma = [&](auto par_c, v16f &acc, v16f &acl) {
        constexpr int P = decltype(par_c)::value;
        wave_lds_fence();  // ring columns were written by other lanes
        const _Float16 *acol = ring + n * RPH + 8 * g;
        v8h a[NKB];
#pragma unroll
        for (int kb = 0; kb < NKB; kb++)
            a[kb] = *reinterpret_cast<const v8h *>(acol + ((32 * P + 16 * kb) & (kRing - 1)));
        wave_a_init<NKB>(g, a);
        pass_mfma(a, bt_v, acc, acl);
    };
    auto v_readout = [&](auto par_c, int s, const v16f &acc, const v16f &acl) {
        constexpr int P = decltype(par_c)::value;
        uint8_t *orow = out_w + n * PO;
        uint32_t mg[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            uint32_t q[4];
#pragma unroll
            for (int i = 0; i < 4; i++) q[i] = __float_as_uint(acc[4 * j + i] + acl[4 * j + i]);
            *reinterpret_cast<uint32_t *>(orow + 8 * j + 4 * g) = pack_bits15(q[0], q[1], q[2], q[3]);
            mg[j] = min_frac4(q);
        }
        if (__ballot(min(min(mg[0], mg[1]), min(mg[2], mg[3])) < thr_v)) {
            uint32_t fl = 0;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                if (mg[j] < thr_v) {
#pragma unroll
                    for (int i = 0; i < 4; i++)
                        fl |= ((__float_as_uint(acc[4 * j + i] + acl[4 * j + i]) & kFxFracMask) < thr_v)
                                  ? (1u << (4 * j + i)) : 0u;
                }
            }
            if (yb0 + 32 * s + n >= yb1) fl = 0;
            while (fl) {
                const int r = __builtin_ctz(fl);
                fl &= fl - 1u;
                const int i = 8 * (r >> 2) + 4 * g + (r & 3);
                const _Float16 *c = ring + i * RPH;
                // ring rows r0 .. r0 + L - 1 (mod 64); read straight when they
                // do not wrap (the rows past L up to the chunk end then lie in
                // the column's zeroed padding rows 64 .. 67)
                const int r0 = (32 * P + n + m) & (kRing - 1);
                const float e =
                    r0 + L <= kRing
                        ? wave_chain(taps_v, L, [&](int t) { return (float)c[r0 + t]; })
                        : wave_chain(taps_v, L, [&](int t) {
                              return (float)c[(r0 + t) & (kRing - 1)];
                          });
                orow[i] = clamp_u8_exact(e);
            }
        }
        wave_lds_fence();  // out_w rows were written by other lanes
        // rows y + lane / 2, columns x0 + 16 (lane % 2) + [0, 16)
        const int yl = lane >> 1, xl = (lane & 1) * 16;
        const int gy = yb0 + 32 * s + yl, gx = x0 + xl;
        if (gy < yb1 && gx < W) {
            const uint4 v = *reinterpret_cast<const uint4 *>(out_w + yl * PO + xl);
            const int o = gy * W + gx;
            if constexpr (VEC) {
                typedef unsigned int v4u __attribute__((ext_vector_type(4)));
                const v4u vv = {v.x, v.y, v.z, v.w};
                __builtin_amdgcn_raw_buffer_store_b128(vv, drs, o, 0, 0);
            } else if (vec_out && gx + 16 <= W) {
                typedef unsigned int v4u __attribute__((ext_vector_type(4)));
                const v4u vv = {v.x, v.y, v.z, v.w};
                __builtin_amdgcn_raw_buffer_store_b128(vv, drs, o, 0, 0);
            } else if (dw_out) {
                // rows 4-aligned but not all 16-aligned: each piece as one
                // 16-byte or two 8-byte stores where its address allows
                // (fewer, wider writes), else four dword stores
                const uint32_t dws[4] = {v.x, v.y, v.z, v.w};
                const uint32_t al = (uint32_t)reinterpret_cast<uintptr_t>(dst + o) & 15u;
                typedef unsigned int v2u __attribute__((ext_vector_type(2)));
                typedef unsigned int v4u __attribute__((ext_vector_type(4)));
                if (gx + 16 <= W && al == 0) {
                    const v4u vv = {v.x, v.y, v.z, v.w};
                    __builtin_amdgcn_raw_buffer_store_b128(vv, drs, o, 0, 0);
                } else if (gx + 16 <= W && (al & 7u) == 0) {
                    const v2u lo = {v.x, v.y}, hi = {v.z, v.w};
                    __builtin_amdgcn_raw_buffer_store_b64(lo, drs, o, 0, 0);
                    __builtin_amdgcn_raw_buffer_store_b64(hi, drs, o + 8, 0, 0);
                } else {
#pragma unroll
                    for (int q = 0; q < 4; q++)
                        if (gx + 4 * q < W)
                            __builtin_amdgcn_raw_buffer_store_b32(dws[q], drs, o + 4 * q, 0, 0);
                }
            } else if constexpr (ALIGNED) {
                // 4-aligned rows from a dst that is not (the register-lean
                // byte loop keeps the aligned kernels free of spills)
                const uint8_t *b = out_w + yl * PO + xl;
                for (int q = 0; q < 16 && gx + q < W; q++) dst[o + q] = b[q];
            } else {
                // rows not 4-aligned: the bytes before the first 4-aligned
                // address, whole dwords, then the bytes left (the dwords at
                // a piece's ends are shared with the neighbouring piece, so
                // those take byte stores)
                const uint8_t *b = out_w + yl * PO + xl;
                const int nb = min(16, W - gx);
                const uint32_t a = (uint32_t)o + ibd;  // from the 4-aligned base dst - ibd
                const int sh = (int)((4u - (a & 3u)) & 3u);
                const int hb = min(sh, nb), nd = (nb - hb) >> 2;
                const uint32_t w5[5] = {v.x, v.y, v.z, v.w, 0u};
                for (int q = 0; q < hb; q++) dst[o + q] = b[q];
#pragma unroll
                for (int k = 0; k < 4; k++)
                    if (k < nd)
                        __builtin_amdgcn_raw_buffer_store_b32(
                            __builtin_amdgcn_alignbyte(w5[k + 1], w5[k], (uint32_t)sh), drs4,
                            (int)(a + (uint32_t)sh + 4u * (uint32_t)k), 0, 0);
                for (int q = hb + 4 * nd; q < nb; q++) dst[o + q] = b[q];
            }
        }
    };

    // Software pipeline over two accumulator sets: X for H passes, Y for V
    // passes. Iteration s converts H(s + 1)'s bytes and issues its MFMAs,
    // reads V(s) out while they run, reads H(s + 1) out and issues V(s + 1)'s
    // MFMAs, which run under the next iteration's conversion. H(t) goes to
    // ring slot (t + 1) & 1; V(s) starts at ring row 32 (s & 1); H(t)'s bytes
    // are loaded two passes ahead, into dA for odd t and dB for even t.
    using P0 = std::integral_constant<int, 0>;
    using P1 = std::integral_constant<int, 1>;
    Chunks dA, dB;
    v16f xc, xl, yc, yl;
    auto h_mfma = [&](Chunks &d) {
        v8h a[NKB];
        stage_h(d, a);
        pass_mfma(a, bt_h, xc, xl);
    };
    // (the loads run two passes ahead unconditionally: rows past the image
    // are clamped, so the last ones read rows that are cached already, and
    // the wait for each load stays a count of the loads issued after it)
    load_h(-1, dA);
    load_h(0, dB);
    h_mfma(dA);
    load_h(1, dA);
    h_readout(P0{}, xc, xl);  // H(-1)
    h_mfma(dB);
    load_h(2, dB);
    h_readout(P1{}, xc, xl);  // H(0)
    v_mfma(P0{}, yc, yl);     // V(0)
    auto iter = [&](auto par_c, int s, Chunks &dx) {
        constexpr int P = decltype(par_c)::value;
        using Q = std::integral_constant<int, 1 - P>;
        const bool more = s + 1 < nsteps;
        if (more) {
            h_mfma(dx);  // H(s + 1)
            load_h(s + 3, dx);
        }
        v_readout(par_c, s, yc, yl);  // V(s)
        if (more) {
            h_readout(par_c, xc, xl);  // H(s + 1) -> slot (s + 2) & 1
            v_mfma(Q{}, yc, yl);       // V(s + 1)
        }
    };
    for (int s = 0; s < nsteps; s += 2) {
        iter(P0{}, s, dA);
        if (s + 1 < nsteps) iter(P1{}, s + 1, dB);
    }
}

// The workgroup's shared tables for job `jb` (uniform across the workgroup),
// then wave unit u of the job, unless `valid` is false (the wave only takes
// part in the table set-up).
template <int NKB, bool SHARE_B, bool ALIGNED>
__device__ __forceinline__ void blur_u8_wave_unit(const WaveJob &jb, int u, bool valid) {
    constexpr int CIN = 16 * NKB, PI = wave_in_pitch<NKB>(), WL = wave_lds_bytes<NKB>();
    extern __shared__ __align__(16) unsigned char smem[];
    float *taps_h = reinterpret_cast<float *>(smem);
    float *taps_v = taps_h + (2 * kTapPad + kMaxTaps);
    v8h *bt_h = reinterpret_cast<v8h *>(smem + 2 * (2 * kTapPad + kMaxTaps) * sizeof(float));
    v8h *bt_v = SHARE_B ? bt_h : bt_h + 2 * NKB * 64;
    unsigned char *wbase = smem + wave_shared_lds<NKB, SHARE_B>() +
                           __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) * WL;
    uint8_t *in_w = wbase;                                           // 32 rows x PI bytes
    _Float16 *ring = reinterpret_cast<_Float16 *>(wbase + 32 * PI);  // 32 columns x kRingPitch
    uint8_t *out_w = wbase + 32 * PI + 32 * kRingPitch;              // 32 rows x PO bytes

    const int L = jb.L;
    if (threadIdx.x < 2 * kTapPad + kMaxTaps) {
        const int t = (int)threadIdx.x - kTapPad;
        const bool in = t >= 0 && t < L;
        taps_h[threadIdx.x] = in ? jb.h[t] : 0.0f;
        taps_v[threadIdx.x] = in ? jb.v[t] : 0.0f;
    }
    __syncthreads();  // taps staged
    if (threadIdx.x < 64) wave_b_table<NKB>(taps_h, (int)threadIdx.x, jb.m, jb.ch, bt_h);
    if (!SHARE_B && threadIdx.x >= 64 && threadIdx.x < 128)
        wave_b_table<NKB>(taps_v, (int)threadIdx.x - 64, jb.m, jb.cv, bt_v);
    __syncthreads();  // Toeplitz tables written; no workgroup barrier after this one
    if (!valid) return;

    // wave unit = (image, band, strip), strip fastest
    const int per_img = jb.bands * jb.strips;
    const int im = u / per_img, rem = u - im * per_img;
    const int band = rem / jb.strips, strip = rem - band * jb.strips;
    const int xs = strip * 32 - L / 2 - jb.m;
    if (!ALIGNED)
        blur_u8_wave_strip<NKB, SHARE_B, kWaveUnaligned>(jb, taps_h, taps_v, bt_h, bt_v, in_w,
                                                         ring, out_w, im, band, strip);
    else if (xs >= 0 && xs + CIN <= jb.W) {
        const bool vec = (jb.W & 15) == 0 && (jb.stride & 15) == 0 &&
                         (reinterpret_cast<uintptr_t>(jb.dst) & 15) == 0;
        if (vec)
            blur_u8_wave_strip<NKB, SHARE_B, kWaveInteriorVec>(jb, taps_h, taps_v, bt_h, bt_v,
                                                               in_w, ring, out_w, im, band, strip);
        else
            blur_u8_wave_strip<NKB, SHARE_B, kWaveInterior>(jb, taps_h, taps_v, bt_h, bt_v, in_w,
                                                            ring, out_w, im, band, strip);
    }
    else
        blur_u8_wave_strip<NKB, SHARE_B, kWaveEdge>(jb, taps_h, taps_v, bt_h, bt_v, in_w, ring,
                                                    out_w, im, band, strip);
}

// A workgroup takes four neighbouring wave units, and contiguous eighths of
// the workgroups go to the eight XCDs (workgroups are dealt round-robin).
// The wave index is read as a wave-uniform value, so the unit, the strip's
// geometry and the buffer descriptors live in scalar registers.
template <int NKB, bool SHARE_B, bool ALIGNED>
__device__ __forceinline__ void blur_u8_wave_body(const WaveJob &jb) {
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int units = jb.n_img * jb.bands * jb.strips;
    const int chunk = (((units + 3) >> 2) + 7) >> 3;
    if ((int)(blockIdx.x >> 3) >= chunk) return;  // the whole workgroup
    const int u = ((int)(blockIdx.x & 7) * chunk + (int)(blockIdx.x >> 3)) * 4 + wave;
    blur_u8_wave_unit<NKB, SHARE_B, ALIGNED>(jb, u, u < units);
}

// Waves per SIMD: 3 (168 VGPRs) for 4-aligned rows; 2 for the others, whose
// five-dword chunks and funnel shifts would spill at 3.
constexpr int wave_kernel_waves(bool aligned) { return aligned ? 3 : 2; }

// The job is read through a reference to the kernel argument: its taps are
// indexed per thread, which must not copy the argument into registers.
template <int NKB, bool SHARE_B, bool ALIGNED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(wave_kernel_waves(ALIGNED), wave_kernel_waves(ALIGNED)))) void
blur_u8_wave_kernel(WaveJob job) {
    blur_u8_wave_body<NKB, SHARE_B, ALIGNED>(job);
}

// The same over a device table of single-image jobs (ragged batches): job j
// owns wave units [unit0, unit0 + bands * strips), unit0 a multiple of 4, so
// the four waves of a workgroup always belong to one job and share its
// tables. nwg = total units / 4.
template <int NKB, bool SHARE_B, bool ALIGNED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(wave_kernel_waves(ALIGNED), wave_kernel_waves(ALIGNED)))) void
blur_u8_wave_jobs_kernel(const WaveJob *__restrict__ jobs, int njobs, int nwg) {
    const int chunk = (nwg + 7) >> 3;
    const int wg = (int)(blockIdx.x & 7) * chunk + (int)(blockIdx.x >> 3);
    if ((int)(blockIdx.x >> 3) >= chunk || wg >= nwg) return;  // the whole workgroup
    int lo = 0, hi = njobs - 1;  // the last job with unit0 <= 4 wg
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].unit0 <= 4 * wg) lo = mid;
        else hi = mid - 1;
    }
    const WaveJob &jb = jobs[lo];
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int u = 4 * wg + wave - jb.unit0;
    blur_u8_wave_unit<NKB, SHARE_B, ALIGNED>(jb, u, u < jb.bands * jb.strips);
}

template <int NKB, bool SHARE_B> constexpr size_t wave_kernel_lds() {
    return (size_t)wave_shared_lds<NKB, SHARE_B>() + 4 * (size_t)wave_lds_bytes<NKB>();
}

iph_status launch_blur_wave(const uint8_t *src, uint8_t *dst, uint32_t w, uint32_t h, uint32_t n,
                            size_t stride, const float *hk, const float *vk, int L, int m,
                            int nkb, const FxPlan &ph, const FxPlan &pv, bool same, bool aligned,
                            hipStream_t s) {
    WaveJob jb{};
    jb.src = src;
    jb.dst = dst;
    jb.stride = (long long)stride;
    jb.W = (int)w;
    jb.H = (int)h;
    jb.n_img = (int)n;
    jb.strips = (int)((w + 31) / 32);
    // bands of rows: enough waves to fill the chip a few times over
    const long long per_band = (long long)n * jb.strips;
    const int steps_total = (int)((h + 31) / 32);
    int bands = 1;
    if (per_band < 8192)
        bands = (int)std::max<long long>(
            1, std::min<long long>(steps_total, (8192 + per_band - 1) / per_band));
    const int steps_per_band = (steps_total + bands - 1) / bands;
    jb.band_rows = steps_per_band * 32;
    jb.bands = (int)((h + jb.band_rows - 1) / jb.band_rows);
    const long long units = per_band * jb.bands;
    if (units == 0) return IPH_OK;
    if (units > (1LL << 30)) {
        set_error("batch of %u images of %ux%u is too large for one launch", n, w, h);
        return IPH_EINVAL;
    }
    jb.L = L;
    jb.m = m;
    jb.ch = ph.c;
    jb.cv = pv.c;
    for (int i = 0; i < L; i++) {
        jb.h[i] = hk[i];
        jb.v[i] = vk[i];
    }
    const long long wgs = (units + 3) / 4;
    const unsigned grid = (unsigned)((wgs + 7) / 8 * 8);
#define WB(K, SAME, AL)                                                                      \
    blur_u8_wave_kernel<K, SAME, AL><<<grid, 256, wave_kernel_lds<K, SAME>(), s>>>(jb)
#define WA(K, SAME)                                                                          \
    if (aligned) WB(K, SAME, true); else WB(K, SAME, false)
#define WK(K)                                                                                \
    if (same) { WA(K, true); } else { WA(K, false); }                                        \
    break
    switch (nkb) {
    case 2: WK(2);
    case 3: WK(3);
    default: WK(4);
    }
#undef WK
#undef WA
#undef WB
    return launch_status("blur_u8_wave_kernel");
}

}  // namespace

// Taps the MFMA path handles: finite, at most 33 (sigma <= 8 for a Gaussian;
// the LDS footprint stays under 64 KiB), not all zero.
bool blur_mfma_supported(const float *hk, const float *vk, uint32_t len) {
    if (len == 0 || len > 33u) return false;
    float m = 0.0f;
    for (uint32_t i = 0; i < len; i++) {
        if (!std::isfinite(hk[i]) || !std::isfinite(vk[i])) return false;
        m = std::fmax(m, std::fmax(std::fabs(hk[i]), std::fabs(vk[i])));
    }
    if (!(m > 0.0f && m < 1.0e30f)) return false;
    // nonzero taps stay far from the subnormal range, so products scaled by
    // a power of two round exactly like the unscaled ones
    for (uint32_t i = 0; i < len; i++) {
        float a = std::fabs(hk[i]), b = std::fabs(vk[i]);
        if ((a != 0.0f && a < 0x1p-90f) || (b != 0.0f && b < 0x1p-90f)) return false;
    }
    return true;
}

iph_status launch_blur_u8_mfma(const uint8_t *src, uint8_t *dst, uint32_t w, uint32_t h,
                               uint32_t n, size_t stride, const float *hk, const float *vk,
                               uint32_t len, hipStream_t s) {
    BlurArgs a{};
    const int L = (int)len;
    const int half = L / 2;
    const int m = (4 - half % 4) % 4;            // x0 is a multiple of 128
    const int nkb = (31 + L + m + 15) / 16;
    float mx = 0.0f;
    for (int i = 0; i < L; i++) {
        a.h[i] = hk[i];
        a.v[i] = vk[i];
        mx = std::fmax(mx, std::fmax(std::fabs(hk[i]), std::fabs(vk[i])));
    }
    // 2^S puts the largest tap in [512, 1024): hi and lo parts stay normal f16.
    int ex = 0;
    std::frexp(mx, &ex);  // mx = f * 2^ex, f in [0.5, 1)
    int S = 10 - ex;
    a.scale = std::ldexp(1.0f, S);
    a.inv_scale = std::ldexp(1.0f, -S);
    a.eps_h = blur_eps(hk, L, nkb);
    a.eps_v = blur_eps(vk, L, nkb);
    a.L = L;
    a.m = m;
    const int tiles_x = (int)((w + kTW - 1) / kTW), tiles_y = (int)((h + kTH - 1) / kTH);
    const long long ntiles = (long long)tiles_x * tiles_y * n;
    if (ntiles > (1LL << 30)) {
        set_error("batch of %u images of %ux%u is too large for one launch", n, w, h);
        return IPH_EINVAL;
    }
    bool same = true;
    for (int i = 0; i < L; i++) same = same && (hk[i] == vk[i]);
    const bool aligned = w >= 4 && (w % 4) == 0 && (stride % 4) == 0 &&
                         (reinterpret_cast<uintptr_t>(src) & 3) == 0;
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
#define LB(K, SAME, AL)                                                                      \
    do {                                                                                     \
        using Sh = BlurShape<K>;                                                             \
        size_t lds = (size_t)kRows * Sh::PIN + (size_t)kTW * Sh::PITCH_T * sizeof(_Float16) +  \
                     (size_t)kTH * kTW + 2 * (2 * kTapPad + kMaxTaps) * sizeof(float) +      \
                     4 * kQCap * sizeof(uint32_t);                                            \
        int per_cu = 0;                                                                      \
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(                                    \
                &per_cu, blur_u8_mfma_kernel<K, SAME, AL>, 256, lds) != hipSuccess ||        \
            per_cu < 1)                                                                      \
            per_cu = 1;                                                                      \
        long long grid = (long long)cus * per_cu;                                            \
        grid = std::min<long long>(grid, (ntiles + 7) / 8 * 8);                              \
        grid = std::max<long long>(8, grid / 8 * 8);                                         \
        blur_u8_mfma_kernel<K, SAME, AL><<<(unsigned)grid, 256, lds, s>>>(                  \
            src, dst, (int)w, (int)h, stride, (int)n, tiles_x, tiles_y, a);                  \
    } while (0)
#define LA(K, SAME)                                                                          \
    if (aligned) LB(K, SAME, true); else LB(K, SAME, false)
#define LK(K)                                                                                \
    if (same) { LA(K, true); } else { LA(K, false); }                                        \
    break
    // the per-wave kernel: two K rows beyond the taps carry the start value
    const int nkb_w = (L + m + 33 + 15) / 16;
    if (L <= kFxMaxL && nkb_w <= 4 && (uint64_t)w * h < (1ull << 31)) {
        const FxPlan sh = blur_fx_plan(hk, L, nkb_w), sv = blur_fx_plan(vk, L, nkb_w);
        if (sh.c > 0 && sv.c > 0)
            return launch_blur_wave(src, dst, w, h, n, stride, hk, vk, L, m, nkb_w, sh, sv, same,
                                    aligned, s);
    }
    if (L <= kFxMaxL && nkb >= 2 && nkb <= 4) {
        const FxPlan sh = blur_fx_plan(hk, L, nkb), sv = blur_fx_plan(vk, L, nkb);
        if (sh.c > 0 && sv.c > 0)
            return launch_blur_stream(src, dst, w, h, n, stride, hk, vk, L, m, nkb, sh, sv, same,
                                      aligned, s);
    }
    const FxPlan ph = blur_fx_plan(hk, L, nkb), pv = blur_fx_plan(vk, L, nkb);
    if (ph.c > 0 && pv.c > 0) {
        FxJob f{};
        f.src = src;
        f.dst = dst;
        f.stride = (long long)stride;
        f.W = (int)w;
        f.H = (int)h;
        f.n_img = (int)n;
        f.tiles_x = (int)((w + kTW - 1) / kTW);
        f.tiles_y = (int)((h + kTH2 - 1) / kTH2);
        f.tile0 = 0;
        for (int i = 0; i < L; i++) {
            f.h[i] = hk[i];
            f.v[i] = vk[i];
        }
        f.L = L;
        f.m = m;
        f.ch = ph.c;
        f.cv = pv.c;
        f.init_h = ph.init;
        f.init_v = pv.init;
        const long long fnt = (long long)f.tiles_x * f.tiles_y * n;
#define FB(K, SAME, AL)                                                                      \
    do {                                                                                     \
        using Sh = RowStage<K, kRows2>;                                                     \
        size_t lds = (size_t)((kRows2 * Sh::PIN + 15) / 16 * 16) + (size_t)kTW * kPitchB +   \
                     2 * (2 * kTapPad + kMaxTaps) * sizeof(float) +                          \
                     4 * kFxQCap * sizeof(uint16_t);                                          \
        int per_cu = 0;                                                                      \
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(                                    \
                &per_cu, blur_u8_fx_kernel<K, SAME, AL>, 256, lds) != hipSuccess ||          \
            per_cu < 1)                                                                      \
            per_cu = 1;                                                                      \
        long long grid = (long long)cus * per_cu;                                            \
        grid = std::min<long long>(grid, (fnt + 7) / 8 * 8);                                 \
        grid = std::max<long long>(8, grid / 8 * 8);                                         \
        blur_u8_fx_kernel<K, SAME, AL><<<(unsigned)grid, 256, lds, s>>>(f);                 \
    } while (0)
#define FA(K, SAME)                                                                          \
    if (aligned) FB(K, SAME, true); else FB(K, SAME, false)
#define FK(K)                                                                                \
    if (same && L + m <= 33) { FA(K, 2); } else if (same) { FA(K, 1); } else { FA(K, 0); }  \
    break
        switch (nkb) {
        case 2: FK(2);
        case 3: FK(3);
        default: FK(4);
        }
#undef FK
#undef FA
#undef FB
        return launch_status("blur_u8_fx_kernel");
    }
    switch (nkb) {
    case 2: LK(2);
    case 3: LK(3);
    default: LK(4);
    }
#undef LK
#undef LA
#undef LB
    return launch_status("blur_u8_mfma_kernel");
}

// ---- batches: jobs grouped by kernel variant -------------------------------------
//
// The device table holds the fixed-point tile kernel's FxJob entries (n slots)
// followed by the per-wave kernel's WaveJob entries (n slots); each group's
// entries are consecutive in their part.

size_t blur_jobs_table_bytes(int n) {
    return (size_t)std::max(n, 1) * (sizeof(FxJob) + sizeof(WaveJob));
}

iph_status blur_jobs_table(const BlurSpec *specs, int n, void *host_table,
                           std::vector<BlurGroup> &groups, std::vector<int> &fallback) {
    struct Entry {
        int key, idx;
        FxJob job;
    };
    struct WEntry {
        int key, idx;
        WaveJob job;
    };
    std::vector<Entry> entries;
    std::vector<WEntry> wentries;
    groups.clear();
    fallback.clear();
    for (int i = 0; i < n; i++) {
        const BlurSpec &b = specs[i];
        const int L = (int)b.len;
        bool same = true;
        if (L < 1 || L > kFxMaxL || !blur_mfma_supported(b.taps, b.taps, b.len)) {
            fallback.push_back(i);
            continue;
        }
        const int half = L / 2, m = (4 - half % 4) % 4, nkb = (31 + L + m + 15) / 16;
        const bool aligned = b.w >= 4 && (b.w % 4) == 0 &&
                             (reinterpret_cast<uintptr_t>(b.src) & 3) == 0;
        // the per-wave kernel where it takes the taps (two K rows beyond them
        // carry the start value)
        const int nkb_w = (L + m + 33 + 15) / 16;
        if (nkb_w <= 4 && (uint64_t)b.w * b.h < (1ull << 31)) {
            const FxPlan pw = blur_fx_plan(b.taps, L, nkb_w);
            if (pw.c > 0) {
                WaveJob j{};
                j.src = b.src;
                j.dst = b.dst;
                j.stride = (long long)b.w * b.h;
                j.W = (int)b.w;
                j.H = (int)b.h;
                j.n_img = 1;
                j.strips = (int)((b.w + 31) / 32);
                // bands of about 24 steps of 32 rows
                const int steps = (int)((b.h + 31) / 32);
                const int bands0 = std::max(1, steps / 24);
                j.band_rows = (steps + bands0 - 1) / bands0 * 32;
                j.bands = (int)((b.h + j.band_rows - 1) / j.band_rows);
                j.L = L;
                j.m = m;
                j.ch = j.cv = pw.c;
                for (int q = 0; q < L; q++) j.h[q] = j.v[q] = b.taps[q];
                wentries.push_back({nkb_w * 2 + (aligned ? 1 : 0), i, j});
                continue;
            }
        }
        const FxPlan pl = blur_fx_plan(b.taps, L, nkb);
        if (pl.c <= 0 || nkb < 2 || nkb > 4) {
            fallback.push_back(i);
            continue;
        }
        const int cls = (same && L + m <= 33) ? 2 : 1;
        FxJob f{};
        f.src = b.src;
        f.dst = b.dst;
        f.stride = (long long)b.w * b.h;
        f.W = (int)b.w;
        f.H = (int)b.h;
        f.n_img = 1;
        f.tiles_x = (int)((b.w + kTW - 1) / kTW);
        f.tiles_y = (int)((b.h + kTH2 - 1) / kTH2);
        for (int q = 0; q < L; q++) f.h[q] = f.v[q] = b.taps[q];
        f.L = L;
        f.m = m;
        f.ch = f.cv = pl.c;
        f.init_h = f.init_v = pl.init;
        entries.push_back({nkb * 8 + cls * 2 + (aligned ? 1 : 0), i, f});
    }
    std::stable_sort(entries.begin(), entries.end(),
                     [](const Entry &a, const Entry &b) { return a.key < b.key; });
    std::stable_sort(wentries.begin(), wentries.end(),
                     [](const WEntry &a, const WEntry &b) { return a.key < b.key; });
    FxJob *table = static_cast<FxJob *>(host_table);
    for (size_t e = 0; e < entries.size(); e++) {
        const int key = entries[e].key;
        if (groups.empty() || groups.back().nkb * 8 + groups.back().taps_class * 2 +
                                      (groups.back().aligned ? 1 : 0) != key) {
            groups.push_back({key / 8, (key % 8) / 2, (key & 1) != 0, false, (int)e, 0, 0});
        }
        BlurGroup &g = groups.back();
        FxJob f = entries[e].job;
        f.tile0 = g.ntiles;
        const long long nt = (long long)g.ntiles + (long long)f.tiles_x * f.tiles_y;
        if (nt > (1LL << 30)) {
            set_error("batch too large for one launch");
            return IPH_EINVAL;
        }
        g.ntiles = (int)nt;
        g.count++;
        table[e] = f;
    }
    WaveJob *wtable = reinterpret_cast<WaveJob *>(
        static_cast<unsigned char *>(host_table) + (size_t)std::max(n, 1) * sizeof(FxJob));
    const size_t nfx = groups.size();
    long long units = 0;
    for (size_t e = 0; e < wentries.size(); e++) {
        const int key = wentries[e].key;
        if (groups.size() == nfx || groups.back().nkb * 2 + (groups.back().aligned ? 1 : 0) != key) {
            groups.push_back({key / 2, 2, (key & 1) != 0, true, (int)e, 0, 0});
            units = 0;
        }
        BlurGroup &g = groups.back();
        WaveJob j = wentries[e].job;
        j.unit0 = (int)units;
        units += ((long long)j.bands * j.strips + 3) / 4 * 4;
        if (units > (1LL << 30)) {
            set_error("batch too large for one launch");
            return IPH_EINVAL;
        }
        g.ntiles = (int)(units / 4);
        g.count++;
        wtable[e] = j;
    }
    return IPH_OK;
}

iph_status launch_blur_jobs(const void *table_dev, int n, const std::vector<BlurGroup> &groups,
                            int grid_pct, hipStream_t s) {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const FxJob *table = static_cast<const FxJob *>(table_dev);
    const WaveJob *wtable = reinterpret_cast<const WaveJob *>(
        static_cast<const unsigned char *>(table_dev) + (size_t)std::max(n, 1) * sizeof(FxJob));
    for (const BlurGroup &g : groups) {
        if (g.ntiles == 0) continue;
        if (g.wave) {
            const unsigned grid = (unsigned)(((long long)g.ntiles + 7) / 8 * 8);
#define WJ(K, AL)                                                                                blur_u8_wave_jobs_kernel<K, true, AL><<<grid, 256, wave_kernel_lds<K, true>(), s>>>(               wtable + g.first, g.count, g.ntiles)
            if (g.nkb == 3) {
                if (g.aligned) WJ(3, true); else WJ(3, false);
            } else {
                if (g.aligned) WJ(4, true); else WJ(4, false);
            }
#undef WJ
            IPH_TRY(launch_status("blur_u8_wave_jobs_kernel"));
            continue;
        }
#define JB(K, CLS, AL)                                                                       \
    do {                                                                                     \
        using Sh = RowStage<K, kRows2>;                                                     \
        size_t lds = (size_t)((kRows2 * Sh::PIN + 15) / 16 * 16) + (size_t)kTW * kPitchB +   \
                     2 * (2 * kTapPad + kMaxTaps) * sizeof(float) +                          \
                     4 * kFxQCap * sizeof(uint16_t);                                          \
        int per_cu = 0;                                                                      \
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(                                    \
                &per_cu, blur_u8_fx_jobs_kernel<K, CLS, AL>, 256, lds) != hipSuccess ||      \
            per_cu < 1)                                                                      \
            per_cu = 1;                                                                      \
        long long grid = (long long)cus * per_cu * std::max(grid_pct, 1) / 100;              \
        grid = std::min<long long>(grid, ((long long)g.ntiles + 7) / 8 * 8);                 \
        grid = std::max<long long>(8, grid / 8 * 8);                                         \
        blur_u8_fx_jobs_kernel<K, CLS, AL><<<(unsigned)grid, 256, lds, s>>>(                \
            table + g.first, g.count, g.ntiles);                                             \
    } while (0)
#define JA(K, CLS)                                                                           \
    if (g.aligned) JB(K, CLS, true); else JB(K, CLS, false)
#define JK(K)                                                                                \
    if (g.taps_class == 2) { JA(K, 2); } else { JA(K, 1); }                                  \
    break
        switch (g.nkb) {
        case 2: JK(2);
        case 3: JK(3);
        default: JK(4);
        }
#undef JK
#undef JA
#undef JB
        IPH_TRY(launch_status("blur_u8_fx_jobs_kernel"));
    }
    return IPH_OK;
}

}  // namespace iph
