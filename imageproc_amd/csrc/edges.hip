// gfx950 kernels for imageproc's canny (edges.rs:30-155).
//
//   1. gaussian_blur_f32(image, 1.4) with the filter module's launcher;
//   2. canny_nms_kernel: per 64 x 16 tile, the Sobel responses of the blurred
//      tile plus a one-pixel ring (filter_clamped::<_, i32, i16>, clamp-to-
//      edge reads), their squared magnitudes in LDS, non_maximum_suppression
//      for the tile's pixels, and the hysteresis class of every pixel: 2 =
//      seed (interior, thinned >= high), 1 = candidate (thinned >= low), 0 =
//      none; plus a flag for each 64 x 64 tile holding a seed;
//   3. canny_hysteresis_kernel, repeated until a launch marks nothing: a
//      candidate becomes marked when one of the pixels whose six-neighbour
//      list (edges.rs:130-137) holds it is marked. Each launch takes a 64 x 64
//      tile as 64 row bitmasks (candidates, marks) and iterates it to a local
//      fixed point, so marks cross one tile border per launch at least; a
//      tile whose 3 x 3 tile neighbourhood marked nothing on its outer ring
//      in the previous launch returns at once. The marked set is the union of
//      what the reference's depth-first searches reach, which does not depend
//      on visiting order. A marked pixel on the image border raises a flag:
//      there the reference's search indexes outside the image and panics;
//   4. canny_finish_kernel: 255 for marked pixels, 0 elsewhere.
//
// Magnitude: (gx as f32).hypot(gy as f32). For the Sobel responses of u8
// images |gx|, |gy| <= 1020, gx^2 + gy^2 <= 2080800 is exact in f32 and glibc's
// hypotf equals the correctly rounded sqrtf of it over that whole domain.
// Direction: the reference bins atan2f(gy, gx) * (180 / pi) (+180 when
// negative) in f32. Over the same domain that equals the exact angle's bin,
// computed here with integer arithmetic (tan 22.5 = sqrt2 - 1, tan 67.5 =
// sqrt2 + 1, and no integer ratio meets either), except at (gx, gy) =
// (-985, -408), where glibc's atan2f (one ulp from correctly rounded) takes
// the f32 angle to 22.5 exactly, bin 45.
#include "iph_internal.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace iph {
namespace {

constexpr int kCannyTW = 64, kCannyTH = 16;
constexpr int kHystT = 64;  // hysteresis tile side
constexpr int kCannyIters = 4;  // hysteresis launches between flag checks
constexpr int kCannyGrid = 1 << 20;  // workgroup cap of the tile kernels

// Branch-free (selects only), so the lanes of a wave never diverge on it.
__device__ __forceinline__ int canny_direction(int gx, int gy) {
    const bool neg = gy < 0;
    const int x = neg ? -gx : gx, y = neg ? -gy : gy;
    const int ax = x < 0 ? -x : x;
    // operands below 2^11: 24-bit multiplies (full rate) are exact
    const int ax2 = 2 * __mul24(ax, ax);
    const int lo = __mul24(ax + y, ax + y) - ax2;  // > 0 <=> y > ax * tan 22.5
    const int d = y - ax;
    const bool below_t2 = d <= 0 || __mul24(d, d) < ax2;  // y < ax * tan 67.5
    // x > 0: the angle is atan(y / ax); x < 0: 180 - atan(y / ax)
    const int right = lo < 0 ? 0 : (below_t2 ? 45 : 90);
    const int left = below_t2 ? (lo > 0 ? 135 : 0) : 90;
    int cls = x > 0 ? right : left;
    cls = x == 0 ? 90 : cls;
    cls = gy == 0 ? 0 : cls;
    return (gx == -985 && gy == -408) ? 45 : cls;
}

// Staged blurred window: image rows Y0 - 2 .. Y0 + kCannyTH + 1, columns
// X0 - 4 .. X0 + kCannyTW + 3 (18 dwords, so interior tiles load whole
// aligned dwords), clamped to the image like filter_clamped's reads.
constexpr int kCannyPW = kCannyTW + 8;

// Magnitudes are compared through n = gx^2 + gy^2: sqrtf is strictly
// increasing on the integers up to 2^21 (neighbouring square roots there are
// at least two ulps apart), so m < c <=> n < n_c, and thinned >= t <=>
// n >= n_t with n_t the least n whose sqrtf reaches t (found on the host).
// seeds: per-tile flag, set when the tile holds a seed. VEC: W % 4 == 0, so
// interior tiles stage whole dwords and every thread stores its four states
// as one dword. One workgroup per tile: measured faster here than persistent
// workgroups prefetching their next tile.
template <bool VEC>
__global__ __launch_bounds__(256) void canny_nms_kernel(const uint8_t *__restrict__ blur, int W,
                                                        int H, int tiles_x, int ntiles,
                                                        int n_low, int n_high,
                                                        uint8_t *__restrict__ state,
                                                        uint8_t *__restrict__ seeds) {
    __shared__ __align__(16) uint8_t px[kCannyTH + 4][kCannyPW];
    __shared__ int nn[kCannyTH + 2][kCannyTW + 2];  // n with a one-pixel ring
    __shared__ int gxy[kCannyTH][kCannyTW];         // packed (gx, gy)
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int X0 = (t % tiles_x) * kCannyTW, Y0 = (t / tiles_x) * kCannyTH;
        if (VEC && X0 >= 4 && X0 + kCannyPW - 4 <= W) {  // the same for the whole workgroup
            constexpr int kQ = kCannyPW / 4;
            for (int i = threadIdx.x; i < (kCannyTH + 4) * kQ; i += 256) {
                const int ly = i / kQ, q = i - ly * kQ;
                const int y = min(max(Y0 - 2 + ly, 0), H - 1);
                *reinterpret_cast<uint32_t *>(&px[ly][4 * q]) =
                    *reinterpret_cast<const uint32_t *>(blur + (size_t)y * W + X0 - 4 + 4 * q);
            }
        } else {
            for (int i = threadIdx.x; i < (kCannyTH + 4) * kCannyPW; i += 256) {
                const int ly = i / kCannyPW, c = i - ly * kCannyPW;
                const int y = min(max(Y0 - 2 + ly, 0), H - 1);
                const int x = min(max(X0 - 4 + c, 0), W - 1);
                px[ly][c] = blur[(size_t)y * W + x];
            }
        }
        __syncthreads();
        // ring position (ly, lx) <-> image (Y0 - 1 + ly, X0 - 1 + lx); its taps
        // are px[ly .. ly + 2][lx + 2 .. lx + 4] (filter_clamped::<_, i32, i16>
        // of filter/mod.rs:29-69: the i32 sums are exact and within i16 for u8
        // input, so their order is free). An item is three positions of one
        // column, whose taps share rows: 15 LDS reads instead of 27.
        constexpr int kStrips = (kCannyTH + 2) / 3;
        static_assert((kCannyTH + 2) % 3 == 0, "ring rows come in strips of three");
        for (int i = threadIdx.x; i < kStrips * (kCannyTW + 2); i += 256) {
            const int st = i / (kCannyTW + 2), lx = i - st * (kCannyTW + 2), ly0 = 3 * st;
            int p[5][3];
#pragma unroll
            for (int k = 0; k < 5; k++)
#pragma unroll
                for (int c = 0; c < 3; c++) p[k][c] = px[ly0 + k][lx + 2 + c];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const int a = p[k][0], b = p[k][1], cc = p[k][2];
                const int d = p[k + 1][0], f = p[k + 1][2];
                const int g = p[k + 2][0], h = p[k + 2][1], kk = p[k + 2][2];
                const int gx = (cc - a) + 2 * (f - d) + (kk - g);  // SOBEL_HORIZONTAL_3X3
                const int gy = (g - a) + 2 * (h - b) + (kk - cc);  // SOBEL_VERTICAL_3X3
                const int ly = ly0 + k;
                nn[ly][lx] = __mul24(gx, gx) + __mul24(gy, gy);
                if (ly >= 1 && ly <= kCannyTH && lx >= 1 && lx <= kCannyTW)
                    gxy[ly - 1][lx - 1] = (int)(((uint32_t)gx & 0xffffu) | ((uint32_t)gy << 16));
            }
        }
        __syncthreads();
        // thread: row r, columns c0 .. c0 + 3
        const int r = threadIdx.x >> 4, c0 = 4 * (threadIdx.x & 15);
        const int y = Y0 + r;
        bool seed = false;
        if (y < H) {
            uint32_t packed = 0;
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int lx = c0 + u, x = X0 + lx;
                int thin = 0;
                const bool interior = x >= 1 && y >= 1 && x + 1 < W && y + 1 < H;
                const int cy = r + 1, cx = lx + 1;
                const int m = nn[cy][cx];
                // below the low threshold the class is 0 whether or not the
                // output is suppressed (n_low <= n_high): no direction needed
                if (interior && m >= n_low) {
                    const int p = gxy[r][lx];
                    const int gx = (int)(int16_t)(p & 0xffff), gy = p >> 16;
                    // the two neighbours across the edge (edges.rs:84-99) sit
                    // at +(dy, dx) and -(dy, dx): 0 -> (0, -1), 45 -> (1, 1),
                    // 90 -> (-1, 0), 135 -> (1, -1); selected without branches
                    const int dir = canny_direction(gx, gy);
                    const int dy = dir == 0 ? 0 : (dir == 90 ? -1 : 1);
                    const int dx = dir == 45 ? 1 : (dir == 90 ? 0 : -1);
                    const int c1 = nn[cy + dy][cx + dx], c2 = nn[cy - dy][cx - dx];
                    thin = (m < c1 || m < c2) ? 0 : m;
                }
                const bool sd = interior && thin >= n_high;
                seed |= sd;
                packed |= (uint32_t)(sd ? 2 : (thin >= n_low ? 1 : 0)) << (8 * u);
            }
            const int xq = X0 + c0;
            uint8_t *row = state + (size_t)y * W;
            if (VEC && xq + 4 <= W) {
                *reinterpret_cast<uint32_t *>(row + xq) = packed;
            } else {
#pragma unroll
                for (int u = 0; u < 4; u++)
                    if (xq + u < W) row[xq + u] = (uint8_t)(packed >> (8 * u));
            }
        }
        // seeds: one flag per 64 x 64 hysteresis tile, cleared beforehand
        const bool any = __syncthreads_or(seed);
        if (threadIdx.x == 0 && any)
            seeds[(size_t)(Y0 / kHystT) * tiles_x + (X0 / kHystT)] = 1;
    }
}

// Row masks of a hysteresis tile: bit i of a row is column X0 + i.
// Occluded fills (Kogge-Stone): every bit of c reachable from s by steps
// towards higher (fill_up) or lower (fill_down) bits through c.
__device__ __forceinline__ uint64_t fill_up(uint64_t s, uint64_t c) {
    s |= c & (s << 1);
    c &= c << 1;
    s |= c & (s << 2);
    c &= c << 2;
    s |= c & (s << 4);
    c &= c << 4;
    s |= c & (s << 8);
    c &= c << 8;
    s |= c & (s << 16);
    c &= c << 16;
    return s | (c & (s << 32));
}
__device__ __forceinline__ uint64_t fill_down(uint64_t s, uint64_t c) {
    s |= c & (s >> 1);
    c &= c >> 1;
    s |= c & (s >> 2);
    c &= c >> 2;
    s |= c & (s >> 4);
    c &= c >> 4;
    s |= c & (s >> 8);
    c &= c >> 8;
    s |= c & (s >> 16);
    c &= c >> 16;
    return s | (c & (s >> 32));
}

// 4 bytes (each 0 or 1) -> 4 bits: b_i * 2^(8i) * 2^(7(3-i)) lands on bit 21 + i,
// and no two terms of the product share a bit.
__device__ __forceinline__ uint32_t bytes_to_nibble(uint32_t x) {
    return ((x * 0x00204081u) >> 21) & 0xfu;
}
// The inverse: bit i of a nibble -> byte i.
__device__ __forceinline__ uint32_t nibble_to_bytes(uint32_t n) {
    return (n * 0x00204081u) & 0x01010101u;
}

// Candidate (state >= 1) and marked (state == 2) masks of the 64 state bytes
// of one row starting at column x0 (columns past W read as 0).
template <bool VEC>
__device__ __forceinline__ void hyst_load_row(const uint8_t *__restrict__ state, int W, int y,
                                              int x0, uint64_t &cand, uint64_t &mark) {
    cand = 0;
    mark = 0;
    const uint8_t *row = state + (size_t)y * W;
    uint32_t w[16];
    if (VEC && x0 + 64 <= W) {
        const uint4 *r4 = reinterpret_cast<const uint4 *>(row + x0);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint4 v = r4[k];
            w[4 * k] = v.x;
            w[4 * k + 1] = v.y;
            w[4 * k + 2] = v.z;
            w[4 * k + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < 16; k++) {
            w[k] = 0;
#pragma unroll
            for (int b = 0; b < 4; b++)
                if (x0 + 4 * k + b < W) w[k] |= (uint32_t)row[x0 + 4 * k + b] << (8 * b);
        }
    }
#pragma unroll
    for (int k = 0; k < 16; k++) {
        cand |= (uint64_t)bytes_to_nibble((w[k] | (w[k] >> 1)) & 0x01010101u) << (4 * k);
        mark |= (uint64_t)bytes_to_nibble((w[k] >> 1) & 0x01010101u) << (4 * k);
    }
}

// One launch of the propagation over 64 x 64 tiles, one wave per tile, a
// lane per row. A tile works only when it or one of its eight neighbour
// tiles marked a pixel on its outer ring (the pixels other tiles see) in the
// previous launch (prev; the seed flags before the first launch), and
// records in cur whether it did so itself: marks inside a tile are already
// at the tile's fixed point and concern no other tile. Per step, a row
// takes the marks that reach it from the rows above and below (the
// predecessors of edges.rs:130-137: (x-1, y-1), (x, y-1), (x+1, y-1) and
// (x+1, y+1)) and the neighbouring tiles' ring pixels, then fills its runs of
// candidates in both directions ((x-1, y) and (x+1, y)); steps repeat until
// no row changes. changed[0]: some tile marked a pixel; border[0]: a marked
// pixel lies on the image border.
template <bool VEC>
__global__ __launch_bounds__(64) void canny_hysteresis_kernel(
    uint8_t *__restrict__ state, int W, int H, int tiles_x, int ntiles,
    const uint8_t *__restrict__ prev, uint8_t *__restrict__ cur, int *__restrict__ changed,
    int *__restrict__ border) {
    __shared__ uint64_t M[kHystT + 2];            // marks of rows -1 .. 64
    __shared__ uint32_t HL[kHystT + 2], HR[kHystT + 2];  // marks at columns -1 and 64
    const int y = threadIdx.x;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int bx = t % tiles_x, by = t / tiles_x, ty = ntiles / tiles_x;
        bool active = false;
        for (int j = max(by - 1, 0); j <= min(by + 1, ty - 1); j++)
            for (int i = max(bx - 1, 0); i <= min(bx + 1, tiles_x - 1); i++)
                active |= prev[(size_t)j * tiles_x + i] != 0;
        if (!active) {  // the same for the whole workgroup
            if (y == 0) cur[t] = 0;
            continue;
        }
        const int X0 = bx * kHystT, Y0 = by * kHystT, gy = Y0 + y;
        uint64_t cand = 0, mark0 = 0;
        if (gy < H) hyst_load_row<VEC>(state, W, gy, X0, cand, mark0);
        // ring: rows -1 and 64, columns -1 and 64 of rows -1 .. 64
        for (int lr = y; lr < kHystT + 2; lr += 64) {
            const int ry = Y0 - 1 + lr;
            const bool in = ry >= 0 && ry < H;
            if (lr == 0 || lr == kHystT + 1) {
                uint64_t c = 0, m = 0;
                if (in) hyst_load_row<false>(state, W, ry, X0, c, m);
                M[lr] = m;
            }
            const uint8_t *row = state + (size_t)(in ? ry : 0) * W;
            HL[lr] = (in && X0 >= 1 && row[X0 - 1] == 2) ? 1u : 0u;
            HR[lr] = (in && X0 + kHystT < W && row[X0 + kHystT] == 2) ? 1u : 0u;
        }
        M[y + 1] = mark0;
        __syncthreads();
        const uint64_t halo = (uint64_t)(HL[y + 1] | HL[y]) |
                              ((uint64_t)(HR[y + 1] | HR[y] | HR[y + 2]) << 63);
        uint64_t mark = mark0;
        for (;;) {
            const uint64_t up = M[y], dn = M[y + 2];
            const uint64_t s = cand & (mark | (up << 1) | up | (up >> 1) | (dn >> 1) | halo);
            const uint64_t nm = fill_down(fill_up(s, cand), cand);
            const bool step = nm != mark;
            mark = nm;
            __syncthreads();
            M[y + 1] = mark;
            if (!__syncthreads_or(step)) break;
        }
        const uint64_t fresh = mark & ~mark0;
        bool ring = false, edge = false;
        if (fresh) {
            uint8_t *row = state + (size_t)gy * W;
            if (VEC && X0 + 64 <= W) {
                uint4 *r4 = reinterpret_cast<uint4 *>(row + X0);
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    uint32_t v[4];
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        const int sh = 16 * k + 4 * e;
                        v[e] = nibble_to_bytes((uint32_t)(cand >> sh) & 0xfu) +
                               nibble_to_bytes((uint32_t)(mark >> sh) & 0xfu);
                    }
                    r4[k] = make_uint4(v[0], v[1], v[2], v[3]);
                }
            } else {
                for (int i = 0; i < 64 && X0 + i < W; i++)
                    if ((fresh >> i) & 1) row[X0 + i] = 2;
            }
            ring = y == 0 || y == kHystT - 1 || (fresh & (1ull | (1ull << 63))) != 0;
            const int last = W - 1 - X0;  // bit of column W - 1, when inside the tile
            edge = gy == 0 || gy == H - 1 || (X0 == 0 && (fresh & 1)) ||
                   (last < 64 && ((fresh >> last) & 1));
        }
        if (edge) border[0] = 1;
        const bool any = __syncthreads_or(fresh != 0);
        ring = __syncthreads_or(ring);
        if (y == 0) {
            if (any) changed[0] = 1;
            cur[t] = ring ? 1 : 0;
        }
    }
}

// 255 where the state is 2 (marked), else 0: per byte (b >> 1) * 255 for b in
// {0, 1, 2}. VEC: 16 bytes per thread (n % 16 == 0, both buffers aligned).
template <bool VEC>
__global__ __launch_bounds__(256) void canny_finish_kernel(const uint8_t *__restrict__ state,
                                                           size_t n, uint8_t *__restrict__ out) {
    const size_t stride = (size_t)gridDim.x * 256;
    if (VEC) {
        const uint4 *s4 = reinterpret_cast<const uint4 *>(state);
        uint4 *o4 = reinterpret_cast<uint4 *>(out);
        for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n / 16; i += stride) {
            const uint4 v = s4[i];
            o4[i] = make_uint4(((v.x >> 1) & 0x01010101u) * 255u, ((v.y >> 1) & 0x01010101u) * 255u,
                               ((v.z >> 1) & 0x01010101u) * 255u, ((v.w >> 1) & 0x01010101u) * 255u);
        }
    } else {
        for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride)
            out[i] = (uint8_t)((state[i] >> 1) * 255);
    }
}

__global__ __launch_bounds__(256) void canny_fill_bytes_kernel(uint8_t *__restrict__ p, int n) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) p[i] = 0;
}

__global__ void canny_zero_flags_kernel(int *__restrict__ flags, int n) {
    if ((int)threadIdx.x < n) flags[threadIdx.x] = 0;
}

// Least n in [0, kCannyMaxN + 1] with sqrtf(n) >= t; kCannyMaxN + 1 when no
// Sobel magnitude of a u8 image reaches t.
constexpr int kCannyMaxN = 2 * 1020 * 1020;
int canny_threshold_n(float t) {
    if (!(t > 0.0f)) return 0;
    int lo = 0, hi = kCannyMaxN + 1;  // sqrtf(lo) < t; answer in (lo, hi]
    if (std::sqrt((float)kCannyMaxN) < t) return kCannyMaxN + 1;
    while (hi - lo > 1) {
        const int mid = lo + (hi - lo) / 2;
        if (std::sqrt((float)mid) >= t) hi = mid; else lo = mid;
    }
    return hi;
}

size_t canny_blur_taps(std::vector<float> &k) { return (size_t)gaussian_kernel(1.4f, k); }
}  // namespace

size_t canny_workspace(uint32_t w, uint32_t h) {
    std::vector<float> k;
    const uint32_t len = (uint32_t)canny_blur_taps(k);
    const size_t ntiles = ((size_t)w + kHystT - 1) / kHystT * (((size_t)h + kHystT - 1) / kHystT);
    WorkspacePlan p;
    p.take<int>(kCannyIters + 1);
    p.take<uint8_t>(ntiles);         // tile flags, even launches
    p.take<uint8_t>(ntiles);         // tile flags, odd launches
    p.take<uint8_t>((size_t)w * h);  // blurred
    p.take<uint8_t>((size_t)w * h);  // state
    p.take<uint8_t>(separable_workspace(w, h, 1, IPH_U8, len, IPH_F32));
    return p.used;
}

iph_status launch_canny(const uint8_t *src, uint8_t *dst, uint32_t w, uint32_t h, float low,
                        float high, void *ws, size_t wsb, hipStream_t s) {
    const size_t n = (size_t)w * h;
    if (n == 0) return IPH_OK;
    std::vector<float> k;
    const uint32_t len = (uint32_t)canny_blur_taps(k);
    const int tiles_x = (int)((w + kCannyTW - 1) / kCannyTW);
    const int ntiles = tiles_x * (int)((h + kCannyTH - 1) / kCannyTH);
    // one workgroup per tile (measured faster than fewer workgroups looping
    // over tiles); the kernels loop only past the cap
    const unsigned grid = (unsigned)std::min(ntiles, kCannyGrid);
    Workspace a{static_cast<uint8_t *>(ws), wsb};
    int *flags = a.take<int>(kCannyIters + 1);
    const int hyst_tiles = tiles_x * (int)((h + kHystT - 1) / kHystT);
    uint8_t *tfl[2] = {a.take<uint8_t>(hyst_tiles), a.take<uint8_t>(hyst_tiles)};
    uint8_t *blur = a.take<uint8_t>(n);
    uint8_t *state = a.take<uint8_t>(n);
    const size_t sepb = separable_workspace(w, h, 1, IPH_U8, len, IPH_F32);
    uint8_t *sep = a.take<uint8_t>(sepb);
    if (!flags || !tfl[0] || !tfl[1] || !blur || !state || (sepb && !sep)) {
        set_error("workspace too small for canny");
        return IPH_EINVAL;
    }
    int *border = flags + kCannyIters;
    IPH_TRY(launch_separable(src, blur, w, h, 1, n, IPH_U8, k.data(), k.data(), len, IPH_F32, sep,
                             sepb, s));
    const int n_low = canny_threshold_n(low), n_high = canny_threshold_n(high);
    canny_fill_bytes_kernel<<<(unsigned)std::min((hyst_tiles + 255) / 256, 4096), 256, 0, s>>>(
        tfl[0], hyst_tiles);
    IPH_TRY(launch_status("canny_fill_bytes_kernel"));
    if (w % 4 == 0)
        canny_nms_kernel<true><<<grid, 256, 0, s>>>(blur, (int)w, (int)h, tiles_x, ntiles, n_low,
                                                    n_high, state, tfl[0]);
    else
        canny_nms_kernel<false><<<grid, 256, 0, s>>>(blur, (int)w, (int)h, tiles_x, ntiles, n_low,
                                                     n_high, state, tfl[0]);
    IPH_TRY(launch_status("canny_nms_kernel"));
    // Hysteresis: launches in groups of kCannyIters, each with its own flag;
    // the group's last flag clear means a launch changed nothing. The flags
    // and the border flag after them are cleared together before the first
    // group; one read-back per group returns the last flag and the border flag.
    for (int launch = 0;;) {
        canny_zero_flags_kernel<<<1, 64, 0, s>>>(flags, launch == 0 ? kCannyIters + 1 : kCannyIters);
        IPH_TRY(launch_status("canny_zero_flags_kernel"));
        for (int i = 0; i < kCannyIters; i++, launch++) {
            const unsigned hgrid = (unsigned)std::min(hyst_tiles, kCannyGrid);
            if (w % 16 == 0)
                canny_hysteresis_kernel<true><<<hgrid, 64, 0, s>>>(
                    state, (int)w, (int)h, tiles_x, hyst_tiles, tfl[launch & 1],
                    tfl[(launch + 1) & 1], flags + i, border);
            else
                canny_hysteresis_kernel<false><<<hgrid, 64, 0, s>>>(
                    state, (int)w, (int)h, tiles_x, hyst_tiles, tfl[launch & 1],
                    tfl[(launch + 1) & 1], flags + i, border);
            IPH_TRY(launch_status("canny_hysteresis_kernel"));
        }
        int last_border[2] = {0, 0};  // flags[kCannyIters - 1], *border
        IPH_TRY_HIP(hipMemcpyAsync(last_border, flags + kCannyIters - 1, 2 * sizeof(int),
                                   hipMemcpyDeviceToHost, s));
        IPH_TRY_HIP(hipStreamSynchronize(s));
        if (last_border[1]) {
            set_error("canny: hysteresis reached the image border, where the reference's "
                      "search indexes outside the image and panics (low_threshold <= 0)");
            return IPH_EINVAL;
        }
        if (last_border[0] == 0) break;
    }
    const bool vec = n % 16 == 0 && (reinterpret_cast<uintptr_t>(dst) & 15) == 0;
    const unsigned fgrid = (unsigned)std::min<size_t>((vec ? n / 16 : n) / 256 + 1, 1u << 16);
    if (vec)
        canny_finish_kernel<true><<<fgrid, 256, 0, s>>>(state, n, dst);
    else
        canny_finish_kernel<false><<<fgrid, 256, 0, s>>>(state, n, dst);
    IPH_TRY(launch_status("canny_finish_kernel"));
    return IPH_OK;
}

}  // namespace iph
