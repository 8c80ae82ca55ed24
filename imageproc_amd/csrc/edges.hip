// gfx950 kernels for imageproc's canny (edges.rs:30-155).
//
//   1. gaussian_blur_f32(image, 1.4) with the filter module's launcher;
//   2. canny_nms_kernel: per 64 x 16 tile, the Sobel responses of the blurred
//      tile plus a one-pixel ring (filter_clamped::<_, i32, i16>, clamp-to-
//      edge reads), their squared magnitudes in LDS, non_maximum_suppression
//      for the tile's pixels, and the hysteresis class of every pixel: 2 =
//      seed (interior, thinned >= high), 1 = candidate (thinned >= low), 0 =
//      none; plus a per-tile flag for tiles holding a seed;
//   3. canny_hysteresis_kernel, repeated until a launch marks nothing: a
//      candidate becomes marked when one of the pixels whose six-neighbour
//      list (edges.rs:130-137) holds it is marked. Each launch iterates its
//      tile in LDS to a local fixed point, so marks cross one tile border per
//      launch at least; a tile whose 3 x 3 tile neighbourhood marked nothing
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
constexpr int kCannyIters = 4;  // hysteresis launches between flag checks
constexpr int kCannyGrid = 1 << 20;  // workgroup cap of the tile kernels

__device__ __forceinline__ int canny_direction(int gx, int gy) {
    if (gx == -985 && gy == -408) return 45;
    if (gy == 0) return 0;
    const int x = gy < 0 ? -gx : gx, y = gy < 0 ? -gy : gy;
    if (x == 0) return 90;
    const int ax = x < 0 ? -x : x;
    const int lo = (ax + y) * (ax + y) - 2 * ax * ax;  // > 0 <=> y > ax * tan 22.5
    const int d = y - ax;
    const bool below_t2 = d <= 0 || d * d < 2 * ax * ax;  // y < ax * tan 67.5
    if (x > 0) {
        if (lo < 0) return 0;
        return below_t2 ? 45 : 90;
    }
    // x < 0: the angle is 180 - atan(y / ax)
    if (!below_t2) return 90;
    return lo > 0 ? 135 : 0;
}

using CannyPx = int16_t[kCannyTW + 4];

// Sobel responses with taps px[ly .. ly + 2][lx .. lx + 2] (filter/mod.rs:29-69
// order; the i32 sums are exact and within i16 for u8 input).
__device__ __forceinline__ void canny_sobel(const CannyPx *px, int ly, int lx, int &gx, int &gy) {
    const int a = px[ly][lx], b = px[ly][lx + 1], c = px[ly][lx + 2];
    const int d = px[ly + 1][lx], f = px[ly + 1][lx + 2];
    const int g = px[ly + 2][lx], h = px[ly + 2][lx + 1], k = px[ly + 2][lx + 2];
    gx = (c - a) + 2 * (f - d) + (k - g);  // SOBEL_HORIZONTAL_3X3
    gy = (g - a) + 2 * (h - b) + (k - c);  // SOBEL_VERTICAL_3X3
}

// Magnitudes are compared through n = gx^2 + gy^2: sqrtf is strictly
// increasing on the integers up to 2^21 (neighbouring square roots there are
// at least two ulps apart), so m < c <=> n < n_c, and thinned >= t <=>
// n >= n_t with n_t the least n whose sqrtf reaches t (found on the host).
// seeds: per-tile flag, set when the tile holds a seed.
__global__ __launch_bounds__(256) void canny_nms_kernel(const uint8_t *__restrict__ blur, int W,
                                                        int H, int tiles_x, int ntiles,
                                                        int n_low, int n_high,
                                                        uint8_t *__restrict__ state,
                                                        uint8_t *__restrict__ seeds) {
    // blurred tile with a two-pixel ring; n and packed (gx, gy) with a one-pixel ring
    __shared__ int16_t px[kCannyTH + 4][kCannyTW + 4];
    __shared__ int nn[kCannyTH + 2][kCannyTW + 2];
    __shared__ int gxy[kCannyTH][kCannyTW];
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int X0 = (t % tiles_x) * kCannyTW, Y0 = (t / tiles_x) * kCannyTH;
        for (int i = threadIdx.x; i < (kCannyTH + 4) * (kCannyTW + 4); i += 256) {
            const int ly = i / (kCannyTW + 4), lx = i % (kCannyTW + 4);
            const int y = min(max(Y0 - 2 + ly, 0), H - 1), x = min(max(X0 - 2 + lx, 0), W - 1);
            px[ly][lx] = blur[(size_t)y * W + x];
        }
        __syncthreads();
        // ring position (ly, lx) <-> image (Y0 - 1 + ly, X0 - 1 + lx)
        for (int i = threadIdx.x; i < (kCannyTH + 2) * (kCannyTW + 2); i += 256) {
            const int ly = i / (kCannyTW + 2), lx = i % (kCannyTW + 2);
            int gx, gy;
            canny_sobel(px, ly, lx, gx, gy);
            nn[ly][lx] = gx * gx + gy * gy;
            if (ly >= 1 && ly <= kCannyTH && lx >= 1 && lx <= kCannyTW)
                gxy[ly - 1][lx - 1] = (int)(((uint32_t)gx & 0xffffu) | ((uint32_t)gy << 16));
        }
        __syncthreads();
        const int lx = threadIdx.x & 63, x = X0 + lx;
        bool seed = false;
        if (x < W) {
            for (int r = threadIdx.x >> 6; r < kCannyTH; r += 4) {
                const int y = Y0 + r;
                if (y >= H) break;
                int thin = 0;
                const bool interior = x >= 1 && y >= 1 && x + 1 < W && y + 1 < H;
                if (interior) {
                    const int p = gxy[r][lx];
                    const int gx = (int)(int16_t)(p & 0xffff), gy = p >> 16;
                    const int cy = r + 1, cx = lx + 1;
                    int c1, c2;
                    switch (canny_direction(gx, gy)) {
                    case 0: c1 = nn[cy][cx - 1]; c2 = nn[cy][cx + 1]; break;
                    case 45: c1 = nn[cy + 1][cx + 1]; c2 = nn[cy - 1][cx - 1]; break;
                    case 90: c1 = nn[cy - 1][cx]; c2 = nn[cy + 1][cx]; break;
                    default: c1 = nn[cy + 1][cx - 1]; c2 = nn[cy - 1][cx + 1]; break;
                    }
                    const int m = nn[cy][cx];
                    thin = (m < c1 || m < c2) ? 0 : m;
                }
                const bool sd = interior && thin >= n_high;
                seed |= sd;
                state[(size_t)y * W + x] = sd ? 2 : (thin >= n_low ? 1 : 0);
            }
        }
        const bool any = __syncthreads_or(seed);
        if (threadIdx.x == 0) seeds[t] = any ? 1 : 0;
    }
}

// One launch of the propagation. A tile works only when it or one of its
// eight neighbour tiles marked a pixel on its outer ring (the pixels other
// tiles see) in the previous launch (prev; the seed flags before the first
// launch), and records in cur whether it did so itself: marks inside a tile
// are already at the tile's fixed point and concern no other tile.
// changed[0]: some tile marked a pixel; border[0]: a marked pixel lies on the
// image border.
__global__ __launch_bounds__(256) void canny_hysteresis_kernel(
    uint8_t *__restrict__ state, int W, int H, int tiles_x, int ntiles,
    const uint8_t *__restrict__ prev, uint8_t *__restrict__ cur, int *__restrict__ changed,
    int *__restrict__ border) {
    __shared__ uint8_t st[kCannyTH + 2][kCannyTW + 2];
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int bx = t % tiles_x, by = t / tiles_x, tx = tiles_x, ty = ntiles / tiles_x;
        bool active = false;
        for (int j = max(by - 1, 0); j <= min(by + 1, ty - 1); j++)
            for (int i = max(bx - 1, 0); i <= min(bx + 1, tx - 1); i++)
                active |= prev[(size_t)j * tx + i] != 0;
        if (!active) {  // the same for the whole workgroup
            if (threadIdx.x == 0) cur[t] = 0;
            continue;
        }
        const int X0 = bx * kCannyTW, Y0 = by * kCannyTH;
        for (int i = threadIdx.x; i < (kCannyTH + 2) * (kCannyTW + 2); i += 256) {
            const int ly = i / (kCannyTW + 2), lx = i % (kCannyTW + 2);
            const int y = Y0 - 1 + ly, x = X0 - 1 + lx;
            st[ly][lx] = (x >= 0 && y >= 0 && x < W && y < H) ? state[(size_t)y * W + x] : 0;
        }
        __syncthreads();
        const int lx = (threadIdx.x & 63) + 1, r0 = threadIdx.x >> 6;
        bool mine = false;  // this thread marked one of its pixels
        for (;;) {
            bool step = false;
            for (int r = r0; r < kCannyTH; r += 4) {
                const int ly = r + 1;
                if (st[ly][lx] != 1) continue;
                // the pixels whose neighbour list (edges.rs:130-137) holds this one
                if (st[ly][lx - 1] == 2 || st[ly - 1][lx - 1] == 2 || st[ly - 1][lx] == 2 ||
                    st[ly + 1][lx + 1] == 2 || st[ly][lx + 1] == 2 || st[ly - 1][lx + 1] == 2) {
                    st[ly][lx] = 2;
                    step = true;
                }
            }
            mine |= step;
            if (!__syncthreads_or(step)) break;
        }
        const bool any = __syncthreads_or(mine);
        bool ring = false;
        if (any) {
            const int x = X0 + lx - 1;
            bool edge = false;
            for (int r = r0; r < kCannyTH; r += 4) {
                const int y = Y0 + r;
                if (x < W && y < H && st[r + 1][lx] == 2) {
                    // only pixels marked in this launch: state still holds 1 there
                    if (state[(size_t)y * W + x] != 2) {
                        state[(size_t)y * W + x] = 2;
                        edge |= x == 0 || y == 0 || x == W - 1 || y == H - 1;
                        ring |= r == 0 || r == kCannyTH - 1 || lx == 1 || lx == kCannyTW;
                    }
                }
            }
            if (edge) border[0] = 1;
            if (threadIdx.x == 0) changed[0] = 1;
        }
        ring = __syncthreads_or(ring);
        if (threadIdx.x == 0) cur[t] = ring ? 1 : 0;
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
    const size_t ntiles = ((size_t)w + kCannyTW - 1) / kCannyTW * (((size_t)h + kCannyTH - 1) / kCannyTH);
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
    uint8_t *tfl[2] = {a.take<uint8_t>(ntiles), a.take<uint8_t>(ntiles)};
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
    canny_nms_kernel<<<grid, 256, 0, s>>>(blur, (int)w, (int)h, tiles_x, ntiles,
                                          canny_threshold_n(low), canny_threshold_n(high), state,
                                          tfl[0]);
    IPH_TRY(launch_status("canny_nms_kernel"));
    canny_zero_flags_kernel<<<1, 64, 0, s>>>(border, 1);
    IPH_TRY(launch_status("canny_zero_flags_kernel"));
    // Hysteresis: launches in groups of kCannyIters, each with its own flag;
    // the group's last flag clear means a launch changed nothing.
    for (int launch = 0;;) {
        canny_zero_flags_kernel<<<1, 64, 0, s>>>(flags, kCannyIters);
        IPH_TRY(launch_status("canny_zero_flags_kernel"));
        for (int i = 0; i < kCannyIters; i++, launch++) {
            canny_hysteresis_kernel<<<grid, 256, 0, s>>>(state, (int)w, (int)h, tiles_x, ntiles,
                                                         tfl[launch & 1], tfl[(launch + 1) & 1],
                                                         flags + i, border);
            IPH_TRY(launch_status("canny_hysteresis_kernel"));
        }
        int last = 0;
        IPH_TRY_HIP(hipMemcpyAsync(&last, flags + kCannyIters - 1, sizeof(int),
                                   hipMemcpyDeviceToHost, s));
        IPH_TRY_HIP(hipStreamSynchronize(s));
        if (last == 0) break;
    }
    const bool vec = n % 16 == 0 && (reinterpret_cast<uintptr_t>(dst) & 15) == 0;
    const unsigned fgrid = (unsigned)std::min<size_t>((vec ? n / 16 : n) / 256 + 1, 1u << 16);
    if (vec)
        canny_finish_kernel<true><<<fgrid, 256, 0, s>>>(state, n, dst);
    else
        canny_finish_kernel<false><<<fgrid, 256, 0, s>>>(state, n, dst);
    IPH_TRY(launch_status("canny_finish_kernel"));
    int hit = 0;
    IPH_TRY_HIP(hipMemcpyAsync(&hit, border, sizeof(int), hipMemcpyDeviceToHost, s));
    IPH_TRY_HIP(hipStreamSynchronize(s));
    if (hit) {
        set_error("canny: hysteresis reached the image border, where the reference's search "
                  "indexes outside the image and panics (low_threshold <= 0)");
        return IPH_EINVAL;
    }
    return IPH_OK;
}

}  // namespace iph
