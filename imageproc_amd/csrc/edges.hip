// gfx950 kernels for imageproc's canny (edges.rs:30-155).
//
//   1. gaussian_blur_f32(image, 1.4) with the filter module's launcher;
//   2. canny_nms_kernel: per 64 x 16 tile, the Sobel responses of the blurred
//      tile plus a one-pixel ring (filter_clamped::<_, i32, i16>, clamp-to-
//      edge reads), their magnitudes in LDS, non_maximum_suppression for the
//      tile's pixels, and the hysteresis class of every pixel: 2 = seed
//      (interior, thinned >= high), 1 = candidate (thinned >= low), 0 = none;
//   3. canny_hysteresis_kernel, repeated until a launch marks nothing: a
//      candidate becomes marked when one of the pixels whose six-neighbour
//      list (edges.rs:130-137) holds it is marked. Each launch iterates its
//      tile in LDS to a local fixed point, so marks cross one tile border per
//      launch at least. The marked set is the union of what the reference's
//      depth-first searches reach, which does not depend on visiting order;
//   4. canny_finish_kernel: 255 for marked pixels, and a flag when a marked
//      pixel lies on the image border, where the reference's search indexes
//      outside the image and panics.
//
// Magnitude: (gx as f32).hypot(gy as f32). For the Sobel responses of u8
// images |gx|, |gy| <= 1020, gx^2 + gy^2 < 2^24 is exact in f32 and glibc's
// hypotf equals the correctly rounded sqrtf of it over that whole domain.
// Direction: the reference bins atan2f(gy, gx) * (180 / pi) (+180 when
// negative) in f32. Over the same domain that equals the exact angle's bin,
// computed here with integer arithmetic (tan 22.5 = sqrt2 - 1, tan 67.5 =
// sqrt2 + 1, and no integer ratio meets either), except at (gx, gy) =
// (-985, -408), where glibc's atan2f (one ulp from correctly rounded) takes
// the f32 angle to 22.5 exactly, bin 45.
#include "iph_internal.hpp"

#include <algorithm>
#include <vector>

namespace iph {
namespace {

constexpr int kCannyTW = 64, kCannyTH = 16;
constexpr int kCannyIters = 4;  // hysteresis launches between flag checks

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

__global__ __launch_bounds__(256) void canny_nms_kernel(const uint8_t *__restrict__ blur, int W,
                                                        int H, float low, float high,
                                                        uint8_t *__restrict__ state) {
    // blurred tile with a two-pixel ring, magnitudes with a one-pixel ring
    __shared__ int16_t px[kCannyTH + 4][kCannyTW + 4];
    __shared__ float mag[kCannyTH + 2][kCannyTW + 2];
    const int X0 = blockIdx.x * kCannyTW, Y0 = blockIdx.y * kCannyTH;
    for (int i = threadIdx.x; i < (kCannyTH + 4) * (kCannyTW + 4); i += 256) {
        const int ly = i / (kCannyTW + 4), lx = i % (kCannyTW + 4);
        const int y = min(max(Y0 - 2 + ly, 0), H - 1), x = min(max(X0 - 2 + lx, 0), W - 1);
        px[ly][lx] = blur[(size_t)y * W + x];
    }
    __syncthreads();
    // magnitude at ring position (ly, lx) <-> image (Y0 - 1 + ly, X0 - 1 + lx)
    for (int i = threadIdx.x; i < (kCannyTH + 2) * (kCannyTW + 2); i += 256) {
        const int ly = i / (kCannyTW + 2), lx = i % (kCannyTW + 2);
        int gx, gy;
        canny_sobel(px, ly, lx, gx, gy);
        mag[ly][lx] = sqrtf((float)(gx * gx + gy * gy));
    }
    __syncthreads();
    const int lx = threadIdx.x & 63, x = X0 + lx;
    if (x >= W) return;
    for (int r = threadIdx.x >> 6; r < kCannyTH; r += 4) {
        const int y = Y0 + r;
        if (y >= H) break;
        float thin = 0.0f;
        const bool interior = x >= 1 && y >= 1 && x + 1 < W && y + 1 < H;
        if (interior) {
            int gx, gy;
            canny_sobel(px, r + 1, lx + 1, gx, gy);
            const int cy = r + 1, cx = lx + 1;
            float c1, c2;
            switch (canny_direction(gx, gy)) {
            case 0: c1 = mag[cy][cx - 1]; c2 = mag[cy][cx + 1]; break;
            case 45: c1 = mag[cy + 1][cx + 1]; c2 = mag[cy - 1][cx - 1]; break;
            case 90: c1 = mag[cy - 1][cx]; c2 = mag[cy + 1][cx]; break;
            default: c1 = mag[cy + 1][cx - 1]; c2 = mag[cy - 1][cx + 1]; break;
            }
            const float m = mag[cy][cx];
            thin = (m < c1 || m < c2) ? 0.0f : m;
        }
        const uint8_t st = (interior && thin >= high) ? 2 : (thin >= low ? 1 : 0);
        state[(size_t)y * W + x] = st;
    }
}

// One launch of the propagation. changed[0] is set when any pixel was marked.
__global__ __launch_bounds__(256) void canny_hysteresis_kernel(uint8_t *__restrict__ state, int W,
                                                               int H, int *__restrict__ changed) {
    __shared__ uint8_t st[kCannyTH + 2][kCannyTW + 2];
    const int X0 = blockIdx.x * kCannyTW, Y0 = blockIdx.y * kCannyTH;
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
    if (__syncthreads_or(mine)) {
        const int x = X0 + lx - 1;
        for (int r = r0; r < kCannyTH; r += 4) {
            const int y = Y0 + r;
            if (x < W && y < H && st[r + 1][lx] == 2) state[(size_t)y * W + x] = 2;
        }
        if (threadIdx.x == 0) changed[0] = 1;
    }
}

__global__ __launch_bounds__(256) void canny_finish_kernel(const uint8_t *__restrict__ state,
                                                           int W, int H, uint8_t *__restrict__ out,
                                                           int *__restrict__ border_hit) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    if (x >= W) return;
    for (int y = blockIdx.y * 4 + (threadIdx.x >> 6); y < H; y += gridDim.y * 4) {
        const bool on = state[(size_t)y * W + x] == 2;
        out[(size_t)y * W + x] = on ? 255 : 0;
        if (on && (x == 0 || y == 0 || x == W - 1 || y == H - 1)) border_hit[0] = 1;
    }
}

__global__ void canny_zero_flags_kernel(int *__restrict__ flags, int n) {
    if ((int)threadIdx.x < n) flags[threadIdx.x] = 0;
}

size_t canny_blur_taps(std::vector<float> &k) { return (size_t)gaussian_kernel(1.4f, k); }
}  // namespace

size_t canny_workspace(uint32_t w, uint32_t h) {
    std::vector<float> k;
    const uint32_t len = (uint32_t)canny_blur_taps(k);
    WorkspacePlan p;
    p.take<int>(kCannyIters + 1);
    p.take<uint8_t>((size_t)w * h);  // blurred
    p.take<uint8_t>((size_t)w * h);  // state
    p.take<uint8_t>(separable_workspace(w, h, 1, IPH_U8, len, IPH_F32));
    return p.used;
}

iph_status launch_canny(const uint8_t *src, uint8_t *dst, uint32_t w, uint32_t h, float low,
                        float high, void *ws, size_t wsb, hipStream_t s) {
    const size_t n = (size_t)w * h;
    if (n == 0) return IPH_OK;
    if (h > 65535u * kCannyTH) {
        set_error("canny supports heights up to %u", 65535u * kCannyTH);
        return IPH_EUNSUPPORTED;
    }
    std::vector<float> k;
    const uint32_t len = (uint32_t)canny_blur_taps(k);
    Workspace a{static_cast<uint8_t *>(ws), wsb};
    int *flags = a.take<int>(kCannyIters + 1);
    uint8_t *blur = a.take<uint8_t>(n);
    uint8_t *state = a.take<uint8_t>(n);
    const size_t sepb = separable_workspace(w, h, 1, IPH_U8, len, IPH_F32);
    uint8_t *sep = a.take<uint8_t>(sepb);
    if (!flags || !blur || !state || (sepb && !sep)) {
        set_error("workspace too small for canny");
        return IPH_EINVAL;
    }
    IPH_TRY(launch_separable(src, blur, w, h, 1, n, IPH_U8, k.data(), k.data(), len, IPH_F32, sep,
                             sepb, s));
    const dim3 tiles((w + kCannyTW - 1) / kCannyTW, (h + kCannyTH - 1) / kCannyTH);
    canny_nms_kernel<<<tiles, 256, 0, s>>>(blur, (int)w, (int)h, low, high, state);
    IPH_TRY(launch_status("canny_nms_kernel"));
    // Hysteresis: launches in groups of kCannyIters, each with its own flag;
    // the group's last flag clear means a launch changed nothing.
    for (;;) {
        canny_zero_flags_kernel<<<1, 64, 0, s>>>(flags, kCannyIters + 1);
        IPH_TRY(launch_status("canny_zero_flags_kernel"));
        for (int i = 0; i < kCannyIters; i++) {
            canny_hysteresis_kernel<<<tiles, 256, 0, s>>>(state, (int)w, (int)h, flags + i);
            IPH_TRY(launch_status("canny_hysteresis_kernel"));
        }
        int last = 0;
        IPH_TRY_HIP(hipMemcpyAsync(&last, flags + kCannyIters - 1, sizeof(int),
                                   hipMemcpyDeviceToHost, s));
        IPH_TRY_HIP(hipStreamSynchronize(s));
        if (last == 0) break;
    }
    const unsigned gy = std::min((h + 3) / 4, 65535u);
    canny_finish_kernel<<<dim3((w + 63) / 64, gy), 256, 0, s>>>(state, (int)w, (int)h, dst,
                                                                flags + kCannyIters);
    IPH_TRY(launch_status("canny_finish_kernel"));
    int hit = 0;
    IPH_TRY_HIP(hipMemcpyAsync(&hit, flags + kCannyIters, sizeof(int), hipMemcpyDeviceToHost, s));
    IPH_TRY_HIP(hipStreamSynchronize(s));
    if (hit) {
        set_error("canny: hysteresis reached the image border, where the reference's search "
                  "indexes outside the image and panics (low_threshold <= 0)");
        return IPH_EINVAL;
    }
    return IPH_OK;
}

}  // namespace iph
