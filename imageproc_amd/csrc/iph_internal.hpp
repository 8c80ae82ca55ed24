// Shared definitions for the gfx950 kernels and the C ABI host code.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "imageproc_hip.h"

namespace iph {

// Records the message returned by iph_last_error() for this thread.
void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
void clear_error();

#define IPH_TRY_HIP(expr)                                                                    \
    do {                                                                                     \
        hipError_t iph_e_ = (expr);                                                          \
        if (iph_e_ != hipSuccess) {                                                          \
            ::iph::set_error("%s: %s", #expr, hipGetErrorString(iph_e_));                    \
            return IPH_EDEVICE;                                                              \
        }                                                                                    \
    } while (0)

#define IPH_TRY(expr)                                                                        \
    do {                                                                                     \
        iph_status iph_s_ = (expr);                                                          \
        if (iph_s_ != IPH_OK) return iph_s_;                                                 \
    } while (0)

// Checks the launch status of the kernel just enqueued.
inline iph_status launch_status(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("launch of %s failed: %s", what, hipGetErrorString(e));
        return IPH_EDEVICE;
    }
    return IPH_OK;
}

inline size_t elem_size(int32_t t) {
    switch (t) {
    case IPH_U8: return 1;
    case IPH_U16:
    case IPH_I16: return 2;
    default: return 4;
    }
}

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Bump allocator over a caller-provided device workspace.
struct Workspace {
    char *base;
    size_t cap, used;
    Workspace(void *p, size_t n) : base(static_cast<char *>(p)), cap(n), used(0) {}
    template <typename T> T *take(size_t count) {
        size_t off = align_up(used, 256);
        size_t bytes = align_up(count * sizeof(T), 256);
        if (off + bytes > cap || base == nullptr) return nullptr;
        used = off + bytes;
        return reinterpret_cast<T *>(base + off);
    }
};

// Grows a workspace size the same way Workspace::take consumes it.
struct WorkspacePlan {
    size_t used = 0;
    template <typename T> void take(size_t count) {
        used = align_up(used, 256) + align_up(count * sizeof(T), 256);
    }
};

// Largest tap count passed by value in a kernel argument block.
constexpr int kMaxArgTaps = 256;
template <typename K> struct Taps {
    K w[kMaxArgTaps];
};

// ---- launchers implemented in filter.hip --------------------------------

iph_status launch_separable(const void *src, void *dst, uint32_t w, uint32_t h, uint32_t n,
                            size_t stride, int32_t p_type, const void *hk, const void *vk,
                            uint32_t len, int32_t k_type, void *ws, size_t ws_bytes,
                            hipStream_t s);
size_t separable_workspace(uint32_t w, uint32_t h, uint32_t n, int32_t p_type, uint32_t len,
                           int32_t k_type);
iph_status launch_filter1d(const void *src, void *dst, uint32_t w, uint32_t h, int32_t p_type,
                           const void *k, uint32_t len, int32_t k_type, bool vertical,
                           void *ws, size_t ws_bytes, hipStream_t s);
iph_status launch_filter_clamped(const void *src, int32_t p_type, uint32_t w, uint32_t h,
                                 const void *kernel, int32_t k_type, uint32_t kw, uint32_t kh,
                                 void *dst, int32_t s_type, void *ws, size_t ws_bytes,
                                 hipStream_t s);
size_t filter_clamped_workspace(int32_t k_type, uint32_t kw, uint32_t kh);
bool blur_mfma_supported(const float *hk, const float *vk, uint32_t len);
iph_status launch_blur_u8_mfma(const uint8_t *src, uint8_t *dst, uint32_t w, uint32_t h,
                               uint32_t n, size_t stride, const float *hk, const float *vk,
                               uint32_t len, hipStream_t s);
// integral_image / integral_squared_image (integral_image.rs:50-152), u32,
// (w+1) x (h+1); adaptive_threshold (contrast.rs:19-47).
size_t integral_workspace(uint32_t w, uint32_t h);
iph_status launch_integral_image(const uint8_t *src, uint32_t w, uint32_t h, int squared,
                                 uint32_t *out, void *ws, size_t wsb, hipStream_t s);
size_t adaptive_threshold_workspace(uint32_t w, uint32_t h, uint32_t block_radius);
iph_status launch_adaptive_threshold(const uint8_t *src, uint32_t w, uint32_t h,
                                     uint32_t block_radius, int32_t delta, uint8_t *dst, void *ws,
                                     size_t wsb, hipStream_t s);
// canny (edges.rs:30-155), implemented in edges.hip; synchronises the stream
// once per few hysteresis launches to test for convergence.
size_t canny_workspace(uint32_t w, uint32_t h);
iph_status launch_canny(const uint8_t *src, uint8_t *dst, uint32_t w, uint32_t h, float low,
                        float high, void *ws, size_t wsb, hipStream_t s);
// gradients (gradients.rs:99-160) of a GrayImage, u16 magnitudes.
iph_status launch_gradients(const uint8_t *src, uint32_t w, uint32_t h, const int32_t *hk,
                            uint32_t hkw, uint32_t hkh, const int32_t *vk, uint32_t vkw,
                            uint32_t vkh, uint16_t *out, hipStream_t s);
// sharpen_gaussian (filter/sharpen.rs:31-38): f32 image, f32 blur, combine.
size_t sharpen_gaussian_workspace(uint32_t w, uint32_t h, uint32_t taps);
iph_status launch_sharpen_gaussian(const uint8_t *src, uint8_t *dst, uint32_t w, uint32_t h,
                                   const float *taps, uint32_t n_taps, float amount, void *ws,
                                   size_t ws_bytes, hipStream_t s);
iph_status launch_box_filter(const uint8_t *src, uint8_t *dst, uint32_t w, uint32_t h,
                             uint32_t rx, uint32_t ry, void *ws, size_t ws_bytes,
                             hipStream_t s);

// ---- argument checks and host helpers shared by the C ABI (capi.hip) -------

// Fails unless a gfx950 device is current.
iph_status require_device();
// gaussian_kernel_f32's taps (host, glibc expf); returns the count or -1.
int32_t gaussian_kernel(float sigma, std::vector<float> &k);
iph_status check_dims(uint32_t w, uint32_t h);
iph_status check_match(int32_t p, uint32_t w, uint32_t h, uint32_t tw, uint32_t th,
                       int32_t method, int32_t path);
// Largest template side of the MFMA match engines (u8 and f32).
constexpr uint32_t kMfmaMaxSide = 128;
// Images the MFMA match engines take: 32-bit byte offsets into the image and
// into 256 rows of the score map.
inline bool mfma_match_fits(uint32_t w, uint32_t h) {
    return (uint64_t)w * h + 64 < (1ull << 32) && w < (1u << 22);
}
int32_t resolve_path(int32_t path, uint32_t w, uint32_t h, uint32_t tw, uint32_t th);

// ---- batches of uneven images (filter.hip, blur_mfma.hip, batch.hip) ------

// filter_clamped<u8, f32, u8> job of a batch (device pointers).
constexpr int kF2MaxJobTaps = 49;
struct FilterSpec {
    const uint8_t *src;
    uint8_t *dst;
    uint32_t w, h, kw, kh;
    float k[kF2MaxJobTaps];
};
int filter_jobs_variants();
const char *filter_jobs_variant_name(int variant);
size_t filter_jobs_table_bytes(int n);
iph_status filter_jobs_table(const FilterSpec *specs, int n, int variant, void *host_table,
                             int *total_tiles);
iph_status launch_filter_jobs(const void *table_dev, int n, int total_tiles, int variant,
                              uint32_t max_kw, uint32_t max_kh, hipStream_t s);

// gaussian_blur_f32 job of a batch (device pointers, host taps).
constexpr int kBlurSpecMaxTaps = 33;
struct BlurSpec {
    const uint8_t *src;
    uint8_t *dst;
    uint32_t w, h, len;
    float taps[kBlurSpecMaxTaps];
};
// Jobs of one kernel variant, consecutive in the device table: the per-wave
// kernel (wave: `first` indexes the table's per-wave entries, ntiles counts
// its workgroups) or the fixed-point tile kernel (ntiles counts its tiles).
struct BlurGroup {
    int nkb, taps_class;
    bool aligned, wave;
    int first, count, ntiles;
};
size_t blur_jobs_table_bytes(int n);
// Fills the host table (grouped by kernel variant) and lists the specs the
// fixed-point kernel does not take in `fallback`.
iph_status blur_jobs_table(const BlurSpec *specs, int n, void *host_table,
                           std::vector<BlurGroup> &groups, std::vector<int> &fallback);
// n: the specs the table was built from (its per-wave part starts after n
// FxJob slots); grid_pct: persistent workgroups of the tile kernel as a
// percentage of the resident maximum.
iph_status launch_blur_jobs(const void *table_dev, int n, const std::vector<BlurGroup> &groups,
                            int grid_pct, hipStream_t s);

// match_template (GrayImage, int8 MFMA with in-kernel window sums) of a batch:
// one entry of the device job table of the fused kernel.
struct MatchJob {
    const uint8_t *img;
    const uint8_t *tpl;
    float *out;
    int W, H, tw, th, ow, oh;
    int tiles_x, tile0;  // 32-column tiles per tile row; first tile in the launch
};
// K blocks of the fused kernel for a template width.
int fused_match_kb(uint32_t tw);
// Largest waves per workgroup (8, 4, 2, 1) whose LDS footprint the fused
// kernel takes for this template and method, 0 if none; and the shape a
// single call picks.
int fused_match_rt_max(uint32_t tw, uint32_t th, int32_t method);
int fused_match_rt_default(uint32_t tw, uint32_t th, int32_t method);
// Template rows per chunk for templates up to th rows high: the fewest
// chunks whose footprint leaves min_waves (12 or 8) waves per CU, else 8;
// and the LDS bytes of one job.
int fused_match_tc(int kb, uint32_t th, int rt, int32_t method, int min_waves = 12);
size_t fused_match_lds(int kb, uint32_t tw, uint32_t th, int rt, int32_t method, int tc);
// Sets tiles_x and tile0 of n consecutive jobs for rt waves per workgroup;
// returns the number of tiles.
long long match_jobs_layout(MatchJob *jobs, int n, int rt);
// One launch over n jobs of one (kb, method) group, device table `jobs`.
iph_status launch_match_jobs(const MatchJob *jobs, int n, long long ntiles, int kb,
                             int32_t method, int rt, int tc, size_t lds, hipStream_t s);

// ---- launchers implemented in match.hip ---------------------------------

size_t match_workspace(int32_t p_type, uint32_t w, uint32_t h, uint32_t tw, uint32_t th,
                       int32_t method, int32_t path);
iph_status launch_match(const void *img, int32_t p_type, uint32_t w, uint32_t h,
                        const void *tpl, uint32_t tw, uint32_t th, int32_t method,
                        int32_t path, float *out, void *ws, size_t ws_bytes, hipStream_t s);
// path: IPH_PATH_DIRECT (the reference's order, bit-identical) or
// IPH_PATH_MFMA (CrossCorrelation and CrossCorrelationNormalized, templates
// up to 128 px: BF16x3 correlations); resolved by the caller.
size_t masked_match_workspace(uint32_t w, uint32_t h, uint32_t tw, uint32_t th, int32_t method,
                              int32_t path);
iph_status launch_masked_match(const uint8_t *img, uint32_t w, uint32_t h, const uint8_t *tpl,
                               uint32_t tw, uint32_t th, const uint8_t *mask, int32_t method,
                               int32_t path, float *out, void *ws, size_t ws_bytes,
                               hipStream_t s);
size_t extremes_workspace();
iph_status launch_extremes(const void *img, int32_t p_type, uint32_t w, uint32_t h,
                           iph_extremes *out, void *ws, size_t ws_bytes, hipStream_t s);
// find_extremes with its key pairs in the library's own fixed device array
// (no allocation); calls that overlap on one device race.
iph_status launch_extremes_noalloc(const void *img, int32_t p_type, uint32_t w, uint32_t h,
                                   iph_extremes *out, hipStream_t s);

}  // namespace iph
