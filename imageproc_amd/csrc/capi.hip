// C ABI entry points (include/imageproc_hip.h): argument checks with the
// reference's panic messages, device memory staging for the host-pointer
// functions, and dispatch to the gfx950 launchers.
#include "iph_internal.hpp"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace iph {

static thread_local std::string g_last_error;

void set_error(const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

void clear_error() { g_last_error.clear(); }


// Fails unless a gfx950 device is current: the library ships gfx950 code only.
iph_status require_device() {
    static thread_local int checked_device = -1;
    int dev = 0;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) {
        set_error("no HIP device available (%s)", hipGetErrorString(e));
        return IPH_EDEVICE;
    }
    IPH_TRY_HIP(hipGetDevice(&dev));
    if (checked_device == dev) return IPH_OK;
    hipDeviceProp_t prop;
    IPH_TRY_HIP(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        set_error("device %d is %s; this library is built for gfx950 only", dev,
                  prop.gcnArchName);
        return IPH_EDEVICE;
    }
    checked_device = dev;
    return IPH_OK;
}

// gaussian_pdf / gaussian_kernel_f32, filter/mod.rs:125-145, in binary32 with
// glibc expf (what f32::exp calls on x86_64-unknown-linux-gnu).
float gaussian_pdf(float x, float sigma) {
    const float pi_f32 = 3.14159274101257324219f;  // f32::consts::PI
    float two_pi = 2.0f * pi_f32;
    float norm = sigma * std::sqrt(two_pi);
    float recip = 1.0f / norm;
    float x2 = x * x;
    float den = 2.0f * (sigma * sigma);
    float arg = -x2 / den;
    return recip * ::expf(arg);
}

constexpr uint32_t kMaxTaps = 1u << 24;

int32_t gaussian_kernel(float sigma, std::vector<float> &k) {
    float r_f = std::ceil(2.0f * sigma);  // (2.0 * sigma).ceil() as usize
    if (!(r_f >= 0.0f)) r_f = 0.0f;
    if (r_f > (float)((kMaxTaps - 1) / 2)) return -1;
    int radius = (int)r_f;
    k.assign(2 * (size_t)radius + 1, 0.0f);
    for (int i = 0; i <= radius; i++) {
        float v = gaussian_pdf((float)i, sigma);
        k[radius + i] = v;
        k[radius - i] = v;
    }
    float sum = -0.0f;  // Iterator::sum for f32: sequential fold from -0.0
    for (float v : k) sum = sum + v;
    for (float &v : k) v = v / sum;
    return (int32_t)k.size();
}

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    iph_status alloc(size_t n) {
        if (n == 0) n = 1;
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            p = nullptr;
            set_error("hipMalloc(%zu) failed: %s", n, hipGetErrorString(e));
            return IPH_ENOMEM;
        }
        return IPH_OK;
    }
};

iph_status upload(DevBuf &b, const void *src, size_t n) {
    IPH_TRY(b.alloc(n));
    if (n) IPH_TRY_HIP(hipMemcpy(b.p, src, n, hipMemcpyHostToDevice));
    return IPH_OK;
}

iph_status finish(DevBuf &d, void *dst, size_t n) {
    IPH_TRY_HIP(hipDeviceSynchronize());
    if (n) IPH_TRY_HIP(hipMemcpy(dst, d.p, n, hipMemcpyDeviceToHost));
    return IPH_OK;
}

bool valid_pk(int32_t p, int32_t k) {
    return (p == IPH_U8 && (k == IPH_F32 || k == IPH_I32)) || (p == IPH_F32 && k == IPH_F32);
}

iph_status check_pk(int32_t p, int32_t k, const char *what) {
    if (!valid_pk(p, k)) {
        set_error("%s: pixel type %d with kernel type %d does not satisfy the reference's "
                  "Into<K> + Clamp<K> bounds",
                  what, p, k);
        return IPH_EUNSUPPORTED;
    }
    return IPH_OK;
}

// filter_clamped needs K: From<P::Subpixel> and S: Clamp<K>
// (filter/mod.rs:202-208, definitions.rs:151-190).
iph_status check_pks(int32_t p, int32_t k, int32_t s) {
    bool ok = false;
    if (k == IPH_F32)
        ok = (p == IPH_U8 || p == IPH_F32) && (s == IPH_U8 || s == IPH_U16 || s == IPH_F32);
    else if (k == IPH_I32)
        ok = p == IPH_U8 && (s == IPH_U8 || s == IPH_U16 || s == IPH_I16 || s == IPH_I32);
    if (!ok) {
        set_error("filter_clamped: (pixel %d, kernel %d, output %d) is not a combination the "
                  "reference accepts",
                  p, k, s);
        return IPH_EUNSUPPORTED;
    }
    return IPH_OK;
}

iph_status check_dims(uint32_t w, uint32_t h) {
    if ((uint64_t)w * h >= (1ull << 32)) {
        set_error("image of %ux%u pixels exceeds 2^32 - 1 pixels", w, h);
        return IPH_EINVAL;
    }
    return IPH_OK;
}

iph_status check_match(int32_t p, uint32_t w, uint32_t h, uint32_t tw, uint32_t th,
                       int32_t method, int32_t path) {
    if (p != IPH_U8 && p != IPH_F32) {
        set_error("match_template takes GrayImage (u8) or Luma<f32> images, not type %d", p);
        return IPH_EUNSUPPORTED;
    }
    // ImageTemplate::new, template_matching.rs:395-402
    if (w < tw) {
        set_error("image width must be greater than or equal to template width");
        return IPH_EINVAL;
    }
    if (h < th) {
        set_error("image height must be greater than or equal to template height");
        return IPH_EINVAL;
    }
    if (method < IPH_SSE || method > IPH_CCORR_NORMED) {
        set_error("unknown MatchTemplateMethod %d", method);
        return IPH_EINVAL;
    }
    if (path < IPH_PATH_AUTO || path > IPH_PATH_DIRECT) {
        set_error("unknown correlation path %d", path);
        return IPH_EINVAL;
    }
    return check_dims(w, h);
}

// Empty templates have no window to correlate: the direct path handles them.
// AUTO takes the MFMA engines where they apply (template sides up to
// kMfmaMaxSide, images within mfma_match_fits) and the direct kernel
// otherwise, so every template the reference accepts
// (template_matching.rs:394-404) has a GPU path.
int32_t resolve_path(int32_t path, uint32_t w, uint32_t h, uint32_t tw, uint32_t th) {
    if (tw == 0 || th == 0) return IPH_PATH_DIRECT;
    if (path != IPH_PATH_AUTO) return path;
    return (tw <= kMfmaMaxSide && th <= kMfmaMaxSide && mfma_match_fits(w, h))
               ? IPH_PATH_MFMA
               : IPH_PATH_DIRECT;
}

}  // namespace iph

using namespace iph;

extern "C" {

IPH_API const char *iph_last_error(void) { return g_last_error.c_str(); }

IPH_API iph_status iph_device_info(char *buf, size_t cap) {
    IPH_TRY(require_device());
    int dev = 0;
    IPH_TRY_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    IPH_TRY_HIP(hipGetDeviceProperties(&prop, dev));
    snprintf(buf, cap, "%s|%s|cus=%d|clock_khz=%d|mem=%zu", prop.name, prop.gcnArchName,
             prop.multiProcessorCount, prop.clockRate, (size_t)prop.totalGlobalMem);
    return IPH_OK;
}

IPH_API int32_t iph_gaussian_kernel_f32(float sigma, float *out, int32_t cap) {
    std::vector<float> k;
    int32_t n = gaussian_kernel(sigma, k);
    if (n < 0) return -1;
    int32_t m = n < cap ? n : cap;
    if (out && m > 0) std::memcpy(out, k.data(), sizeof(float) * (size_t)m);
    return n;
}

// ---- host-pointer API ----------------------------------------------------------

IPH_API iph_status iph_separable_filter(const void *src, void *dst, uint32_t w, uint32_t h,
                                        int32_t p, const void *hk, uint32_t hlen,
                                        const void *vk, uint32_t vlen, int32_t k) {
    clear_error();
    if (hlen != vlen) {  // filter/mod.rs:174-180
        set_error("the two 1D kernels must be the same length");
        return IPH_EINVAL;
    }
    IPH_TRY(check_pk(p, k, "separable_filter"));
    IPH_TRY(check_dims(w, h));
    if (w == 0 || h == 0) return IPH_OK;
    IPH_TRY(require_device());
    size_t bytes = (size_t)w * h * elem_size(p);
    DevBuf din, dout, dws;
    IPH_TRY(upload(din, src, bytes));
    IPH_TRY(dout.alloc(bytes));
    size_t wsb = separable_workspace(w, h, 1, p, hlen, k);
    IPH_TRY(dws.alloc(wsb));
    IPH_TRY(launch_separable(din.p, dout.p, w, h, 1, (size_t)w * h, p, hk, vk, hlen, k, dws.p,
                             wsb, nullptr));
    return finish(dout, dst, bytes);
}

IPH_API iph_status iph_gaussian_blur_f32(const void *src, void *dst, uint32_t w, uint32_t h,
                                         int32_t p, float sigma) {
    clear_error();
    if (!(sigma > 0.0f)) {  // filter/mod.rs:161
        set_error("sigma must be > 0.0");
        return IPH_EINVAL;
    }
    if (p != IPH_U8 && p != IPH_F32) {
        set_error("gaussian_blur_f32 takes Luma<u8> or Luma<f32>, not type %d", p);
        return IPH_EUNSUPPORTED;
    }
    std::vector<float> k;
    if (gaussian_kernel(sigma, k) < 0) {
        set_error("sigma %g gives more than %u taps", (double)sigma, kMaxTaps);
        return IPH_EINVAL;
    }
    uint32_t n = (uint32_t)k.size();
    return iph_separable_filter(src, dst, w, h, p, k.data(), n, k.data(), n, IPH_F32);
}

namespace {
iph_status check_integral(uint32_t w, uint32_t h) {
    if ((uint64_t)(w + 1ull) * (h + 1ull) >= (1ull << 32)) {
        set_error("integral image of %ux%u pixels is too large", w, h);
        return IPH_EINVAL;
    }
    return IPH_OK;
}

iph_status check_adaptive(uint32_t w, uint32_t h, uint32_t block_radius) {
    if (block_radius == 0) {  // contrast.rs:20
        set_error("assertion failed: block_radius > 0");
        return IPH_EINVAL;
    }
    if (block_radius > 0x7fffffffu) {  // the reference's `block_radius as i32` turns negative
        set_error("block_radius %u is above 2^31 - 1", block_radius);
        return IPH_EINVAL;
    }
    return check_integral(w, h);
}
}  // namespace

IPH_API iph_status iph_integral_image(const uint8_t *src, uint32_t w, uint32_t h, int32_t squared,
                                      uint32_t *out) {
    clear_error();
    IPH_TRY(check_integral(w, h));
    IPH_TRY(require_device());
    const size_t n = (size_t)w * h, obytes = ((size_t)w + 1) * ((size_t)h + 1) * sizeof(uint32_t);
    const size_t wsb = integral_workspace(w, h);
    DevBuf din, dout, dws;
    if (n) IPH_TRY(upload(din, src, n));
    IPH_TRY(dout.alloc(obytes));
    IPH_TRY(dws.alloc(wsb));
    IPH_TRY(launch_integral_image(static_cast<const uint8_t *>(din.p), w, h, squared ? 1 : 0,
                                  static_cast<uint32_t *>(dout.p), dws.p, wsb, nullptr));
    return finish(dout, out, obytes);
}

IPH_API iph_status iph_adaptive_threshold(const uint8_t *src, uint8_t *dst, uint32_t w, uint32_t h,
                                          uint32_t block_radius, int32_t delta) {
    clear_error();
    IPH_TRY(check_adaptive(w, h, block_radius));
    const size_t n = (size_t)w * h;
    if (n == 0) return IPH_OK;
    IPH_TRY(require_device());
    const size_t wsb = adaptive_threshold_workspace(w, h, block_radius);
    DevBuf din, dout, dws;
    IPH_TRY(upload(din, src, n));
    IPH_TRY(dout.alloc(n));
    IPH_TRY(dws.alloc(wsb));
    IPH_TRY(launch_adaptive_threshold(static_cast<const uint8_t *>(din.p), w, h, block_radius,
                                      delta, static_cast<uint8_t *>(dout.p), dws.p, wsb, nullptr));
    return finish(dout, dst, n);
}

IPH_API size_t iph_dev_integral_image_workspace(uint32_t w, uint32_t h) {
    return integral_workspace(w, h);
}

IPH_API iph_status iph_dev_integral_image(const uint8_t *src, uint32_t w, uint32_t h,
                                          int32_t squared, uint32_t *out, void *ws, size_t wsb,
                                          void *stream) {
    clear_error();
    IPH_TRY(check_integral(w, h));
    const bool empty = (size_t)w * h == 0;
    if (out == nullptr || (!empty && (src == nullptr || ws == nullptr))) {
        set_error("null image, table or workspace");
        return IPH_EINVAL;
    }
    IPH_TRY(require_device());
    return launch_integral_image(src, w, h, squared ? 1 : 0, out, ws, wsb,
                                 static_cast<hipStream_t>(stream));
}

IPH_API size_t iph_dev_adaptive_threshold_workspace(uint32_t w, uint32_t h,
                                                    uint32_t block_radius) {
    return adaptive_threshold_workspace(w, h, block_radius);
}

IPH_API iph_status iph_dev_adaptive_threshold(const uint8_t *src, uint8_t *dst, uint32_t w,
                                              uint32_t h, uint32_t block_radius, int32_t delta,
                                              void *ws, size_t wsb, void *stream) {
    clear_error();
    IPH_TRY(check_adaptive(w, h, block_radius));
    if ((size_t)w * h == 0) return IPH_OK;
    if (src == nullptr || dst == nullptr ||
        (ws == nullptr && adaptive_threshold_workspace(w, h, block_radius) > 0)) {
        set_error("null image, output or workspace");
        return IPH_EINVAL;
    }
    IPH_TRY(require_device());
    return launch_adaptive_threshold(src, w, h, block_radius, delta, dst, ws, wsb,
                                     static_cast<hipStream_t>(stream));
}

namespace {
iph_status check_canny(uint32_t w, uint32_t h, float low, float high) {
    if (!(high >= low)) {  // edges.rs:31
        set_error("assertion failed: high_threshold >= low_threshold");
        return IPH_EINVAL;
    }
    if (w == 0 || h == 0) {  // edges.rs:62, 119: height - 1 / width - 1 on an empty image
        set_error("attempt to subtract with overflow");
        return IPH_EINVAL;
    }
    return IPH_OK;
}
}  // namespace

IPH_API iph_status iph_canny(const uint8_t *src, uint8_t *dst, uint32_t w, uint32_t h, float low,
                             float high) {
    clear_error();
    IPH_TRY(check_canny(w, h, low, high));
    IPH_TRY(require_device());
    const size_t n = (size_t)w * h, wsb = canny_workspace(w, h);
    DevBuf din, dout, dws;
    IPH_TRY(upload(din, src, n));
    IPH_TRY(dout.alloc(n));
    IPH_TRY(dws.alloc(wsb));
    IPH_TRY(launch_canny(static_cast<const uint8_t *>(din.p), static_cast<uint8_t *>(dout.p), w, h,
                         low, high, dws.p, wsb, nullptr));
    return finish(dout, dst, n);
}

IPH_API size_t iph_dev_canny_workspace(uint32_t w, uint32_t h) { return canny_workspace(w, h); }

IPH_API iph_status iph_dev_canny(const uint8_t *src, uint8_t *dst, uint32_t w, uint32_t h,
                                 float low, float high, void *ws, size_t wsb, void *stream) {
    clear_error();
    IPH_TRY(check_canny(w, h, low, high));
    if (src == nullptr || dst == nullptr || ws == nullptr) {
        set_error("null image, output or workspace");
        return IPH_EINVAL;
    }
    IPH_TRY(require_device());
    return launch_canny(src, dst, w, h, low, high, ws, wsb, static_cast<hipStream_t>(stream));
}

IPH_API iph_status iph_gradients(const uint8_t *src, uint32_t w, uint32_t h, const int32_t *hk,
                                 uint32_t hkw, uint32_t hkh, const int32_t *vk, uint32_t vkw,
                                 uint32_t vkh, uint16_t *out) {
    clear_error();
    if (hkw == 0 || hkh == 0 || vkw == 0 || vkh == 0) {  // Kernel::new, kernel.rs:24-35
        set_error("width and height must be non-zero");
        return IPH_EINVAL;
    }
    IPH_TRY(check_dims(w, h));
    const size_t n = (size_t)w * h;
    if (n == 0) return IPH_OK;
    IPH_TRY(require_device());
    DevBuf din, dout;
    IPH_TRY(upload(din, src, n));
    IPH_TRY(dout.alloc(n * sizeof(uint16_t)));
    IPH_TRY(launch_gradients(static_cast<const uint8_t *>(din.p), w, h, hk, hkw, hkh, vk, vkw,
                             vkh, static_cast<uint16_t *>(dout.p), nullptr));
    return finish(dout, out, n * sizeof(uint16_t));
}

IPH_API iph_status iph_sharpen3x3(const uint8_t *src, uint8_t *dst, uint32_t w, uint32_t h) {
    // identity minus the 3x3 Laplacian, filter/sharpen.rs:14
    static const int32_t k[9] = {0, -1, 0, -1, 5, -1, 0, -1, 0};
    return iph_filter_clamped(src, IPH_U8, w, h, k, IPH_I32, 3, 3, dst, IPH_U8);
}

IPH_API iph_status iph_sharpen_gaussian(const uint8_t *src, uint8_t *dst, uint32_t w, uint32_t h,
                                        float sigma, float amount) {
    clear_error();
    if (!(sigma > 0.0f)) {  // gaussian_blur_f32's assert, filter/mod.rs:161
        set_error("sigma must be > 0.0");
        return IPH_EINVAL;
    }
    std::vector<float> k;
    if (gaussian_kernel(sigma, k) < 0) {
        set_error("sigma %g gives more than %u taps", (double)sigma, kMaxTaps);
        return IPH_EINVAL;
    }
    IPH_TRY(check_dims(w, h));
    const size_t bytes = (size_t)w * h;
    if (bytes == 0) return IPH_OK;
    IPH_TRY(require_device());
    DevBuf din, dout, dws;
    IPH_TRY(upload(din, src, bytes));
    IPH_TRY(dout.alloc(bytes));
    const uint32_t n = (uint32_t)k.size();
    const size_t wsb = sharpen_gaussian_workspace(w, h, n);
    IPH_TRY(dws.alloc(wsb));
    IPH_TRY(launch_sharpen_gaussian(static_cast<const uint8_t *>(din.p),
                                    static_cast<uint8_t *>(dout.p), w, h, k.data(), n, amount,
                                    dws.p, wsb, nullptr));
    return finish(dout, dst, bytes);
}

IPH_API iph_status iph_dev_gradients(const uint8_t *src, uint32_t w, uint32_t h, const int32_t *hk,
                                     uint32_t hkw, uint32_t hkh, const int32_t *vk, uint32_t vkw,
                                     uint32_t vkh, uint16_t *out, void *stream) {
    clear_error();
    if (hkw == 0 || hkh == 0 || vkw == 0 || vkh == 0) {  // Kernel::new, kernel.rs:24-35
        set_error("width and height must be non-zero");
        return IPH_EINVAL;
    }
    IPH_TRY(check_dims(w, h));
    if ((size_t)w * h == 0) return IPH_OK;
    if (src == nullptr || out == nullptr || hk == nullptr || vk == nullptr) {
        set_error("null image, output or kernel");
        return IPH_EINVAL;
    }
    IPH_TRY(require_device());
    return launch_gradients(src, w, h, hk, hkw, hkh, vk, vkw, vkh, out,
                            static_cast<hipStream_t>(stream));
}

IPH_API size_t iph_dev_sharpen_gaussian_workspace(uint32_t w, uint32_t h, float sigma) {
    std::vector<float> k;
    if (!(sigma > 0.0f) || gaussian_kernel(sigma, k) < 0) return 0;
    return sharpen_gaussian_workspace(w, h, (uint32_t)k.size());
}

IPH_API iph_status iph_dev_sharpen_gaussian(const uint8_t *src, uint8_t *dst, uint32_t w,
                                            uint32_t h, float sigma, float amount, void *ws,
                                            size_t wsb, void *stream) {
    clear_error();
    if (!(sigma > 0.0f)) {  // gaussian_blur_f32's assert, filter/mod.rs:161
        set_error("sigma must be > 0.0");
        return IPH_EINVAL;
    }
    std::vector<float> k;
    if (gaussian_kernel(sigma, k) < 0) {
        set_error("sigma %g gives more than %u taps", (double)sigma, kMaxTaps);
        return IPH_EINVAL;
    }
    IPH_TRY(check_dims(w, h));
    if ((size_t)w * h == 0) return IPH_OK;
    if (src == nullptr || dst == nullptr ||
        (ws == nullptr && sharpen_gaussian_workspace(w, h, (uint32_t)k.size()) > 0)) {
        set_error("null image, output or workspace");
        return IPH_EINVAL;
    }
    IPH_TRY(require_device());
    return launch_sharpen_gaussian(src, dst, w, h, k.data(), (uint32_t)k.size(), amount, ws, wsb,
                                   static_cast<hipStream_t>(stream));
}

static iph_status filter1d_host(const void *src, void *dst, uint32_t w, uint32_t h, int32_t p,
                                const void *k, uint32_t len, int32_t kt, bool vertical) {
    clear_error();
    IPH_TRY(check_pk(p, kt, vertical ? "vertical_filter" : "horizontal_filter"));
    IPH_TRY(check_dims(w, h));
    if (w == 0 || h == 0) return IPH_OK;
    IPH_TRY(require_device());
    size_t bytes = (size_t)w * h * elem_size(p);
    DevBuf din, dout, dws;
    IPH_TRY(upload(din, src, bytes));
    IPH_TRY(dout.alloc(bytes));
    size_t wsb = align_up((size_t)len * 4, 256);
    IPH_TRY(dws.alloc(wsb));
    IPH_TRY(launch_filter1d(din.p, dout.p, w, h, p, k, len, kt, vertical, dws.p, wsb, nullptr));
    return finish(dout, dst, bytes);
}

IPH_API iph_status iph_horizontal_filter(const void *src, void *dst, uint32_t w, uint32_t h,
                                         int32_t p, const void *k, uint32_t len, int32_t kt) {
    return filter1d_host(src, dst, w, h, p, k, len, kt, false);
}

IPH_API iph_status iph_vertical_filter(const void *src, void *dst, uint32_t w, uint32_t h,
                                       int32_t p, const void *k, uint32_t len, int32_t kt) {
    return filter1d_host(src, dst, w, h, p, k, len, kt, true);
}

IPH_API iph_status iph_filter_clamped(const void *src, int32_t p, uint32_t w, uint32_t h,
                                      const void *kernel, int32_t kt, uint32_t kw, uint32_t kh,
                                      void *dst, int32_t st) {
    clear_error();
    if (kw == 0 || kh == 0) {  // Kernel::new, kernel.rs:25
        set_error("width and height must be non-zero");
        return IPH_EINVAL;
    }
    IPH_TRY(check_pks(p, kt, st));
    IPH_TRY(check_dims(w, h));
    if (w == 0 || h == 0) return IPH_OK;
    IPH_TRY(require_device());
    DevBuf din, dout, dws;
    IPH_TRY(upload(din, src, (size_t)w * h * elem_size(p)));
    size_t obytes = (size_t)w * h * elem_size(st);
    IPH_TRY(dout.alloc(obytes));
    size_t wsb = filter_clamped_workspace(kt, kw, kh);
    IPH_TRY(dws.alloc(wsb));
    IPH_TRY(launch_filter_clamped(din.p, p, w, h, kernel, kt, kw, kh, dout.p, st, dws.p, wsb,
                                  nullptr));
    return finish(dout, dst, obytes);
}

IPH_API iph_status iph_box_filter(const uint8_t *src, uint8_t *dst, uint32_t w, uint32_t h,
                                  uint32_t rx, uint32_t ry) {
    clear_error();
    if (rx >= kMaxTaps || ry >= kMaxTaps) {
        set_error("box_filter radius above %u is not supported", kMaxTaps - 1);
        return IPH_EINVAL;
    }
    IPH_TRY(check_dims(w, h));
    if (w == 0 || h == 0) return IPH_OK;
    IPH_TRY(require_device());
    size_t bytes = (size_t)w * h;
    DevBuf din, dout, dws;
    IPH_TRY(upload(din, src, bytes));
    IPH_TRY(dout.alloc(bytes));
    size_t wsb = align_up(bytes, 256);
    IPH_TRY(dws.alloc(wsb));
    IPH_TRY(launch_box_filter(static_cast<uint8_t *>(din.p), static_cast<uint8_t *>(dout.p), w,
                              h, rx, ry, dws.p, wsb, nullptr));
    return finish(dout, dst, bytes);
}

IPH_API iph_status iph_match_template(const void *image, int32_t p, uint32_t w, uint32_t h,
                                      const void *tmpl, uint32_t tw, uint32_t th,
                                      int32_t method, int32_t path, float *out) {
    clear_error();
    IPH_TRY(check_match(p, w, h, tw, th, method, path));
    path = resolve_path(path, w, h, tw, th);
    uint32_t ow = w - tw + 1, oh = h - th + 1;
    size_t obytes = (size_t)ow * oh * sizeof(float);
    if (obytes == 0) return IPH_OK;
    IPH_TRY(require_device());
    DevBuf di, dt, dout, dws;
    IPH_TRY(upload(di, image, (size_t)w * h * elem_size(p)));
    IPH_TRY(upload(dt, tmpl, (size_t)tw * th * elem_size(p)));
    IPH_TRY(dout.alloc(obytes));
    size_t wsb = match_workspace(p, w, h, tw, th, method, path);
    IPH_TRY(dws.alloc(wsb));
    IPH_TRY(launch_match(di.p, p, w, h, dt.p, tw, th, method, path,
                         static_cast<float *>(dout.p), dws.p, wsb, nullptr));
    return finish(dout, out, obytes);
}

// ImageTemplateMask::new, template_matching.rs:430-440: the mask check comes
// before the template-vs-image checks.
static iph_status check_masked(uint32_t w, uint32_t h, uint32_t tw, uint32_t th, uint32_t mw,
                               uint32_t mh, int32_t method, int32_t path) {
    if (mw != tw || mh != th) {
        set_error("assertion `left == right` failed: the template and mask must be the same "
                  "size\n  left: (%u, %u)\n right: (%u, %u)", tw, th, mw, mh);
        return IPH_EINVAL;
    }
    return check_match(IPH_U8, w, h, tw, th, method, path);
}

// The default masked path: the MFMA correlations for CrossCorrelation and
// CrossCorrelationNormalized with templates up to 128 px, else the direct
// kernel in the reference's order.
static int32_t resolve_masked_path(int32_t path, uint32_t w, uint32_t h, uint32_t tw,
                                   uint32_t th, int32_t method) {
    if (path != IPH_PATH_AUTO) return path;
    if (tw == 0 || th == 0 || (method != IPH_CCORR && method != IPH_CCORR_NORMED))
        return IPH_PATH_DIRECT;
    return (tw <= kMfmaMaxSide && th <= kMfmaMaxSide && mfma_match_fits(w, h)) ? IPH_PATH_MFMA
                                                                               : IPH_PATH_DIRECT;
}

IPH_API iph_status iph_match_template_with_mask(const uint8_t *image, uint32_t w, uint32_t h,
                                                const uint8_t *tmpl, uint32_t tw, uint32_t th,
                                                const uint8_t *mask, uint32_t mw, uint32_t mh,
                                                int32_t method, int32_t path, float *out) {
    clear_error();
    IPH_TRY(check_masked(w, h, tw, th, mw, mh, method, path));
    path = resolve_masked_path(path, w, h, tw, th, method);
    const uint32_t ow = w - tw + 1, oh = h - th + 1;
    const size_t obytes = (size_t)ow * oh * sizeof(float);
    if (obytes == 0) return IPH_OK;
    IPH_TRY(require_device());
    DevBuf di, dt, dm, dout, dws;
    IPH_TRY(upload(di, image, (size_t)w * h));
    IPH_TRY(upload(dt, tmpl, (size_t)tw * th));
    IPH_TRY(upload(dm, mask, (size_t)tw * th));
    IPH_TRY(dout.alloc(obytes));
    const size_t wsb = masked_match_workspace(w, h, tw, th, method, path);
    IPH_TRY(dws.alloc(wsb));
    IPH_TRY(launch_masked_match(static_cast<const uint8_t *>(di.p), w, h,
                                static_cast<const uint8_t *>(dt.p), tw, th,
                                static_cast<const uint8_t *>(dm.p), method, path,
                                static_cast<float *>(dout.p), dws.p, wsb, nullptr));
    return finish(dout, out, obytes);
}

IPH_API iph_status iph_find_extremes(const void *image, int32_t p, uint32_t w, uint32_t h,
                                     iph_extremes *out) {
    clear_error();
    if (w == 0 || h == 0) {  // template_matching.rs:483-486
        set_error("image must be non-empty");
        return IPH_EINVAL;
    }
    if (p != IPH_U8 && p != IPH_F32) {
        set_error("find_extremes is built for Luma<u8> and Luma<f32>, not type %d", p);
        return IPH_EUNSUPPORTED;
    }
    IPH_TRY(check_dims(w, h));
    IPH_TRY(require_device());
    DevBuf di, dout, dws;
    IPH_TRY(upload(di, image, (size_t)w * h * elem_size(p)));
    IPH_TRY(dout.alloc(sizeof(iph_extremes)));
    IPH_TRY(dws.alloc(extremes_workspace()));
    IPH_TRY(launch_extremes(di.p, p, w, h, static_cast<iph_extremes *>(dout.p), dws.p,
                            extremes_workspace(), nullptr));
    return finish(dout, out, sizeof(iph_extremes));
}

// ---- device-pointer API --------------------------------------------------------

IPH_API size_t iph_dev_gaussian_blur_workspace(uint32_t w, uint32_t h, uint32_t n, int32_t p,
                                               float sigma) {
    std::vector<float> k;
    if (!(sigma > 0.0f) || gaussian_kernel(sigma, k) < 0) return 0;
    return separable_workspace(w, h, n, p, (uint32_t)k.size(), IPH_F32);
}

IPH_API iph_status iph_dev_separable_filter(const void *src, void *dst, uint32_t w, uint32_t h,
                                            uint32_t n, size_t stride, int32_t p,
                                            const void *hk, uint32_t hlen, const void *vk,
                                            uint32_t vlen, int32_t k, void *ws, size_t wsb,
                                            void *stream) {
    clear_error();
    if (hlen != vlen) {
        set_error("the two 1D kernels must be the same length");
        return IPH_EINVAL;
    }
    IPH_TRY(check_pk(p, k, "separable_filter"));
    IPH_TRY(check_dims(w, h));
    if ((uint64_t)w * h > stride && n > 1) {
        set_error("image_stride %zu is smaller than one image (%ux%u)", stride, w, h);
        return IPH_EINVAL;
    }
    IPH_TRY(require_device());
    return launch_separable(src, dst, w, h, n, stride, p, hk, vk, hlen, k, ws, wsb,
                            static_cast<hipStream_t>(stream));
}

IPH_API size_t iph_dev_separable_filter_workspace(uint32_t w, uint32_t h, uint32_t n,
                                                  int32_t p, uint32_t len, int32_t k) {
    return separable_workspace(w, h, n, p, len, k);
}

IPH_API iph_status iph_dev_gaussian_blur_f32(const void *src, void *dst, uint32_t w, uint32_t h,
                                             uint32_t n, size_t stride, int32_t p, float sigma,
                                             void *ws, size_t wsb, void *stream) {
    clear_error();
    if (!(sigma > 0.0f)) {
        set_error("sigma must be > 0.0");
        return IPH_EINVAL;
    }
    if (p != IPH_U8 && p != IPH_F32) {
        set_error("gaussian_blur_f32 takes Luma<u8> or Luma<f32>, not type %d", p);
        return IPH_EUNSUPPORTED;
    }
    std::vector<float> k;
    if (gaussian_kernel(sigma, k) < 0) {
        set_error("sigma %g gives more than %u taps", (double)sigma, kMaxTaps);
        return IPH_EINVAL;
    }
    uint32_t len = (uint32_t)k.size();
    return iph_dev_separable_filter(src, dst, w, h, n, stride, p, k.data(), len, k.data(), len,
                                    IPH_F32, ws, wsb, stream);
}

IPH_API size_t iph_dev_filter_clamped_workspace(int32_t k, uint32_t kw, uint32_t kh) {
    return filter_clamped_workspace(k, kw, kh);
}

IPH_API iph_status iph_dev_filter_clamped(const void *src, int32_t p, uint32_t w, uint32_t h,
                                          const void *kernel, int32_t kt, uint32_t kw,
                                          uint32_t kh, void *dst, int32_t st, void *ws,
                                          size_t wsb, void *stream) {
    clear_error();
    if (kw == 0 || kh == 0) {
        set_error("width and height must be non-zero");
        return IPH_EINVAL;
    }
    IPH_TRY(check_pks(p, kt, st));
    IPH_TRY(check_dims(w, h));
    IPH_TRY(require_device());
    return launch_filter_clamped(src, p, w, h, kernel, kt, kw, kh, dst, st, ws, wsb,
                                 static_cast<hipStream_t>(stream));
}

IPH_API size_t iph_dev_box_filter_workspace(uint32_t w, uint32_t h) {
    return align_up((size_t)w * h, 256);
}

IPH_API iph_status iph_dev_box_filter(const uint8_t *src, uint8_t *dst, uint32_t w, uint32_t h,
                                      uint32_t rx, uint32_t ry, void *ws, size_t wsb,
                                      void *stream) {
    clear_error();
    if (rx >= kMaxTaps || ry >= kMaxTaps) {
        set_error("box_filter radius above %u is not supported", kMaxTaps - 1);
        return IPH_EINVAL;
    }
    IPH_TRY(check_dims(w, h));
    IPH_TRY(require_device());
    return launch_box_filter(src, dst, w, h, rx, ry, ws, wsb, static_cast<hipStream_t>(stream));
}

IPH_API size_t iph_dev_match_template_with_mask_workspace(uint32_t w, uint32_t h, uint32_t tw,
                                                          uint32_t th, int32_t method,
                                                          int32_t path) {
    return masked_match_workspace(w, h, tw, th, method,
                                  resolve_masked_path(path, w, h, tw, th, method));
}

IPH_API iph_status iph_dev_match_template_with_mask(const uint8_t *image, uint32_t w, uint32_t h,
                                                    const uint8_t *tmpl, uint32_t tw,
                                                    uint32_t th, const uint8_t *mask,
                                                    uint32_t mw, uint32_t mh, int32_t method,
                                                    int32_t path, float *out, void *ws,
                                                    size_t wsb, void *stream) {
    clear_error();
    IPH_TRY(check_masked(w, h, tw, th, mw, mh, method, path));
    IPH_TRY(require_device());
    return launch_masked_match(image, w, h, tmpl, tw, th, mask, method,
                               resolve_masked_path(path, w, h, tw, th, method), out, ws, wsb,
                               static_cast<hipStream_t>(stream));
}

IPH_API size_t iph_dev_match_template_workspace(int32_t p, uint32_t w, uint32_t h, uint32_t tw,
                                                uint32_t th, int32_t method, int32_t path) {
    if (w < tw || h < th) return 0;
    return match_workspace(p, w, h, tw, th, method, resolve_path(path, w, h, tw, th));
}

IPH_API iph_status iph_dev_match_template(const void *image, int32_t p, uint32_t w, uint32_t h,
                                          const void *tmpl, uint32_t tw, uint32_t th,
                                          int32_t method, int32_t path, float *out, void *ws,
                                          size_t wsb, void *stream) {
    clear_error();
    IPH_TRY(check_match(p, w, h, tw, th, method, path));
    IPH_TRY(require_device());
    return launch_match(image, p, w, h, tmpl, tw, th, method, resolve_path(path, w, h, tw, th), out,
                        ws, wsb, static_cast<hipStream_t>(stream));
}

IPH_API size_t iph_dev_find_extremes_workspace(void) { return extremes_workspace(); }

IPH_API iph_status iph_dev_find_extremes(const void *image, int32_t p, uint32_t w, uint32_t h,
                                         iph_extremes *out, void *ws, size_t wsb,
                                         void *stream) {
    clear_error();
    if (w == 0 || h == 0) {
        set_error("image must be non-empty");
        return IPH_EINVAL;
    }
    if (p != IPH_U8 && p != IPH_F32) {
        set_error("find_extremes is built for Luma<u8> and Luma<f32>, not type %d", p);
        return IPH_EUNSUPPORTED;
    }
    IPH_TRY(check_dims(w, h));
    IPH_TRY(require_device());
    return launch_extremes(image, p, w, h, out, ws, wsb, static_cast<hipStream_t>(stream));
}

IPH_API iph_status iph_dev_find_extremes_noalloc(const void *image, int32_t p, uint32_t w,
                                                 uint32_t h, iph_extremes *out, void *stream) {
    clear_error();
    if (w == 0 || h == 0) {
        set_error("image must be non-empty");
        return IPH_EINVAL;
    }
    if (p != IPH_U8 && p != IPH_F32) {
        set_error("find_extremes is built for Luma<u8> and Luma<f32>, not type %d", p);
        return IPH_EUNSUPPORTED;
    }
    if (image == nullptr || out == nullptr) {
        set_error("null image or result");
        return IPH_EINVAL;
    }
    IPH_TRY(require_device());
    return launch_extremes_noalloc(image, p, w, h, out, static_cast<hipStream_t>(stream));
}

}  // extern "C"
