// Entry points of the MFMA correlation engine (match_mfma.hip).
#pragma once

#include "iph_internal.hpp"

namespace iph {

// A window's score from its exact integer sums: c = sum(i t), sii = sum(i^2),
// stt = sum(t^2) over n template pixels (template_matching.rs:213-293). The
// normalised methods divide in f32 like the reference below 64 x 64 template
// pixels and in f64 from there on.
template <int METHOD>
__device__ __forceinline__ float finish_score(long long c, long long sii, long long stt,
                                              long long n) {
    const bool sse = (METHOD == IPH_SSE || METHOD == IPH_SSE_NORMED);
    long long num = sse ? (stt - 2 * c + sii) : c;
    if (METHOD == IPH_SSE || METHOD == IPH_CCORR) return (float)num;
    if (n < 64 * 64) {
        // the reference's normalisation in f32 (template_matching.rs:249-250)
        float norm = sqrtf((float)sii * (float)stt);
        float s = (float)num;
        return norm > 0.0f ? s / norm : s;
    }
    double norm = sqrt((double)sii * (double)stt);
    return norm > 0.0 ? (float)((double)num / norm) : (float)num;
}

size_t mfma_match_workspace(int32_t p_type, uint32_t w, uint32_t h, uint32_t tw, uint32_t th,
                            int32_t method);
iph_status launch_mfma_match(const void *img, int32_t p_type, uint32_t w, uint32_t h,
                             const void *tpl, uint32_t tw, uint32_t th, int32_t method,
                             float *out, void *ws, size_t ws_bytes, hipStream_t s);

}  // namespace iph
