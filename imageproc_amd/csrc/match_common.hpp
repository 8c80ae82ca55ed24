// Entry points of the MFMA correlation engine (match_mfma.hip).
#pragma once

#include "iph_internal.hpp"

namespace iph {

size_t mfma_match_workspace(int32_t p_type, uint32_t w, uint32_t h, uint32_t tw, uint32_t th,
                            int32_t method);
iph_status launch_mfma_match(const void *img, int32_t p_type, uint32_t w, uint32_t h,
                             const void *tpl, uint32_t tw, uint32_t th, int32_t method,
                             float *out, void *ws, size_t ws_bytes, hipStream_t s);

}  // namespace iph
