// MFMA correlation engine for match_template (work in progress).
#include "match_common.hpp"

namespace iph {

size_t mfma_match_workspace(int32_t, uint32_t, uint32_t, uint32_t, uint32_t, int32_t) {
    return 256;
}

iph_status launch_mfma_match(const void *, int32_t, uint32_t, uint32_t, const void *, uint32_t,
                             uint32_t, int32_t, float *, void *, size_t, hipStream_t) {
    set_error("the MFMA correlation path is not built yet");
    return IPH_EUNSUPPORTED;
}

}  // namespace iph
