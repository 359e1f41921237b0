// Split-precision FP32 gravity walk (placeholder until the FP32 kernels land).
#include "ag2_internal.h"

namespace ag2 {

void forces_fp32(Ctx &c, int mode) {
    (void)c;
    (void)mode;
    throw HipError("AG2_FP32_SPLIT is not available in this build");
}

}  // namespace ag2
