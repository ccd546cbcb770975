// Learned message-passing preconditioner (BF16 MFMA): see zj_chorin.hpp.
#include <hip/hip_runtime.h>

#include "../../include/zoomy_hip.h"
#include "zj_chorin.hpp"

namespace zj {
int mpnn_apply_f32(zj_mpnn*, zj_chorin*, const float*, float*, const int*, hipStream_t) {
    return ZJ_EUNSUP;
}
}  // namespace zj
