// Stage kernels of the coulomb theory in FP32 storage.
#include "skx_stage.h"

namespace skx {
SKX_INSTANTIATE_STAGES(float, COULOMB)
}  // namespace skx
