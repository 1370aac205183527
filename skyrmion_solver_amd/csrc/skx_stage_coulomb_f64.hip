// Stage kernels of the coulomb theory in FP64 storage.
#include "skx_stage.h"

namespace skx {
SKX_INSTANTIATE_STAGES(double, COULOMB)
}  // namespace skx
