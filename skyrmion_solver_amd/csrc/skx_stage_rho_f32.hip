// Stage kernels (curl pass + stage kernel) of the rho theory in FP32 storage.
#include "skx_stage.h"

namespace skx {
SKX_INSTANTIATE_STAGES(float, RHO)
}  // namespace skx
