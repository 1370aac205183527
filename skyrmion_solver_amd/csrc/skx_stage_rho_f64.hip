// Stage kernels (curl pass + stage kernel) of the rho theory in FP64 storage.
#include "skx_stage.h"

namespace skx {
SKX_INSTANTIATE_STAGES(double, RHO)
}  // namespace skx
