// Stage kernels of the pion theory in FP64 storage.
#include "skx_stage.h"

namespace skx {
SKX_INSTANTIATE_STAGES(double, PION)
}  // namespace skx
