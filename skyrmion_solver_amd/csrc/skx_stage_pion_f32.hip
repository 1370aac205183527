// Stage kernels of the pion theory in FP32 storage.
#include "skx_stage.h"

namespace skx {
SKX_INSTANTIATE_STAGES(float, PION)
}  // namespace skx
