// Stage kernels of the omega theory in FP32 storage.
#include "skx_stage.h"

namespace skx {
SKX_INSTANTIATE_STAGES(float, OMEGA)
}  // namespace skx
