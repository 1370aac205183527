// Stage kernels of the lightly bound Skyrme theory in FP32 storage.
#include "skx_stage.h"

namespace skx {
SKX_INSTANTIATE_STAGES(float, LIGHTLY)
}  // namespace skx
