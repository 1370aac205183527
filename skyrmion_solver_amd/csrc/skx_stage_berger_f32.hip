// Stage kernels of the Berger-Skyrme theory in FP32 storage.
#include "skx_stage.h"

namespace skx {
SKX_INSTANTIATE_STAGES(float, BERGER)
}  // namespace skx
