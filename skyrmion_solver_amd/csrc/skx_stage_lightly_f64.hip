// Stage kernels of the lightly bound Skyrme theory in FP64.
#include "skx_stage.h"

namespace skx {
SKX_INSTANTIATE_STAGES(double, LIGHTLY)
}  // namespace skx
