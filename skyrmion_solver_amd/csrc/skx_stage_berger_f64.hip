// Stage kernels of the Berger-Skyrme theory in FP64.
#include "skx_stage.h"

namespace skx {
SKX_INSTANTIATE_STAGES(double, BERGER)
}  // namespace skx
