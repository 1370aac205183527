// Stage kernels of the omega theory in FP64 storage.
#include "skx_stage.h"

namespace skx {
SKX_INSTANTIATE_STAGES(double, OMEGA)
}  // namespace skx
