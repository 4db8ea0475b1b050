// DCFR kernels of the SM_F32 storage mode (device code in dcfr_kernels.hpp).
#include "dcfr_kernels.hpp"

namespace psr {
PSR_INSTANTIATE_STORAGE(SM_F32)
}  // namespace psr
