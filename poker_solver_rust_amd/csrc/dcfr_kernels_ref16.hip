// DCFR kernels of the SM_REF16 storage mode (device code in dcfr_kernels.hpp).
#include "dcfr_kernels.hpp"

namespace psr {
PSR_INSTANTIATE_STORAGE(SM_REF16)
}  // namespace psr
