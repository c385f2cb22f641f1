// Transposed-step variant lists and kernels for float precision.
#include "variants_impl.hpp"

namespace pfc_amd {
PFC_AMD_DEFINE_T_LISTS(float)
} // namespace pfc_amd
