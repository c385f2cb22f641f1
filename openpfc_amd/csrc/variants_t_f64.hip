// Transposed-step variant lists and kernels for double precision.
#include "variants_impl.hpp"

namespace pfc_amd {
PFC_AMD_DEFINE_T_LISTS(double)
} // namespace pfc_amd
