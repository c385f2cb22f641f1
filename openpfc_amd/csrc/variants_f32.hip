// Variant lists and kernels for single precision.
#include "variants_impl.hpp"

namespace pfc_amd {
PFC_AMD_DEFINE_LISTS(float)
} // namespace pfc_amd
