// Transposed PFC passes on the double-butterfly plans (single-precision
// fields, FP64 twiddle products and in-register DFTs).
#include "variants_impl.hpp"

namespace pfc_amd {
const std::vector<TV<float>> &t_variants_xp(int kind, int n) {
  static const std::vector<TV<float>> empty;
  if (kind < TK_X_FWD || kind > TK_X_INV) return empty;
  PFC_AMD_SWITCH_LEN(n, static const std::vector<TV<float>> v[5] = {
      make_t_kind_xp<N_, TK_X_FWD>(), make_t_kind_xp<N_, TK_Y_FWD>(),
      make_t_kind_xp<N_, TK_Z>(), make_t_kind_xp<N_, TK_Y_INV>(),
      make_t_kind_xp<N_, TK_X_INV>()};
                        return v[kind];)
  return empty;
}
} // namespace pfc_amd
