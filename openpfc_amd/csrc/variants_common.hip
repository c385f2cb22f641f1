// Precision-independent plan queries.
#include "variants_impl.hpp"

namespace pfc_amd {

bool supported_len(int n) {
  PFC_AMD_SWITCH_LEN(n, return true;)
  return false;
}

template <typename... Ps> int count_of(PlanList<Ps...>) { return (int)sizeof...(Ps); }

template <typename... Ps>
std::vector<int> radices_in(PlanList<Ps...>, int pid) {
  std::vector<std::vector<int>> all = {RadixOf<typename Ps::List>::get()...};
  return (pid >= 0 && pid < (int)all.size()) ? all[pid] : std::vector<int>{};
}

int plan_count(int n) {
  PFC_AMD_SWITCH_LEN(n, return count_of(typename Plans<N_>::L{});)
  return 0;
}

int plan_count_xp(int n) {
  PFC_AMD_SWITCH_LEN(n, return count_of(typename PlansXP<N_>::L{});)
  return 0;
}

std::vector<int> radices_for(int n, int pid) {
  if (pid >= kXpPid) {
    PFC_AMD_SWITCH_LEN(n, return radices_in(typename PlansXP<N_>::L{}, pid - kXpPid);)
    return {};
  }
  PFC_AMD_SWITCH_LEN(n, return radices_in(typename Plans<N_>::L{}, pid);)
  return {};
}

} // namespace pfc_amd
