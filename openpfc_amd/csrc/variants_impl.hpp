// variants_impl.hpp -- builds the variant lists (included once per precision).
#pragma once

#include "variants.hpp"

#include <utility>

namespace pfc_amd {

template <typename... Ps> struct PlanList {};

// Line plans per length; index 0 is the default.
template <int N> struct Plans;
template <> struct Plans<1> { using L = PlanList<FftPlan<1, 1, RadixList<1>>>; };
template <> struct Plans<2> { using L = PlanList<FftPlan<2, 2, RadixList<2>>>; };
template <> struct Plans<4> { using L = PlanList<FftPlan<4, 4, RadixList<4>>>; };
template <> struct Plans<8> { using L = PlanList<FftPlan<8, 8, RadixList<8>>>; };
template <> struct Plans<16> {
  using L = PlanList<FftPlan<16, 16, RadixList<16>>>;
};
template <> struct Plans<32> { using L = PlanList<FftPlan<32, 8, RadixList<8, 4>>>; };
template <> struct Plans<64> { using L = PlanList<FftPlan<64, 8, RadixList<8, 8>>>; };
template <> struct Plans<128> {
  using L = PlanList<FftPlan<128, 16, RadixList<16, 8>>, FftPlan<128, 8, RadixList<8, 4, 4>>>;
};
template <> struct Plans<256> {
  using L = PlanList<FftPlan<256, 16, RadixList<16, 16>>,
                     FftPlan<256, 8, RadixList<8, 8, 4>>,
                     FftPlan<256, 4, RadixList<4, 4, 4, 4>>>;
};
template <> struct Plans<512> {
  using L = PlanList<FftPlan<512, 8, RadixList<8, 8, 8>>,
                     FftPlan<512, 16, RadixList<16, 16, 2>>,
                     FftPlan<512, 4, RadixList<4, 4, 4, 4, 2>>,
                     FftPlan<512, 32, RadixList<16, 16, 2>>>;
};
// Mixed-radix lengths of the uneven-size ensembles (radix 12 = 4 x 3 and
// 6 = 2 x 3 in registers, radix 3 as its own pass).
template <> struct Plans<12> {
  using L = PlanList<FftPlan<12, 12, RadixList<12>>, FftPlan<12, 12, RadixList<4, 3>>>;
};
template <> struct Plans<24> {
  using L = PlanList<FftPlan<24, 12, RadixList<12, 2>>, FftPlan<24, 12, RadixList<6, 4>>>;
};
template <> struct Plans<48> {
  using L = PlanList<FftPlan<48, 12, RadixList<12, 4>>, FftPlan<48, 12, RadixList<4, 4, 3>>>;
};
template <> struct Plans<96> {
  using L = PlanList<FftPlan<96, 24, RadixList<12, 8>>, FftPlan<96, 12, RadixList<4, 4, 6>>>;
};
template <> struct Plans<192> {
  using L = PlanList<FftPlan<192, 12, RadixList<12, 4, 4>>,
                     FftPlan<192, 24, RadixList<8, 8, 3>>>;
};
template <> struct Plans<1024> {
  using L = PlanList<FftPlan<1024, 16, RadixList<16, 16, 4>>,
                     FftPlan<1024, 8, RadixList<8, 8, 4, 4>>,
                     FftPlan<1024, 4, RadixList<4, 4, 4, 4, 4>>,
                     FftPlan<1024, 32, RadixList<16, 16, 4>>>;
};

// Plans with double butterflies (FftPlan XP = true) for the transposed PFC
// passes of single-precision plans made with PFC_HIP_FLAG_F64_BUTTERFLIES;
// their twiddle tables hold double entries and sit at plan id kXpPid + i.
template <int N> struct PlansXP { using L = PlanList<>; };
template <> struct PlansXP<128> {
  using L = PlanList<FftPlan<128, 8, RadixList<8, 4, 4>, true>,
                     FftPlan<128, 16, RadixList<16, 8>, true>>;
};
template <> struct PlansXP<256> {
  using L = PlanList<FftPlan<256, 8, RadixList<8, 8, 4>, true>,
                     FftPlan<256, 16, RadixList<16, 16>, true>,
                     FftPlan<256, 4, RadixList<4, 4, 4, 4>, true>>;
};
template <> struct PlansXP<512> {
  using L = PlanList<FftPlan<512, 8, RadixList<8, 8, 8>, true>,
                     FftPlan<512, 16, RadixList<16, 16, 2>, true>,
                     FftPlan<512, 4, RadixList<4, 4, 4, 4, 2>, true>>;
};
template <> struct PlansXP<1024> {
  using L = PlanList<FftPlan<1024, 8, RadixList<8, 8, 4, 4>, true>,
                     FftPlan<1024, 16, RadixList<16, 16, 4>, true>,
                     FftPlan<1024, 4, RadixList<4, 4, 4, 4, 4>, true>>;
};

template <typename List> struct RadixOf;
template <int... Rs> struct RadixOf<RadixList<Rs...>> {
  static std::vector<int> get() { return {Rs...}; }
};

template <typename Plan> std::string plan_name() {
  std::string s = "r";
  auto r = RadixOf<typename Plan::List>::get();
  for (size_t i = 0; i < r.size(); ++i) s += (i ? "x" : "") + std::to_string(r[i]);
  return s + "/E" + std::to_string(Plan::E) + (Plan::XP ? "/xp" : "");
}

constexpr int kMaxLds = 147456; // bytes of static LDS a block may use
constexpr int kMaxThreads = 1024;

// ---- launchers ---------------------------------------------------------------
template <typename T, typename Plan, int L, bool S>
hipError_t col_launch(int dir, ColArgs<T> a, long n_outer, int nfields,
                      hipStream_t s) {
  a.ntx = (a.W + L - 1) / L;
  dim3 grid((unsigned)(a.ntx * n_outer), (unsigned)nfields), block(L * Plan::TL);
  if (dir < 0)
    hipLaunchKernelGGL((k_col<T, Plan, -1, L, S>), grid, block, 0, s, a);
  else
    hipLaunchKernelGGL((k_col<T, Plan, +1, L, S>), grid, block, 0, s, a);
  return hipGetLastError();
}

template <typename T, typename Plan, int L, bool S>
hipError_t pfcz_launch(PfcZArgs<T> a, long ny, hipStream_t s) {
  a.ntx = (a.W + L - 1) / L;
  dim3 grid((unsigned)(a.ntx * ny)), block(L * Plan::TL);
  switch (a.mode) {
  case ZM_PFC:
    hipLaunchKernelGGL((k_pfc_z<T, Plan, L, S, ZM_PFC>), grid, block, 0, s, a);
    break;
  case ZM_TMF:
    hipLaunchKernelGGL((k_pfc_z<T, Plan, L, S, ZM_TMF>), grid, block, 0, s, a);
    break;
  case ZM_TCOMB:
    hipLaunchKernelGGL((k_pfc_z<T, Plan, L, S, ZM_TCOMB>), grid, block, 0, s, a);
    break;
  default:
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T, typename Plan, int LPB>
hipError_t xr2c_launch(XArgs<T> a, hipStream_t s) {
  dim3 grid((unsigned)((a.rows + LPB - 1) / LPB)), block(LPB * Plan::TL);
  hipLaunchKernelGGL((k_x_r2c<T, Plan, LPB>), grid, block, 0, s, a);
  return hipGetLastError();
}

template <typename T, typename Plan, int LPB>
hipError_t xc2r_launch(XArgs<T> a, hipStream_t s) {
  dim3 grid((unsigned)((a.rows + LPB - 1) / LPB)), block(LPB * Plan::TL);
  hipLaunchKernelGGL((k_x_c2r<T, Plan, LPB>), grid, block, 0, s, a);
  return hipGetLastError();
}

template <typename T, typename Plan, int LPB>
hipError_t pfcx_launch(PfcXArgs<T> a, hipStream_t s) {
  dim3 grid((unsigned)((a.rows + LPB - 1) / LPB)), block(LPB * Plan::TL);
  hipLaunchKernelGGL((k_pfc_x<T, Plan, LPB>), grid, block, 0, s, a);
  return hipGetLastError();
}

template <typename T, typename Plan, int LPB>
hipError_t tungx_launch(TungXArgs<T> a, hipStream_t s) {
  dim3 grid((unsigned)((a.rows + LPB - 1) / LPB)), block(LPB * Plan::TL);
  hipLaunchKernelGGL((k_x_tung<T, Plan, LPB>), grid, block, 0, s, a);
  return hipGetLastError();
}

// ---- list builders ------------------------------------------------------------
// Column-shaped passes: tile widths 8, 4, 16 (128/64/256-byte rows for
// double complex), both exchange modes, kept when the block fits.
template <typename T, typename Plan, int L, bool S> constexpr bool col_fits() {
  return L * Plan::TL <= kMaxThreads && L * Plan::TL >= 1 &&
         (S ? 1 : 2) * L * row_padded(Plan::N) * (int)sizeof(T) <= kMaxLds;
}

template <typename T, typename Plan, int L, bool S>
void add_col(std::vector<ColV<T>> &v, int pid) {
  if constexpr (col_fits<T, Plan, L, S>())
    v.push_back({plan_name<Plan>() + "/L" + std::to_string(L) + (S ? "/split" : "/full"),
                 pid, &col_launch<T, Plan, L, S>});
}
template <typename T, typename Plan, int L, bool S>
void add_z(std::vector<ZV<T>> &v, int pid) {
  if constexpr (col_fits<T, Plan, L, S>())
    v.push_back({plan_name<Plan>() + "/L" + std::to_string(L) + (S ? "/split" : "/full"),
                 pid, &pfcz_launch<T, Plan, L, S>});
}

// Default first: 128-byte tile rows, exchange mode by LDS budget.
template <typename T, typename Plan> constexpr int default_tile() {
  constexpr int l = 128 / (2 * (int)sizeof(T));
  return (l * Plan::TL > kMaxThreads) ? kMaxThreads / Plan::TL : l;
}
template <typename T, typename Plan> constexpr bool default_split() {
  return 2 * default_tile<T, Plan>() * row_padded(Plan::N) * (int)sizeof(T) > 65536;
}

template <typename T, typename Plan, bool ALL>
void col_for_plan(std::vector<ColV<T>> &v, int pid) {
  constexpr int L0 = default_tile<T, Plan>();
  constexpr bool S0 = default_split<T, Plan>();
  if (pid == 0) add_col<T, Plan, L0, S0>(v, pid);
  if constexpr (ALL) {
    if (pid != 0 || L0 != 8 || S0 != true) add_col<T, Plan, 8, true>(v, pid);
    if (pid != 0 || L0 != 8 || S0 != false) add_col<T, Plan, 8, false>(v, pid);
    if (pid != 0 || L0 != 4 || S0 != true) add_col<T, Plan, 4, true>(v, pid);
    if (pid != 0 || L0 != 4 || S0 != false) add_col<T, Plan, 4, false>(v, pid);
    if (pid != 0 || L0 != 16 || S0 != true) add_col<T, Plan, 16, true>(v, pid);
    if (pid != 0 || L0 != 16 || S0 != false) add_col<T, Plan, 16, false>(v, pid);
  }
}
template <typename T, typename Plan, bool ALL>
void z_for_plan(std::vector<ZV<T>> &v, int pid) {
  constexpr int L0 = default_tile<T, Plan>();
  constexpr bool S0 = default_split<T, Plan>();
  if (pid == 0) add_z<T, Plan, L0, S0>(v, pid);
  if constexpr (ALL) {
    if (pid != 0 || L0 != 8 || S0 != true) add_z<T, Plan, 8, true>(v, pid);
    if (pid != 0 || L0 != 8 || S0 != false) add_z<T, Plan, 8, false>(v, pid);
    if (pid != 0 || L0 != 4 || S0 != true) add_z<T, Plan, 4, true>(v, pid);
    if (pid != 0 || L0 != 4 || S0 != false) add_z<T, Plan, 4, false>(v, pid);
    if (pid != 0 || L0 != 16 || S0 != true) add_z<T, Plan, 16, true>(v, pid);
    if (pid != 0 || L0 != 16 || S0 != false) add_z<T, Plan, 16, false>(v, pid);
  }
}

// Row-shaped passes: 1, 2 or 4 rows per block (or enough rows to fill a wave).
template <typename Plan> constexpr int min_rows() {
  return Plan::TL >= 64 ? 1 : 64 / Plan::TL;
}
template <typename T, typename Plan, int LPB> constexpr bool row_fits() {
  return LPB * Plan::TL <= kMaxThreads &&
         2 * LPB * row_padded(Plan::N) * (int)sizeof(T) <= kMaxLds;
}
template <typename T, typename Plan, bool ALL, typename V, typename Make>
void rows_for_plan(std::vector<V> &v, int pid, Make make) {
  constexpr int R0 = min_rows<Plan>();
  make(std::integral_constant<int, R0>{});
  if constexpr (ALL) {
    if constexpr (row_fits<T, Plan, 2 * R0>()) make(std::integral_constant<int, 2 * R0>{});
    if constexpr (row_fits<T, Plan, 4 * R0>()) make(std::integral_constant<int, 4 * R0>{});
  }
  (void)pid;
}

// Plans with 32 values per thread serve the single-precision transposed
// kernels only (in double they would hold 64 registers of data per thread).
template <typename T, typename P, bool TRANSPOSED> constexpr bool use_plan() {
  return P::E < 32 || (TRANSPOSED && sizeof(T) == 4);
}
template <typename T, bool TRANSPOSED, typename P, typename F>
void call_plan(F &f, int pid) {
  if constexpr (use_plan<T, P, TRANSPOSED>()) f(P{}, pid);
}
template <typename T, int N, typename F, typename... Ps>
void for_plans(PlanList<Ps...>, F &&f) {
  int pid = 0;
  ((call_plan<T, false, Ps>(f, pid), ++pid), ...);
}
template <typename T, int N, typename F, typename... Ps>
void for_t_plans(PlanList<Ps...>, F &&f) {
  int pid = 0;
  ((call_plan<T, true, Ps>(f, pid), ++pid), ...);
}
template <typename F, typename... Ps> void for_t_plans_xp(PlanList<Ps...>, F &&f) {
  (f(Ps{}), ...);
}

// Lengths with several plans and tile shapes get the full sweep; the rest a
// single default variant.
template <int N> constexpr bool sweep_len() { return N >= 48; }

template <typename T, int N> std::vector<ColV<T>> make_col() {
  std::vector<ColV<T>> v;
  for_plans<T, N>(typename Plans<N>::L{}, [&](auto p, int pid) {
    col_for_plan<T, decltype(p), sweep_len<N>()>(v, pid);
  });
  return v;
}
template <typename T, int N> std::vector<ZV<T>> make_z() {
  std::vector<ZV<T>> v;
  for_plans<T, N>(typename Plans<N>::L{}, [&](auto p, int pid) {
    z_for_plan<T, decltype(p), sweep_len<N>()>(v, pid);
  });
  return v;
}
template <typename T, int N> std::vector<XV<T>> make_xr2c() {
  std::vector<XV<T>> v;
  for_plans<T, N>(typename Plans<N>::L{}, [&](auto p, int pid) {
    using P = decltype(p);
    rows_for_plan<T, P, sweep_len<N>()>(v, pid, [&](auto r) {
      constexpr int R = decltype(r)::value;
      v.push_back({plan_name<P>() + "/rows" + std::to_string(R), pid,
                   &xr2c_launch<T, P, R>});
    });
  });
  return v;
}
template <typename T, int N> std::vector<XV<T>> make_xc2r() {
  std::vector<XV<T>> v;
  for_plans<T, N>(typename Plans<N>::L{}, [&](auto p, int pid) {
    using P = decltype(p);
    rows_for_plan<T, P, sweep_len<N>()>(v, pid, [&](auto r) {
      constexpr int R = decltype(r)::value;
      v.push_back({plan_name<P>() + "/rows" + std::to_string(R), pid,
                   &xc2r_launch<T, P, R>});
    });
  });
  return v;
}
template <typename T, int N> std::vector<TXV<T>> make_tungx() {
  std::vector<TXV<T>> v;
  for_plans<T, N>(typename Plans<N>::L{}, [&](auto p, int pid) {
    using P = decltype(p);
    rows_for_plan<T, P, sweep_len<N>()>(v, pid, [&](auto r) {
      constexpr int R = decltype(r)::value;
      v.push_back({plan_name<P>() + "/rows" + std::to_string(R), pid,
                   &tungx_launch<T, P, R>});
    });
  });
  return v;
}
template <typename T, int N> std::vector<PXV<T>> make_pfcx() {
  std::vector<PXV<T>> v;
  for_plans<T, N>(typename Plans<N>::L{}, [&](auto p, int pid) {
    using P = decltype(p);
    rows_for_plan<T, P, sweep_len<N>()>(v, pid, [&](auto r) {
      constexpr int R = decltype(r)::value;
      v.push_back({plan_name<P>() + "/rows" + std::to_string(R), pid,
                   &pfcx_launch<T, P, R>});
    });
  });
  return v;
}

// ---- transposed-step kernels ----------------------------------------------------
template <typename T, typename Plan, int G, bool S, bool TILE>
hipError_t tx_fwd_launch(TStepArgs<T> a, hipStream_t s) {
  dim3 grid((unsigned)(a.nz * ((a.ny + G - 1) / G))), block(G * Plan::TL);
  hipLaunchKernelGGL((k_tx_fwd<T, Plan, G, S, TILE>), grid, block, 0, s, a);
  return hipGetLastError();
}
template <typename T, typename Plan, int G, bool S, bool TILE>
hipError_t ty_fwd_launch(TStepArgs<T> a, hipStream_t s) {
  if (a.nfields != 1 && a.nfields != 2) return hipErrorInvalidValue;
  dim3 grid((unsigned)(a.W * ((a.nz + G - 1) / G)), (unsigned)a.nfields),
      block(G * Plan::TL);
  hipLaunchKernelGGL((k_ty_fwd<T, Plan, G, S, TILE>), grid, block, 0, s, a);
  return hipGetLastError();
}
template <typename T, typename Plan, int G, bool S, bool TILE>
hipError_t tz_launch(TStepArgs<T> a, hipStream_t s) {
  const long lines = (long)a.W * a.ny;
  dim3 grid((unsigned)((lines + G - 1) / G)), block(G * Plan::TL);
  switch (a.mode) {
  case ZM_PFC:
    hipLaunchKernelGGL((k_tz<T, Plan, G, S, ZM_PFC>), grid, block, 0, s, a);
    break;
  case ZM_TMF:
    hipLaunchKernelGGL((k_tz<T, Plan, G, S, ZM_TMF>), grid, block, 0, s, a);
    break;
  case ZM_TCOMB:
    hipLaunchKernelGGL((k_tz<T, Plan, G, S, ZM_TCOMB>), grid, block, 0, s, a);
    break;
  default:
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
template <typename T, typename Plan, int G, bool S, bool TILE>
hipError_t ty_inv_launch(TStepArgs<T> a, hipStream_t s) {
  dim3 grid((unsigned)(a.W * ((a.nz + G - 1) / G))), block(G * Plan::TL);
  hipLaunchKernelGGL((k_ty_inv<T, Plan, G, S, TILE>), grid, block, 0, s, a);
  return hipGetLastError();
}
template <typename T, typename Plan, int G, bool S, bool TILE, bool NT = false>
hipError_t tx_inv_launch(TStepArgs<T> a, hipStream_t s) {
  dim3 grid((unsigned)(a.nz * ((a.ny + G - 1) / G))), block(G * Plan::TL);
  if constexpr (NT)
    hipLaunchKernelGGL((k_tx_inv<T, Plan, G, TILE, true>), grid, block, 0, s, a);
  else
    hipLaunchKernelGGL((k_tx_inv<T, Plan, G, TILE>), grid, block, 0, s, a);
  return hipGetLastError();
}

template <typename T, typename Plan, int G, bool S, bool TILE>
hipError_t tx_r2c_launch(TStepArgs<T> a, hipStream_t s) {
  dim3 grid((unsigned)(a.nz * ((a.ny + G - 1) / G))), block(G * Plan::TL);
  hipLaunchKernelGGL((k_tx_r2c<T, Plan, G, TILE>), grid, block, 0, s, a);
  return hipGetLastError();
}
template <typename T, typename Plan, int G, bool S, bool TILE>
hipError_t tx_tung_launch(TStepArgs<T> a, hipStream_t s) {
  dim3 grid((unsigned)(a.nz * ((a.ny + G - 1) / G))), block(G * Plan::TL);
  hipLaunchKernelGGL((k_tx_tung<T, Plan, G, TILE>), grid, block, 0, s, a);
  return hipGetLastError();
}

template <typename T, typename Plan, int G, bool S, bool TILE>
hipError_t tx_fwd2_launch(TStepArgs<T> a, hipStream_t s) {
  dim3 grid((unsigned)(a.nz * ((a.ny + G - 1) / G))), block(G * Plan::TL);
  hipLaunchKernelGGL((k_tx_fwd2<T, Plan, G, TILE>), grid, block, 0, s, a);
  return hipGetLastError();
}

template <typename T, typename Plan, int G, bool S, bool NT = false>
hipError_t tx_fwd_w_launch(TStepArgs<T> a, hipStream_t s) {
  dim3 grid((unsigned)(a.nz * ((a.ny + G - 1) / G))), block(G * Plan::TL);
  if constexpr (NT)
    hipLaunchKernelGGL((k_tx_fwd_w<T, Plan, G, S, true>), grid, block, 0, s, a);
  else
    hipLaunchKernelGGL((k_tx_fwd_w<T, Plan, G, S>), grid, block, 0, s, a);
  return hipGetLastError();
}

template <typename T, typename Plan, int G, bool S>
hipError_t tz_pre_launch(TStepArgs<T> a, hipStream_t s) {
  const long lines = (long)a.W * a.ny;
  dim3 grid((unsigned)((lines + G - 1) / G)), block(G * Plan::TL);
  switch (a.mode) {
  case ZM_PFC:
    hipLaunchKernelGGL((k_tz<T, Plan, G, S, ZM_PFC, true>), grid, block, 0, s, a);
    break;
  case ZM_TMF:
    hipLaunchKernelGGL((k_tz<T, Plan, G, S, ZM_TMF, true>), grid, block, 0, s, a);
    break;
  case ZM_TCOMB:
    hipLaunchKernelGGL((k_tz<T, Plan, G, S, ZM_TCOMB, true>), grid, block, 0, s, a);
    break;
  default:
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
// z pass with non-temporal line loads and stores (k_tz<..., NT = true>)
template <typename T, typename Plan, int G, bool S, bool PRE>
hipError_t tz_nt_launch(TStepArgs<T> a, hipStream_t s) {
  const long lines = (long)a.W * a.ny;
  dim3 grid((unsigned)((lines + G - 1) / G)), block(G * Plan::TL);
  switch (a.mode) {
  case ZM_PFC:
    hipLaunchKernelGGL((k_tz<T, Plan, G, S, ZM_PFC, PRE, true>), grid, block, 0, s, a);
    break;
  case ZM_TMF:
    hipLaunchKernelGGL((k_tz<T, Plan, G, S, ZM_TMF, PRE, true>), grid, block, 0, s, a);
    break;
  case ZM_TCOMB:
    hipLaunchKernelGGL((k_tz<T, Plan, G, S, ZM_TCOMB, PRE, true>), grid, block, 0, s, a);
    break;
  default:
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
// z pass with the multipliers tabulated per line in LDS (k_tz_mt)
// (PAIR: lines ky and ny - ky share one table, G = 2)
template <typename T, typename Plan, int G, bool S, bool PAIR = false>
hipError_t tz_mt_launch(TStepArgs<T> a, hipStream_t s) {
  const long lines = (long)a.W * a.ny;
  const long blocks = PAIR ? (long)a.W * ((a.ny + 1) / 2) : (lines + G - 1) / G;
  dim3 grid((unsigned)blocks), block(G * Plan::TL);
  switch (a.mode) {
  case ZM_PFC:
    hipLaunchKernelGGL((k_tz_mt<T, Plan, G, S, ZM_PFC, PAIR>), grid, block, 0, s, a);
    break;
  case ZM_TMF:
    hipLaunchKernelGGL((k_tz_mt<T, Plan, G, S, ZM_TMF, PAIR>), grid, block, 0, s, a);
    break;
  case ZM_TCOMB:
    hipLaunchKernelGGL((k_tz_mt<T, Plan, G, S, ZM_TCOMB, PAIR>), grid, block, 0, s, a);
    break;
  default:
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
template <typename T, typename Plan, int G>
hipError_t tx_r2c_wide_launch(TStepArgs<T> a, hipStream_t s) {
  dim3 grid((unsigned)(a.nz * ((a.ny + G - 1) / G))), block(G * Plan::TL);
  hipLaunchKernelGGL((k_tx_r2c<T, Plan, G, true, true>), grid, block, 0, s, a);
  return hipGetLastError();
}
template <typename T, typename Plan, int G>
hipError_t tx_tung_wide_launch(TStepArgs<T> a, hipStream_t s) {
  dim3 grid((unsigned)(a.nz * ((a.ny + G - 1) / G))), block(G * Plan::TL);
  hipLaunchKernelGGL((k_tx_tung<T, Plan, G, true, true>), grid, block, 0, s, a);
  return hipGetLastError();
}
template <typename T, typename Plan, int G, int WPE>
hipError_t tx_tung_wide_o_launch(TStepArgs<T> a, hipStream_t s) {
  dim3 grid((unsigned)(a.nz * ((a.ny + G - 1) / G))), block(G * Plan::TL);
  hipLaunchKernelGGL((k_tx_tung_o<T, Plan, G, true, true, WPE>), grid, block, 0, s, a);
  return hipGetLastError();
}

template <typename T, typename PH, typename PF, int G, bool S, int WPE = 1>
hipError_t tx_fuse_launch(TStepArgs<T> a, hipStream_t s) {
  dim3 grid((unsigned)(a.nz * ((a.ny + G - 1) / G))), block(G * PF::TL);
  hipLaunchKernelGGL((k_tx_fuse<T, PH, PF, G, S, WPE>), grid, block, 0, s, a);
  return hipGetLastError();
}

template <typename T, typename Plan, int G, bool S, int NR, bool NT = false>
hipError_t ty_fwd_w_launch(TStepArgs<T> a, hipStream_t s) {
  if (a.nfields != 1 && a.nfields != 2) return hipErrorInvalidValue;
  dim3 grid((unsigned)(a.W * ((a.nz + G - 1) / G)), (unsigned)a.nfields), block(G * Plan::TL);
  if constexpr (NT)
    hipLaunchKernelGGL((k_ty_fwd_w<T, Plan, G, S, NR, true>), grid, block, 0, s, a);
  else
    hipLaunchKernelGGL((k_ty_fwd_w<T, Plan, G, S, NR>), grid, block, 0, s, a);
  return hipGetLastError();
}
template <typename T, typename Plan, int G, bool S, int NR, bool NT = false>
hipError_t ty_inv_w_launch(TStepArgs<T> a, hipStream_t s) {
  dim3 grid((unsigned)(a.W * ((a.nz + G - 1) / G))), block(G * Plan::TL);
  if constexpr (NT)
    hipLaunchKernelGGL((k_ty_inv_w<T, Plan, G, S, NR, true>), grid, block, 0, s, a);
  else
    hipLaunchKernelGGL((k_ty_inv_w<T, Plan, G, S, NR>), grid, block, 0, s, a);
  return hipGetLastError();
}

template <typename T, typename Plan, int G, bool S, bool TILE, int KIND>
constexpr bool t_fits() {
  if (G * Plan::TL > kMaxThreads || G * Plan::TL < 64) return false;
  if (KIND == TK_Z && !TILE) return false; // the z pass has no transposition
  constexpr int fft = (S ? 1 : 2) * G * row_padded(Plan::N) * (int)sizeof(T);
  // row kernels on the nx/2-point plan use the full exchange only
  if (KIND == TK_X_INV || KIND == TK_X_TUNG) {
    constexpr int stage = TILE ? 2 * (Plan::N + 1) * tile_pitch<G, T>() * (int)sizeof(T) : 0;
    constexpr int full = 2 * G * row_padded(Plan::N) * (int)sizeof(T);
    return !S && (stage > full ? stage : full) <= kMaxLds;
  }
  if (KIND == TK_X_R2C) return !S && 2 * G * row_padded(Plan::N) * (int)sizeof(T) <= kMaxLds;
  return fft <= kMaxLds;
}

template <typename T, typename Plan, int G, bool S, bool TILE, int KIND>
void add_t(std::vector<TV<T>> &v, int pid) {
  if constexpr (t_fits<T, Plan, G, S, TILE, KIND>()) {
    constexpr bool rows_m = KIND == TK_X_INV || KIND == TK_X_R2C || KIND == TK_X_TUNG;
    const std::string nm = plan_name<Plan>() + "/G" + std::to_string(G) +
                           (rows_m ? "" : (S ? "/split" : "/full")) +
                           (TILE ? "" : "/direct");
    if constexpr (KIND == TK_X_FWD)
      v.push_back({nm, pid, &tx_fwd_launch<T, Plan, G, S, TILE>});
    if constexpr (KIND == TK_Y_FWD)
      v.push_back({nm, pid, &ty_fwd_launch<T, Plan, G, S, TILE>});
    if constexpr (KIND == TK_Z) v.push_back({nm, pid, &tz_launch<T, Plan, G, S, TILE>});
    if constexpr (KIND == TK_Y_INV)
      v.push_back({nm, pid, &ty_inv_launch<T, Plan, G, S, TILE>});
    if constexpr (KIND == TK_X_INV)
      v.push_back({nm, pid, &tx_inv_launch<T, Plan, G, S, TILE>});
    if constexpr (KIND == TK_X_R2C)
      v.push_back({nm, pid, &tx_r2c_launch<T, Plan, G, S, TILE>});
    if constexpr (KIND == TK_X_TUNG)
      v.push_back({nm, pid, &tx_tung_launch<T, Plan, G, S, TILE>});
  }
}

// Lines per block giving 128-byte chunks first (8 in double, 16 in float),
// then half and double that; LDS-tiled and direct transpositions.
template <typename T, typename Plan, int KIND>
void t_for_plan(std::vector<TV<T>> &v, int pid) {
  constexpr int G0 = 128 / (2 * (int)sizeof(T));
  add_t<T, Plan, G0, true, true, KIND>(v, pid);
  add_t<T, Plan, G0, false, true, KIND>(v, pid);
  add_t<T, Plan, G0 / 2, true, true, KIND>(v, pid);
  add_t<T, Plan, G0 / 2, false, true, KIND>(v, pid);
  add_t<T, Plan, 2 * G0, true, true, KIND>(v, pid);
  add_t<T, Plan, G0, true, false, KIND>(v, pid);
  add_t<T, Plan, G0, false, false, KIND>(v, pid);
  add_t<T, Plan, G0 / 2, false, false, KIND>(v, pid);
}

// PFC forward x on the nx/2-point plans (k_tx_fwd2), appended to the
// forward-x list of length N with half = 1.
template <typename T, typename Plan, int G, bool TILE>
void add_t_fwd2(std::vector<TV<T>> &v, int pid) {
  if constexpr (t_fits<T, Plan, G, false, TILE, TK_X_R2C>()) {
    TV<T> e{plan_name<Plan>() + "/G" + std::to_string(G) + (TILE ? "" : "/direct") + "/r2c2",
            pid, &tx_fwd2_launch<T, Plan, G, false, TILE>};
    e.half = 1;
    v.push_back(e);
  }
}

// PFC forward x with the wide transposed store (k_tx_fwd_w).
// NT: non-temporal loads and stores of the lines (the kernels' NT parameter).
template <typename T, typename Plan, int G, bool S, bool NT = false>
void add_t_fwd_w(std::vector<TV<T>> &v, int pid) {
  if constexpr (G >= 1 && G * Plan::TL <= kMaxThreads && G * Plan::TL >= 64 &&
                tx_fwd_w_lds<T, Plan, G, S>() * (int)sizeof(T) <= kMaxLds)
    v.push_back({plan_name<Plan>() + "/G" + std::to_string(G) + (S ? "/split" : "/full") +
                     "/wide" + (NT ? "/nt" : ""),
                 pid, &tx_fwd_w_launch<T, Plan, G, S, NT>});
}

// y passes with NR-round transpositions (k_ty_fwd_w / k_ty_inv_w).
template <typename T, typename Plan, int G, bool S, int NR, int KIND, bool NT = false>
void add_t_y_w(std::vector<TV<T>> &v, int pid) {
  if constexpr (Plan::E % NR == 0 && G * Plan::TL <= kMaxThreads && G * Plan::TL >= 64 &&
                ty_w_lds<T, Plan, G, S, NR>() * (int)sizeof(T) <= kMaxLds) {
    const std::string nm = plan_name<Plan>() + "/G" + std::to_string(G) +
                           (S ? "/split" : "/full") + "/wide" + std::to_string(NR) +
                           (NT ? "/nt" : "");
    if constexpr (KIND == TK_Y_FWD)
      v.push_back({nm, pid, &ty_fwd_w_launch<T, Plan, G, S, NR, NT>});
    if constexpr (KIND == TK_Y_INV)
      v.push_back({nm, pid, &ty_inv_w_launch<T, Plan, G, S, NR, NT>});
  }
}
// inverse x with the LDS-tiled transposed load and non-temporal lines
template <typename T, typename Plan, int G>
void add_t_x_inv_nt(std::vector<TV<T>> &v, int pid) {
  if constexpr (t_fits<T, Plan, G, false, true, TK_X_INV>())
    v.push_back({plan_name<Plan>() + "/G" + std::to_string(G) + "/nt", pid,
                 &tx_inv_launch<T, Plan, G, false, true, true>});
}

// z pass with the loads issued before the multipliers (k_tz<..., PRE>)
template <typename T, typename Plan, int G, bool S>
void add_t_z_pre(std::vector<TV<T>> &v, int pid) {
  if constexpr (t_fits<T, Plan, G, S, true, TK_Z>())
    v.push_back({plan_name<Plan>() + "/G" + std::to_string(G) + (S ? "/split" : "/full") +
                     "/pre",
                 pid, &tz_pre_launch<T, Plan, G, S>});
}
template <typename T, typename Plan, int G, bool S, bool PRE>
void add_t_z_nt(std::vector<TV<T>> &v, int pid) {
  if constexpr (t_fits<T, Plan, G, S, true, TK_Z>())
    v.push_back({plan_name<Plan>() + "/G" + std::to_string(G) + (S ? "/split" : "/full") +
                     (PRE ? "/pre" : "") + "/nt",
                 pid, &tz_nt_launch<T, Plan, G, S, PRE>});
}
template <typename T, typename Plan, int G, bool S, bool PAIR = false>
void add_t_z_mt(std::vector<TV<T>> &v, int pid) {
  if constexpr (t_fits<T, Plan, G, S, true, TK_Z>() && (!PAIR || G == 2) &&
                (S ? 1 : 2) * G * row_padded(Plan::N) * (int)sizeof(T) +
                        tz_table_bytes<T, Plan, G>() <=
                    kMaxLds)
    v.push_back({plan_name<Plan>() + "/G" + std::to_string(G) + (S ? "/split" : "/full") +
                     (PAIR ? "/mtp" : "/mt"),
                 pid, &tz_mt_launch<T, Plan, G, S, PAIR>});
}
// Fused step boundary (k_tx_fuse): an nx/2-point plan PH paired with an
// nx-point plan PF of the same threads per line; WPE > 1 caps the registers
// for WPE waves per SIMD (named /wN).
template <typename T, typename PH, typename PF, int G, bool S, int WPE = 1>
void add_t_fuse(std::vector<TV<T>> &v, int pid, int pid2) {
  if constexpr (PH::TL == PF::TL && G * PF::TL <= kMaxThreads && G * PF::TL >= 64 &&
                tx_fuse_lds<T, PH, PF, G, S>() * (int)sizeof(T) <= kMaxLds) {
    TV<T> e{plan_name<PH>() + "+" + plan_name<PF>() + "/G" + std::to_string(G) +
                (S ? "/split" : "/full") + (WPE > 1 ? "/w" + std::to_string(WPE) : ""),
            pid, &tx_fuse_launch<T, PH, PF, G, S, WPE>};
    e.plan_id2 = pid2;
    v.push_back(e);
  }
}
// Tungsten fused x pass, wide store, registers capped for WPE waves per SIMD
template <typename T, typename Plan, int G, int WPE>
void add_t_tung_wide_o(std::vector<TV<T>> &v, int pid) {
  if constexpr (G * Plan::TL <= kMaxThreads && G * Plan::TL >= 64 &&
                tx_inv_lds<T, Plan, G, true>() * (int)sizeof(T) <= kMaxLds)
    v.push_back({plan_name<Plan>() + "/G" + std::to_string(G) + "/wide/w" + std::to_string(WPE),
                 pid, &tx_tung_wide_o_launch<T, Plan, G, WPE>});
}
// Tungsten row kernels with the wide transposed store
template <typename T, typename Plan, int G, int KIND>
void add_t_rows_wide(std::vector<TV<T>> &v, int pid) {
  if constexpr (G * Plan::TL <= kMaxThreads && G * Plan::TL >= 64 &&
                tx_inv_lds<T, Plan, G, true>() * (int)sizeof(T) <= kMaxLds) {
    const std::string nm = plan_name<Plan>() + "/G" + std::to_string(G) + "/wide";
    if constexpr (KIND == TK_X_R2C) v.push_back({nm, pid, &tx_r2c_wide_launch<T, Plan, G>});
    if constexpr (KIND == TK_X_TUNG) v.push_back({nm, pid, &tx_tung_wide_launch<T, Plan, G>});
  }
}

// Transposed-step kernels exist for power-of-two lengths from 128 up.
template <typename T, int N, int KIND> std::vector<TV<T>> make_t_kind() {
  std::vector<TV<T>> v;
  if constexpr (N >= 128 && (N & (N - 1)) == 0) {
    for_t_plans<T, N>(typename Plans<N>::L{}, [&](auto p, int pid) {
      t_for_plan<T, decltype(p), KIND>(v, pid);
    });
    if constexpr (KIND == TK_X_FWD) {
      for_t_plans<T, N / 2>(typename Plans<N / 2>::L{}, [&](auto p, int pid) {
        using P = decltype(p);
        constexpr int G0 = 128 / (2 * (int)sizeof(T));
        add_t_fwd2<T, P, G0, true>(v, pid);
        add_t_fwd2<T, P, G0 / 2, true>(v, pid);
        add_t_fwd2<T, P, 2 * G0, true>(v, pid);
        add_t_fwd2<T, P, G0, false>(v, pid);
      });
    }
    if constexpr (KIND == TK_X_FWD) {
      for_t_plans<T, N>(typename Plans<N>::L{}, [&](auto p, int pid) {
        using P = decltype(p);
        constexpr int G0 = 128 / (2 * (int)sizeof(T));
        add_t_fwd_w<T, P, G0 / 2, true>(v, pid);
        add_t_fwd_w<T, P, G0 / 2, false>(v, pid);
        add_t_fwd_w<T, P, G0, true>(v, pid);
        add_t_fwd_w<T, P, G0, false>(v, pid);
        add_t_fwd_w<T, P, G0 / 4, true>(v, pid);
        add_t_fwd_w<T, P, G0, true, true>(v, pid);
        add_t_fwd_w<T, P, G0, false, true>(v, pid);
      });
    }
    if constexpr (KIND == TK_Y_FWD || KIND == TK_Y_INV) {
      for_t_plans<T, N>(typename Plans<N>::L{}, [&](auto p, int pid) {
        using P = decltype(p);
        constexpr int G0 = 128 / (2 * (int)sizeof(T));
        add_t_y_w<T, P, G0, true, 2, KIND>(v, pid);
        add_t_y_w<T, P, G0, true, 4, KIND>(v, pid);
        add_t_y_w<T, P, G0, false, 2, KIND>(v, pid);
        add_t_y_w<T, P, G0 / 2, true, 1, KIND>(v, pid);
        add_t_y_w<T, P, G0 / 2, true, 2, KIND>(v, pid);
        add_t_y_w<T, P, 2 * G0, true, 4, KIND>(v, pid);
        add_t_y_w<T, P, G0, true, 1, KIND, true>(v, pid);
        add_t_y_w<T, P, G0, true, 2, KIND, true>(v, pid);
        add_t_y_w<T, P, G0, true, 4, KIND, true>(v, pid);
      });
    }
    if constexpr (KIND == TK_X_INV) {
      for_t_plans<T, N>(typename Plans<N>::L{}, [&](auto p, int pid) {
        add_t_x_inv_nt<T, decltype(p), 128 / (2 * (int)sizeof(T))>(v, pid);
      });
    }
    if constexpr (KIND == TK_Z) {
      // one or two lines per block: waves of different blocks drift apart, so
      // loads of one overlap the transforms of another
      for_t_plans<T, N>(typename Plans<N>::L{}, [&](auto p, int pid) {
        using P = decltype(p);
        add_t<T, P, 1, true, true, KIND>(v, pid);
        add_t<T, P, 1, false, true, KIND>(v, pid);
        add_t<T, P, 2, true, true, KIND>(v, pid);
        add_t<T, P, 2, false, true, KIND>(v, pid);
        add_t_z_pre<T, P, 1, true>(v, pid);
        add_t_z_pre<T, P, 1, false>(v, pid);
        add_t_z_pre<T, P, 2, true>(v, pid);
        add_t_z_nt<T, P, 1, true, false>(v, pid);
        add_t_z_nt<T, P, 1, false, false>(v, pid);
        add_t_z_nt<T, P, 1, true, true>(v, pid);
        add_t_z_nt<T, P, 1, false, true>(v, pid);
        // tabulated multipliers: one wavefront of lines per block at least
        constexpr int GM = P::TL >= 64 ? 1 : 64 / P::TL;
        add_t_z_mt<T, P, GM, true>(v, pid);
        add_t_z_mt<T, P, GM, false>(v, pid);
        add_t_z_mt<T, P, 2 * GM, true>(v, pid);
        add_t_z_mt<T, P, 2, true, true>(v, pid);
        add_t_z_mt<T, P, 2, false, true>(v, pid);
      });
    }
    if constexpr (KIND == TK_X_FUSE) {
      // N is nx/2 here; the forward half of the pass runs the 2N-point plans
      if constexpr (2 * N <= 1024) {
        for_t_plans<T, N>(typename Plans<N>::L{}, [&](auto ph, int pid) {
          using PH = decltype(ph);
          for_t_plans<T, 2 * N>(typename Plans<2 * N>::L{}, [&](auto pf, int pid2) {
            using PF = decltype(pf);
            constexpr int G0 = 128 / (2 * (int)sizeof(T));
            add_t_fuse<T, PH, PF, G0 / 2, true>(v, pid, pid2);
            add_t_fuse<T, PH, PF, G0 / 2, false>(v, pid, pid2);
            add_t_fuse<T, PH, PF, G0, true>(v, pid, pid2);
            add_t_fuse<T, PH, PF, G0, false>(v, pid, pid2);
            // registers capped so that the blocks the LDS allows all fit: four of
            // G0 / 2 lines or two of G0 lines (128-byte stores) per CU
            if constexpr (4 * tx_fuse_lds<T, PH, PF, G0 / 2, true>() * (int)sizeof(T) <=
                              160 * 1024 &&
                          (G0 / 2) * PF::TL / 64 <= 8)
              add_t_fuse<T, PH, PF, G0 / 2, true, 4 * (G0 / 2) * PF::TL / 256>(v, pid, pid2);
            if constexpr (2 * tx_fuse_lds<T, PH, PF, G0, true>() * (int)sizeof(T) <=
                              160 * 1024 &&
                          G0 * PF::TL / 128 <= 8 && G0 * PF::TL / 128 > 1)
              add_t_fuse<T, PH, PF, G0, true, 2 * G0 * PF::TL / 256>(v, pid, pid2);
          });
        });
      }
    }
    if constexpr (KIND == TK_X_R2C || KIND == TK_X_TUNG) {
      for_t_plans<T, N>(typename Plans<N>::L{}, [&](auto p, int pid) {
        using P = decltype(p);
        constexpr int G0 = 128 / (2 * (int)sizeof(T));
        add_t_rows_wide<T, P, G0 / 2, KIND>(v, pid);
        add_t_rows_wide<T, P, G0, KIND>(v, pid);
        if constexpr (KIND == TK_X_TUNG && P::E == 8 && sizeof(T) == 8)
          add_t_tung_wide_o<T, P, 8, 4>(v, pid);
      });
    }
  }
  return v;
}

// Transposed PFC passes on the double-butterfly plans: the launch shapes
// the single-precision autotuner favours, a few per kind (16 or 8 lines per
// block for the row and transposing passes, 4 when 16 lines exceed a block;
// the z pass with the fewest lines that fill a wave, and twice that).
template <int N, int KIND> std::vector<TV<float>> make_t_kind_xp() {
  using T = float;
  std::vector<TV<T>> v;
  int pid = kXpPid;
  auto each = [&](auto f) {
    pid = kXpPid;
    for_t_plans_xp(typename PlansXP<N>::L{}, [&](auto p) { f(p, pid++); });
  };
  if constexpr (KIND == TK_X_FWD) {
    each([&](auto p, int id) {
      using P = decltype(p);
      add_t_fwd_w<T, P, 16, true>(v, id);
      add_t_fwd_w<T, P, 16, false>(v, id);
      add_t_fwd_w<T, P, 8, true>(v, id);
      add_t_fwd_w<T, P, 16, true, true>(v, id);
      add_t<T, P, 16, true, true, KIND>(v, id);
      if constexpr (16 * P::TL > kMaxThreads) add_t_fwd_w<T, P, 4, true>(v, id);
    });
  }
  if constexpr (KIND == TK_Y_FWD || KIND == TK_Y_INV) {
    each([&](auto p, int id) {
      using P = decltype(p);
      add_t_y_w<T, P, 16, true, 1, KIND>(v, id);
      add_t_y_w<T, P, 16, true, 2, KIND>(v, id);
      add_t_y_w<T, P, 8, true, 2, KIND>(v, id);
      add_t_y_w<T, P, 16, true, 2, KIND, true>(v, id);
      add_t<T, P, 16, true, true, KIND>(v, id);
      if constexpr (16 * P::TL > kMaxThreads) add_t_y_w<T, P, 4, true, 2, KIND>(v, id);
    });
  }
  if constexpr (KIND == TK_Z) {
    each([&](auto p, int id) {
      using P = decltype(p);
      constexpr int G = P::TL >= 64 ? 1 : 64 / P::TL;
      add_t<T, P, G, true, true, KIND>(v, id);
      add_t<T, P, G, false, true, KIND>(v, id);
      add_t_z_pre<T, P, G, false>(v, id);
      add_t_z_nt<T, P, G, false, true>(v, id);
      add_t_z_nt<T, P, G, true, false>(v, id);
      add_t<T, P, 2 * G, true, true, KIND>(v, id);
      add_t_z_mt<T, P, G, false>(v, id);
      add_t_z_mt<T, P, G, true>(v, id);
      add_t_z_mt<T, P, 2, true, true>(v, id);
    });
  }
  if constexpr (KIND == TK_X_INV) {
    each([&](auto p, int id) {
      using P = decltype(p);
      add_t_x_inv_nt<T, P, 16>(v, id);
      add_t<T, P, 16, false, true, KIND>(v, id);
      add_t<T, P, 8, false, true, KIND>(v, id);
      if constexpr (16 * P::TL > kMaxThreads) add_t<T, P, 4, false, true, KIND>(v, id);
    });
  }
  return v;
}

#define PFC_AMD_SWITCH_LEN(n, ...)                                            \
  switch (n) {                                                                 \
  case 1: { constexpr int N_ = 1; __VA_ARGS__ }                                      \
  case 2: { constexpr int N_ = 2; __VA_ARGS__ }                                      \
  case 4: { constexpr int N_ = 4; __VA_ARGS__ }                                      \
  case 8: { constexpr int N_ = 8; __VA_ARGS__ }                                      \
  case 16: { constexpr int N_ = 16; __VA_ARGS__ }                                    \
  case 32: { constexpr int N_ = 32; __VA_ARGS__ }                                    \
  case 64: { constexpr int N_ = 64; __VA_ARGS__ }                                    \
  case 12: { constexpr int N_ = 12; __VA_ARGS__ }                                    \
  case 24: { constexpr int N_ = 24; __VA_ARGS__ }                                    \
  case 48: { constexpr int N_ = 48; __VA_ARGS__ }                                    \
  case 96: { constexpr int N_ = 96; __VA_ARGS__ }                                    \
  case 128: { constexpr int N_ = 128; __VA_ARGS__ }                                  \
  case 192: { constexpr int N_ = 192; __VA_ARGS__ }                                  \
  case 256: { constexpr int N_ = 256; __VA_ARGS__ }                                  \
  case 512: { constexpr int N_ = 512; __VA_ARGS__ }                                  \
  case 1024: { constexpr int N_ = 1024; __VA_ARGS__ }                                \
  default: break;                                                              \
  }

// Lists of the column- and row-shaped passes (one translation unit per
// precision) and of the transposed-step kernels (another one).
#define PFC_AMD_DEFINE_LISTS(T)                                                \
  template <> const std::vector<ColV<T>> &col_variants<T>(int n) {             \
    static const std::vector<ColV<T>> empty;                                   \
    PFC_AMD_SWITCH_LEN(n, static const auto v = make_col<T, N_>(); return v;)  \
    return empty;                                                              \
  }                                                                            \
  template <> const std::vector<ZV<T>> &pfcz_variants<T>(int n) {              \
    static const std::vector<ZV<T>> empty;                                     \
    PFC_AMD_SWITCH_LEN(n, static const auto v = make_z<T, N_>(); return v;)    \
    return empty;                                                              \
  }                                                                            \
  template <> const std::vector<XV<T>> &xr2c_variants<T>(int n) {              \
    static const std::vector<XV<T>> empty;                                     \
    PFC_AMD_SWITCH_LEN(n, static const auto v = make_xr2c<T, N_>(); return v;) \
    return empty;                                                              \
  }                                                                            \
  template <> const std::vector<XV<T>> &xc2r_variants<T>(int n) {              \
    static const std::vector<XV<T>> empty;                                     \
    PFC_AMD_SWITCH_LEN(n, static const auto v = make_xc2r<T, N_>(); return v;) \
    return empty;                                                              \
  }                                                                            \
  template <> const std::vector<PXV<T>> &pfcx_variants<T>(int n) {             \
    static const std::vector<PXV<T>> empty;                                    \
    PFC_AMD_SWITCH_LEN(n, static const auto v = make_pfcx<T, N_>(); return v;) \
    return empty;                                                              \
  }                                                                            \
  template <> const std::vector<TXV<T>> &tungx_variants<T>(int n) {            \
    static const std::vector<TXV<T>> empty;                                    \
    PFC_AMD_SWITCH_LEN(n, static const auto v = make_tungx<T, N_>(); return v;) \
    return empty;                                                              \
  }

#define PFC_AMD_DEFINE_T_LISTS(T)                                              \
  template <> const std::vector<TV<T>> &t_variants<T>(int kind, int n) {       \
    static const std::vector<TV<T>> empty;                                     \
    if (kind < 0 || kind >= TK_COUNT) return empty;                            \
    PFC_AMD_SWITCH_LEN(n, static const std::vector<TV<T>> v[TK_COUNT] = {      \
        make_t_kind<T, N_, 0>(), make_t_kind<T, N_, 1>(),                      \
        make_t_kind<T, N_, 2>(), make_t_kind<T, N_, 3>(),                      \
        make_t_kind<T, N_, 4>(), make_t_kind<T, N_, 5>(),                      \
        make_t_kind<T, N_, 6>(), make_t_kind<T, N_, 7>()};                     \
                          return v[kind];)                                     \
    return empty;                                                              \
  }

} // namespace pfc_amd
