// tungsten_model_hip.hpp -- the Tungsten mean-field PFC model, device-resident.
//
// TungstenParams mirrors apps/tungsten/include/tungsten/common/
// tungsten_params.hpp:34-412 (base parameters with the tungsten_single_seed
// defaults, derived vapor-model coefficients computed on access in the
// reference's formulas).  TungstenHIP<RealType> is the drop-in for
// TungstenHIP (apps/tungsten/include/tungsten/hip/tungsten_model.hpp:57-330):
// initialize(dt) sets the operators (evaluated per mode on the device by the
// fused step), step() advances psi by one ETD1 step
//   psi_MF = F^-1(fMF F(psi)),
//   N = p3 psi^2 + p4 psi^3 + q3 psi_MF^2 + q4 psi_MF^3 - stabP psi,
//   psi <- F^-1(exp(L dt) F(psi) + n_weight F(N)),
// and get_psi()/get_psiMF() return the device buffers (psi_MF of the last
// step's input, as the reference holds after step()).  The registered host
// field "psi" is a mirror refreshed by the residency hooks.
#pragma once

#include "autotune.hpp"
#include "model.hpp"

#include <cmath>
#include <type_traits>

namespace pfc {

struct TungstenParams {
  // base parameters (tungsten_params.hpp:36-57)
  double n0 = -0.10, n_sol = -0.047, n_vap = -0.464, T = 3300.0, T0 = 156000.0;
  double Bx = 0.8582, alpha = 0.50, alpha_farTol = 0.001;
  int alpha_highOrd = 4;
  double lambda = 0.22, stabP = 0.2, shift_u = 0.3341, shift_s = 0.1898;
  double p2 = 1.0, p3 = -0.5, p4 = 0.333333333;
  double q20 = -0.0037, q21 = 1.0, q30 = -12.4567, q31 = 20.0, q40 = 45.0;

  // setters/getters of the reference's interface
  void set_n0(double v) { n0 = v; }
  void set_T(double v) { T = v; }
  void set_T0(double v) { T0 = v; }
  void set_Bx(double v) { Bx = v; }
  void set_alpha(double v) { alpha = v; }
  void set_alpha_farTol(double v) { alpha_farTol = v; }
  void set_alpha_highOrd(int v) { alpha_highOrd = v; }
  void set_lambda(double v) { lambda = v; }
  void set_stabP(double v) { stabP = v; }
  void set_shift_u(double v) { shift_u = v; }
  void set_shift_s(double v) { shift_s = v; }
  void set_p2(double v) { p2 = v; }
  void set_p3(double v) { p3 = v; }
  void set_p4(double v) { p4 = v; }
  void set_q20(double v) { q20 = v; }
  void set_q21(double v) { q21 = v; }
  void set_q30(double v) { q30 = v; }
  void set_q31(double v) { q31 = v; }
  void set_q40(double v) { q40 = v; }
  double get_n0() const { return n0; }
  double get_stabP() const { return stabP; }

  // derived (tungsten_params.hpp:339-411)
  double get_tau() const { return T / T0; }
  double get_p2_bar() const {
    return p2 + 2 * shift_s * p3 + 3 * std::pow(shift_s, 2) * p4;
  }
  double get_p3_bar() const { return shift_u * (p3 + 3 * shift_s * p4); }
  double get_p4_bar() const { return std::pow(shift_u, 2) * p4; }
  double get_q20_bar() const {
    return q20 + 2.0 * shift_s * q30 + 3.0 * std::pow(shift_s, 2) * q40;
  }
  double get_q21_bar() const { return q21 + 2.0 * shift_s * q31; }
  double get_q30_bar() const { return shift_u * (q30 + 3.0 * shift_s * q40); }
  double get_q31_bar() const { return shift_u * q31; }
  double get_q40_bar() const { return std::pow(shift_u, 2) * q40; }
  double get_q2_bar() const { return get_q21_bar() * get_tau() + get_q20_bar(); }
  double get_q3_bar() const { return get_q31_bar() * get_tau() + get_q30_bar(); }
  double get_q4_bar() const { return get_q40_bar(); }

  // The PFC_HIP_MODEL_TUNGSTEN parameter vector: spacing, the OperatorParams
  // of make_operator_params (tungsten_spectral.hpp:68-73), p3..q4 bars.
  void abi_params(const World &w, double (&out)[PFC_HIP_TUNGSTEN_NPARAMS]) const {
    const double v[PFC_HIP_TUNGSTEN_NPARAMS] = {
        w.spacing[0], w.spacing[1], w.spacing[2], 2.0 * alpha * alpha,
        2.0 * lambda * lambda, alpha_farTol, (double)alpha_highOrd, Bx, T, T0, stabP,
        get_p2_bar(), get_q2_bar(), get_p3_bar(), get_p4_bar(), get_q3_bar(),
        get_q4_bar()};
    for (int i = 0; i < PFC_HIP_TUNGSTEN_NPARAMS; ++i) out[i] = v[i];
  }
};

template <typename RealType = double> class TungstenHIP : public Model {
  static_assert(std::is_same_v<RealType, double> || std::is_same_v<RealType, float>,
                "TungstenHIP<double> or TungstenHIP<float>");

public:
  TungstenParams params;

  // TungstenHIP(fft::IFFT&, const World&, MPI_Comm = MPI_COMM_WORLD)
  // (apps/tungsten/include/tungsten/hip/tungsten_model.hpp:131-136); fft must
  // be a HipFft of RealType's precision.
  explicit TungstenHIP(fft::IFFT &fft, const World &world, MPI_Comm comm = MPI_COMM_WORLD)
      : Model(fft, world, comm), m_hfft(fft::detail::as_hip_fft(fft, "TungstenHIP")) {
    const int want = std::is_same_v<RealType, double> ? PFC_HIP_F64 : PFC_HIP_F32;
    if (m_hfft.precision() != want)
      throw std::invalid_argument("TungstenHIP: the HipFft plan precision does not "
                                  "match RealType");
  }

  void initialize(double dt) override {
    const size_t n = m_hfft.size_inbox();
    m_psi_cpu.assign(n, 0.0);
    m_psi = core::DataBuffer<backend::HipTag, RealType>(n);
    m_psiMF = core::DataBuffer<backend::HipTag, RealType>(n);
    add_real_field("psi", m_psi_cpu);
    double p[PFC_HIP_TUNGSTEN_NPARAMS];
    params.abi_params(get_world(), p);
    pfc::detail::check_abi(pfc_hip_pfc_set_operators(m_hfft.plan(), PFC_HIP_MODEL_TUNGSTEN,
                                                     p, PFC_HIP_TUNGSTEN_NPARAMS, dt),
                           "TungstenHIP::initialize");
    finalize_after_field_modifiers();
    // step variants from the autotune cache, or tuned now (autotune.hpp)
    m_tuned = gpu::tune_step(
        m_hfft.plan(), m_psi.data(),
        gpu::step_key("tungsten", std::is_same_v<RealType, double> ? PFC_HIP_F64 : PFC_HIP_F32,
                      get_world().size, false),
        m_hfft.stream());
  }

  // How initialize() chose the step's launch variants.
  gpu::TuneOutcome autotune_outcome() const noexcept { return m_tuned; }

  void step(double) override { step_n(1); }
  void step_many(const double *, int n) override { step_n(n); }

  // advance n steps without returning to the host in between
  void step_n(int n) {
    pfc::detail::check_abi(pfc_hip_tungsten_step(m_hfft.plan(), m_psi.data(), n,
                                                 m_psiMF.data(), m_hfft.stream()),
                           "TungstenHIP::step");
    m_host_stale = true;
  }

  void prepare_for_field_modifiers() override {
    if (m_host_stale) {
      pfc::detail::check_hip(hipStreamSynchronize(m_hfft.stream()), "TungstenHIP: sync");
      const auto h = m_psi.to_host();
      m_psi_cpu.assign(h.begin(), h.end());
      m_host_stale = false;
    }
  }
  void finalize_after_field_modifiers() override {
    std::vector<RealType> tmp(m_psi_cpu.begin(), m_psi_cpu.end());
    m_psi.copy_from_host(tmp);
    m_host_stale = false;
  }

  bool get_device_field(const std::string &name, DeviceField &d) override {
    if (name != "psi" && name != "psiMF") return false;
    d.ptr = name == "psi" ? m_psi.data() : m_psiMF.data();
    d.precision = std::is_same_v<RealType, double> ? PFC_HIP_F64 : PFC_HIP_F32;
    d.plan = m_hfft.plan();
    d.stream = m_hfft.stream();
    return true;
  }
  void device_field_modified(const std::string &name) override {
    if (name == "psi") m_host_stale = true;
  }

  core::DataBuffer<backend::HipTag, RealType> &get_psi() { return m_psi; }
  core::DataBuffer<backend::HipTag, RealType> &get_psiMF() { return m_psiMF; }
  RealField &get_psi_for_writer() {
    prepare_for_field_modifiers();
    return m_psi_cpu;
  }

  size_t get_allocated_memory_bytes() const override {
    return m_psi_cpu.size() * sizeof(double) + 2 * m_psi.size() * sizeof(RealType);
  }

private:
  fft::HipFft &m_hfft;
  RealField m_psi_cpu;
  core::DataBuffer<backend::HipTag, RealType> m_psi;
  core::DataBuffer<backend::HipTag, RealType> m_psiMF;
  bool m_host_stale = false;
  gpu::TuneOutcome m_tuned = gpu::TuneOutcome::defaults;
};

} // namespace pfc
