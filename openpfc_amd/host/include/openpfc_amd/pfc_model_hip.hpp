// pfc_model_hip.hpp -- the plain conserved PFC model, device-resident.
//
// Drop-in for PFCModel of docs/api/examples/10_complete_pfc_model.cpp:62-214:
// same constructor arguments (FFT, domain, epsilon), initialize(dt) builds
// opL = exp(L dt), opN = (opL-1)/L*(-k^2) with L = -k^2(eps + (1-k^2)^2)
// (evaluated per mode on the device) and picks the step's launch variants
// (autotune cache, or tuned on the spot), step() advances psi by one ETD1
// step psi_hat <- opL*F(psi) + opN*F(psi^3), psi <- F^-1(psi_hat).
// psi lives on the device; the registered host field "density" (and
// "default") is a mirror refreshed by the residency hooks, so field
// modifiers, writers and get_real_field() keep their reference meaning.
#pragma once

#include "autotune.hpp"
#include "model.hpp"

namespace pfc {

class PFCModelHIP : public Model {
public:
  // fft must be a HipFft (any IFFT& the HIP factories return)
  PFCModelHIP(fft::IFFT &fft, const World &world, double epsilon,
              MPI_Comm comm = MPI_COMM_WORLD)
      : Model(fft, world, comm), m_hfft(fft::detail::as_hip_fft(fft, "PFCModelHIP")),
        m_epsilon(epsilon) {}

  void initialize(double dt) override {
    m_dt = dt;
    m_density.assign(m_hfft.size_inbox(), 0.0);
    m_psi = core::DataBuffer<backend::HipTag, double>(m_hfft.size_inbox());
    add_real_field("density", m_density);
    add_real_field("default", m_density);
    const World &w = get_world();
    const double params[4] = {w.spacing[0], w.spacing[1], w.spacing[2], m_epsilon};
    pfc::detail::check_abi(
        pfc_hip_pfc_set_operators(m_hfft.plan(), PFC_HIP_MODEL_PFC, params, 4, dt),
        "PFCModelHIP::initialize");
    m_host_stale = false;
    m_psi.copy_from_host(m_density);
    // step variants from the autotune cache, or tuned now (autotune.hpp)
    m_tuned = gpu::tune_step(m_hfft.plan(), m_psi.data(),
                             gpu::step_key("pfc", PFC_HIP_F64, w.size, false), m_hfft.stream());
  }

  // How initialize() chose the step's launch variants.
  gpu::TuneOutcome autotune_outcome() const noexcept { return m_tuned; }

  void step(double) override {
    pfc::detail::check_abi(pfc_hip_pfc_step(m_hfft.plan(), m_psi.data(), 1, m_hfft.stream()),
                           "PFCModelHIP::step");
    m_host_stale = true;
  }

  // the step does not depend on t: one call, fused step boundaries
  void step_many(const double *, int n) override { step_n(n); }

  // advance n steps without returning to the host in between
  void step_n(int n) {
    pfc::detail::check_abi(pfc_hip_pfc_step(m_hfft.plan(), m_psi.data(), n, m_hfft.stream()),
                           "PFCModelHIP::step_n");
    m_host_stale = true;
  }

  void prepare_for_field_modifiers() override {
    if (m_host_stale) {
      pfc::detail::check_hip(hipStreamSynchronize(m_hfft.stream()), "PFCModelHIP: sync");
      m_density = m_psi.to_host();
      m_host_stale = false;
    }
  }
  void finalize_after_field_modifiers() override { m_psi.copy_from_host(m_density); }

  size_t get_allocated_memory_bytes() const override {
    return m_density.size() * sizeof(double) + m_psi.size() * sizeof(double);
  }

  bool get_device_field(const std::string &name, DeviceField &d) override {
    if (name != "density" && name != "default") return false;
    d.ptr = m_psi.data();
    d.precision = PFC_HIP_F64;
    d.plan = m_hfft.plan();
    d.stream = m_hfft.stream();
    return true;
  }
  void device_field_modified(const std::string &) override { m_host_stale = true; }

  // Free energy F = dV sum [psi (eps + (1 + lap)^2) psi / 2 + psi^4 / 4] of
  // the current field, summed in double on the device (the reference model
  // has no such method; the functional is the one its step descends).
  double free_energy() {
    double F = 0;
    pfc::detail::check_abi(pfc_hip_pfc_free_energy(m_hfft.plan(), m_psi.data(), &F,
                                                   m_hfft.stream()),
                           "PFCModelHIP::free_energy");
    return F;
  }

  double get_epsilon() const { return m_epsilon; }
  core::DataBuffer<backend::HipTag, double> &device_density() { return m_psi; }

private:
  fft::HipFft &m_hfft;
  double m_epsilon;
  double m_dt = 0;
  RealField m_density;
  core::DataBuffer<backend::HipTag, double> m_psi;
  bool m_host_stale = false;
  gpu::TuneOutcome m_tuned = gpu::TuneOutcome::defaults;
};

} // namespace pfc
