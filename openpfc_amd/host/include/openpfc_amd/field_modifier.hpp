// field_modifier.hpp -- FieldModifier and the initial conditions on the path.
//
// FieldModifier::apply(Model&, double) (field_modifier.hpp:251-452),
// Constant (initial_conditions/constant.hpp:82-89), SingleSeed
// (initial_conditions/single_seed.hpp:51-83, radius 64.0 physical units,
// writes 0.0 outside).  UniformNoise has no reference counterpart: it adds
// lo + (hi-lo)*u_i with the counter-based splitmix64 stream of the device IC
// pfc_hip_ic_add_noise, so host and device runs start from the same field.
#pragma once

#include "model.hpp"

#include <array>
#include <cmath>
#include <cstdint>
#include <string>

namespace pfc {

class FieldModifier {
public:
  virtual ~FieldModifier() = default;
  virtual void apply(Model &m, double t) = 0;
  // Device path: applies the modifier to the model's device field when the
  // model has one and the modifier has a device implementation; returns
  // false (nothing done) otherwise.
  virtual bool apply_device(Model &, double) { return false; }
  void set_field_name(const std::string &n) { m_field = n; }
  const std::string &get_field_name() const noexcept { return m_field; }

private:
  std::string m_field = "default";
};

class Constant : public FieldModifier {
public:
  explicit Constant(double n0 = 0.0) : m_n0(n0) {}
  void apply(Model &m, double) override {
    const double n0 = m_n0;
    field::apply(m.get_real_field(get_field_name()), m.get_world(), m.get_fft(),
                 [n0](const Real3 &) { return n0; });
  }
  double get_density() const noexcept { return m_n0; }
  bool apply_device(Model &m, double) override {
    DeviceField d;
    if (!m.get_device_field(get_field_name(), d)) return false;
    pfc::detail::check_abi(pfc_hip_ic_constant(d.plan, d.ptr, m_n0, d.stream), "Constant");
    m.device_field_modified(get_field_name());
    return true;
  }

private:
  double m_n0;
};

class SingleSeed : public FieldModifier {
public:
  void set_amplitude(double a) { m_amp = a; }
  double get_amplitude() const { return m_amp; }
  void set_density(double d) { m_rho = d; }
  double get_density() const { return m_rho; }
  // The reference fixes the radius to 64.0; it is a parameter here (default 64).
  void set_radius(double r) { m_radius = r; }
  double get_radius() const { return m_radius; }

  void apply(Model &m, double) override {
    const double s = 1.0 / std::sqrt(2.0);
    const std::array<std::array<double, 3>, 6> q = {{{s, s, 0},
                                                     {s, 0, s},
                                                     {0, s, s},
                                                     {s, 0, -s},
                                                     {s, -s, 0},
                                                     {0, s, -s}}};
    const double r2 = std::pow(m_radius, 2);
    const double amp = m_amp, rho = m_rho;
    field::apply(m.get_real_field(get_field_name()), m.get_world(), m.get_fft(),
                 [=](const Real3 &X) {
                   const double x = X[0], y = X[1], z = X[2];
                   if (x * x + y * y + z * z >= r2) return 0.0;
                   double u = rho;
                   for (int qi = 0; qi < 6; ++qi)
                     u += 2.0 * amp * std::cos((q[qi][0] * x) + (q[qi][1] * y) + (q[qi][2] * z));
                   return u;
                 });
  }

  bool apply_device(Model &m, double) override {
    DeviceField d;
    if (!m.get_device_field(get_field_name(), d)) return false;
    const World &w = m.get_world();
    pfc::detail::check_abi(pfc_hip_ic_single_seed(d.plan, d.ptr, w.origin.data(),
                                                  w.spacing.data(), m_rho, m_amp, m_radius, 0,
                                                  d.stream),
                           "SingleSeed");
    m.device_field_modified(get_field_name());
    return true;
  }

private:
  double m_amp = 0.0, m_rho = 0.0, m_radius = 64.0;
};

class UniformNoise : public FieldModifier {
public:
  UniformNoise(std::uint64_t seed, double lo, double hi) : m_seed(seed), m_lo(lo), m_hi(hi) {}
  void apply(Model &m, double) override {
    RealField &f = m.get_real_field(get_field_name());
    const std::uint64_t s0 = mix(m_seed);
    const double w = m_hi - m_lo;
    for (size_t i = 0; i < f.size(); ++i) {
      const double u = static_cast<double>(mix(s0 + i) >> 11) * 0x1.0p-53;
      f[i] = f[i] + (m_lo + w * u);
    }
  }

  bool apply_device(Model &m, double) override {
    DeviceField d;
    if (!m.get_device_field(get_field_name(), d)) return false;
    pfc::detail::check_abi(pfc_hip_ic_add_noise(d.plan, d.ptr, m_seed, m_lo, m_hi, d.stream),
                           "UniformNoise");
    m.device_field_modified(get_field_name());
    return true;
  }

private:
  static std::uint64_t mix(std::uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  std::uint64_t m_seed;
  double m_lo, m_hi;
};

// FixedBC (include/openpfc/kernel/simulation/boundary_conditions/fixed_bc.hpp):
// a sigmoid from rho_high to rho_low across a band of half-width xwidth = 20
// around xpos = Lx*dx - xwidth, in world x coordinates.
class FixedBC : public FieldModifier {
public:
  FixedBC() = default;
  FixedBC(double rho_low, double rho_high) : m_rho_low(rho_low), m_rho_high(rho_high) {}
  void set_rho_low(double v) { m_rho_low = v; }
  void set_rho_high(double v) { m_rho_high = v; }
  double get_rho_low() const { return m_rho_low; }
  double get_rho_high() const { return m_rho_high; }

  void apply(Model &m, double) override {
    const World &w = m.get_world();
    const double xpos = ((double)w.size[0] * w.spacing[0]) - m_xwidth;
    RealField &f = m.get_real_field(get_field_name());
    const Box3i in = m.get_fft().get_inbox_bounds();
    size_t idx = 0;
    for (int k = in.low[2]; k <= in.high[2]; ++k)
      for (int j = in.low[1]; j <= in.high[1]; ++j)
        for (int i = in.low[0]; i <= in.high[0]; ++i, ++idx) {
          const double x = domain::to_coords(w, {i, j, k})[0];
          if (std::abs(x - xpos) < m_xwidth) {
            const double S = 1.0 / (1.0 + std::exp(-m_alpha * (x - xpos)));
            f[idx] = (m_rho_low * S) + (m_rho_high * (1.0 - S));
          }
        }
  }
  bool apply_device(Model &m, double) override {
    DeviceField d;
    if (!m.get_device_field(get_field_name(), d)) return false;
    const World &w = m.get_world();
    pfc::detail::check_abi(pfc_hip_bc_fixed(d.plan, d.ptr, w.origin[0], w.spacing[0],
                                            m_rho_low, m_rho_high, m_xwidth, m_alpha, d.stream),
                           "FixedBC");
    m.device_field_modified(get_field_name());
    return true;
  }

private:
  double m_xwidth = 20.0, m_alpha = 1.0;
  double m_rho_low = 0.0, m_rho_high = 0.0;
};

} // namespace pfc
