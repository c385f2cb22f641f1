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

} // namespace pfc
