// field_modifier.hpp -- FieldModifier and the initial conditions on the path.
//
// FieldModifier::apply(Model&, double) (field_modifier.hpp:251-452),
// Constant (initial_conditions/constant.hpp:82-89), SingleSeed
// (initial_conditions/single_seed.hpp:51-83, radius 64.0 physical units,
// writes 0.0 outside), the rotated Seed helper (initial_conditions/seed.hpp),
// SeedGrid (seed_grid.hpp), RandomSeeds (random_seeds.hpp), FixedBC
// (boundary_conditions/fixed_bc.hpp) and MovingBC
// (boundary_conditions/moving_bc.hpp).  UniformNoise has no reference
// counterpart: it adds lo + (hi-lo)*u_i with the counter-based splitmix64
// stream of the device IC pfc_hip_ic_add_noise, so host and device runs start
// from the same field.
#pragma once

#include "model.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <limits>
#include <random>
#include <string>
#include <vector>

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

// Seed (initial_conditions/seed.hpp:39-143): a spherical bcc one-mode seed
// at `location` whose six <110>/sqrt(2) wave vectors are turned by
// R = yaw(o[0]) * pitch(o[1]) * roll(o[2]) (Euler angles about z, y, x).
// Inside |X - location| < radius the value is rho + sum_i 2*amp*cos(q_i . X)
// with X the absolute coordinate, as the reference evaluates it.
class Seed {
public:
  using vec3 = std::array<double, 3>;

  Seed(const vec3 &location, const vec3 &orientation, double radius, double rho,
       double amplitude)
      : m_loc(location), m_q(rotated(orientation)), m_rho(rho), m_radius(radius),
        m_amp(amplitude) {}

  bool is_inside(const vec3 &X) const {
    const double x = X[0] - m_loc[0], y = X[1] - m_loc[1], z = X[2] - m_loc[2];
    return x * x + y * y + z * z < m_radius * m_radius;
  }
  double get_value(const vec3 &X) const {
    double u = m_rho;
    for (const vec3 &q : m_q) u += 2.0 * m_amp * std::cos((q[0] * X[0]) + (q[1] * X[1]) + (q[2] * X[2]));
    return u;
  }
  const vec3 &get_location() const noexcept { return m_loc; }
  const std::array<vec3, 6> &get_q() const noexcept { return m_q; }
  double get_radius() const noexcept { return m_radius; }
  double get_rho() const noexcept { return m_rho; }
  double get_amplitude() const noexcept { return m_amp; }

  // The 24-double row of pfc_hip_ic_seeds: location, q_1..q_6, rho, radius,
  // amplitude.
  void to_row(double *out) const {
    for (int d = 0; d < 3; ++d) out[d] = m_loc[d];
    for (int i = 0; i < 6; ++i)
      for (int d = 0; d < 3; ++d) out[3 + 3 * i + d] = m_q[i][d];
    out[21] = m_rho;
    out[22] = m_radius;
    out[23] = m_amp;
  }

private:
  using mat3 = std::array<vec3, 3>;
  // 3x3 products summed from 0.0 over the inner index in order, as the
  // reference's mult3 does (so the rotated vectors round the same way)
  static mat3 mul(const mat3 &A, const mat3 &B) {
    mat3 C{};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double c = 0.0;
        for (int k = 0; k < 3; ++k) c += A[i][k] * B[k][j];
        C[i][j] = c;
      }
    return C;
  }
  static std::array<vec3, 6> rotated(const vec3 &o) {
    const double ca = std::cos(o[0]), sa = std::sin(o[0]);
    const double cb = std::cos(o[1]), sb = std::sin(o[1]);
    const double cc = std::cos(o[2]), sc = std::sin(o[2]);
    const mat3 yaw{vec3{ca, -sa, 0.0}, vec3{sa, ca, 0.0}, vec3{0.0, 0.0, 1.0}};
    const mat3 pitch{vec3{cb, 0.0, sb}, vec3{0.0, 1.0, 0.0}, vec3{-sb, 0.0, cb}};
    const mat3 roll{vec3{1.0, 0.0, 0.0}, vec3{0.0, cc, -sc}, vec3{0.0, sc, cc}};
    const mat3 R = mul(yaw, mul(pitch, roll));
    const double s = 1.0 / std::sqrt(2.0);
    const std::array<vec3, 6> q0{vec3{s, s, 0}, vec3{s, 0, s},  vec3{0, s, s},
                                 vec3{s, 0, -s}, vec3{s, -s, 0}, vec3{0, s, -s}};
    std::array<vec3, 6> q{};
    for (int n = 0; n < 6; ++n)
      for (int i = 0; i < 3; ++i) {
        double c = 0.0;
        for (int j = 0; j < 3; ++j) c += R[i][j] * q0[n][j];
        q[n][i] = c;
      }
    return q;
  }

  vec3 m_loc;
  std::array<vec3, 6> m_q;
  double m_rho, m_radius, m_amp;
};

namespace detail {
// field of a seed list: the first seed holding X gives the value, 0.0 outside
inline void apply_seeds(Model &m, const std::string &field, const std::vector<Seed> &seeds) {
  field::apply(m.get_real_field(field), m.get_world(), m.get_fft(), [&seeds](const Real3 &X) {
    for (const Seed &sd : seeds)
      if (sd.is_inside(X)) return sd.get_value(X);
    return 0.0;
  });
}
inline bool apply_seeds_device(Model &m, const std::string &field, const std::vector<Seed> &seeds,
                               const char *who) {
  DeviceField d;
  if (!m.get_device_field(field, d)) return false;
  std::vector<double> rows(24 * seeds.size());
  for (size_t i = 0; i < seeds.size(); ++i) seeds[i].to_row(rows.data() + 24 * i);
  const World &w = m.get_world();
  check_abi(pfc_hip_ic_seeds(d.plan, d.ptr, w.origin.data(), w.spacing.data(), rows.data(),
                             (int)seeds.size(), 0, d.stream),
            who);
  m.device_field_modified(field);
  return true;
}
} // namespace detail

// SeedGrid (initial_conditions/seed_grid.hpp:46-123): Ny x Nz seeds on a
// regular (y, z) grid at x = X0, each jittered by U(-0.2 radius, 0.2 radius)
// per axis and turned by three U(0, 2 pi) angles, all drawn from
// std::mt19937_64(42) in the reference's order; 0.0 outside every seed.
class SeedGrid : public FieldModifier {
public:
  SeedGrid() = default;
  SeedGrid(int Ny, int Nz, double X0, double radius)
      : m_Ny(Ny), m_Nz(Nz), m_X0(X0), m_radius(radius) {}
  void set_Nx(int Nx) { m_Nx = Nx; }
  void set_Ny(int Ny) { m_Ny = Ny; }
  void set_Nz(int Nz) { m_Nz = Nz; }
  void set_X0(double X0) { m_X0 = X0; }
  void set_radius(double radius) { m_radius = radius; }
  void set_density(double rho) { m_rho = rho; }
  void set_amplitude(double amplitude) { m_amp = amplitude; }
  int get_Nx() const { return m_Nx; }
  int get_Ny() const { return m_Ny; }
  int get_Nz() const { return m_Nz; }
  double get_X0() const { return m_X0; }
  double get_radius() const { return m_radius; }
  double get_density() const { return m_rho; }
  double get_amplitude() const { return m_amp; }

  std::vector<Seed> seeds(const World &w) const {
    const double Dy = w.spacing[1] * w.size[1] / m_Ny;
    const double Dz = w.spacing[2] * w.size[2] / m_Nz;
    const double Y0 = Dy / 2.0, Z0 = Dz / 2.0;
    std::mt19937_64 re(42);
    std::uniform_real_distribution<double> rt(-0.2 * m_radius, 0.2 * m_radius);
    std::uniform_real_distribution<double> rr(0.0, 8.0 * std::atan(1.0));
    std::vector<Seed> out;
    for (int j = 0; j < m_Ny; j++)
      for (int k = 0; k < m_Nz; k++) {
        // braced lists evaluate left to right: x, y, z jitter, then 3 angles
        const Seed::vec3 loc{m_X0 + rt(re), Y0 + (Dy * j) + rt(re), Z0 + (Dz * k) + rt(re)};
        const Seed::vec3 ori{rr(re), rr(re), rr(re)};
        out.emplace_back(loc, ori, m_radius, m_rho, m_amp);
      }
    return out;
  }
  void apply(Model &m, double) override {
    detail::apply_seeds(m, get_field_name(), seeds(m.get_world()));
  }
  bool apply_device(Model &m, double) override {
    return detail::apply_seeds_device(m, get_field_name(), seeds(m.get_world()), "SeedGrid");
  }

private:
  int m_Nx = 1, m_Ny = 2, m_Nz = 2;
  double m_X0 = 0.0, m_radius = 0.0, m_rho = 0.0, m_amp = 0.0;
};

// RandomSeeds (initial_conditions/random_seeds.hpp:46-95): 150 seeds of
// radius 20 at x ~ U(-108, -68), y, z ~ U(-128, 128), each turned by three
// U(0, 2 pi) angles, drawn from std::mt19937_64(42); 0.0 outside every seed.
class RandomSeeds : public FieldModifier {
public:
  void set_amplitude(double amplitude) { m_amp = amplitude; }
  double get_amplitude() const { return m_amp; }
  void set_density(double density) { m_rho = density; }
  double get_density() const { return m_rho; }

  std::vector<Seed> seeds() const {
    const int nseeds = 150;
    const double radius = 20.0;
    std::mt19937_64 re(42);
    std::uniform_real_distribution<double> rx(-128.0 + radius, -128.0 + (3 * radius));
    std::uniform_real_distribution<double> ry(-128.0, 128.0), rz(-128.0, 128.0);
    std::uniform_real_distribution<double> ro(0.0, 8.0 * std::atan(1.0));
    std::vector<Seed> out;
    for (int i = 0; i < nseeds; i++) {
      const Seed::vec3 loc{rx(re), ry(re), rz(re)};
      const Seed::vec3 ori{ro(re), ro(re), ro(re)};
      out.emplace_back(loc, ori, radius, m_rho, m_amp);
    }
    return out;
  }
  void apply(Model &m, double) override { detail::apply_seeds(m, get_field_name(), seeds()); }
  bool apply_device(Model &m, double) override {
    return detail::apply_seeds_device(m, get_field_name(), seeds(), "RandomSeeds");
  }

private:
  double m_rho = 0.0, m_amp = 0.0;
};

// MovingBC (boundary_conditions/moving_bc.hpp:55-219) at one rank: the
// column maxima over (y, z) locate the front (first call: the last column
// above threshold; later calls advance while the column ahead stays above
// it, at most one period), the boundary sits disp ahead of it and never
// moves back, and a sigmoid band of half-width xwidth round
// fmod(xpos, Lx*dx) (wrapping round the periodic box) is written.
class MovingBC : public FieldModifier {
public:
  MovingBC() = default;
  MovingBC(double rho_low, double rho_high) : m_rho_low(rho_low), m_rho_high(rho_high) {}
  void set_rho_low(double v) { m_rho_low = v; }
  void set_rho_high(double v) { m_rho_high = v; }
  void set_xpos(double v) { m_xpos = v; }
  double get_xpos() const { return m_xpos; }
  void set_xwidth(double v) { m_xwidth = v; }
  double get_xwidth() const { return m_xwidth; }
  void set_alpha(double v) { m_alpha = v; }
  void set_disp(double v) { m_disp = v; }
  void set_threshold(double v) { m_threshold = v; }
  double get_threshold() const { return m_threshold; }
  int get_index() const { return m_idx; }

  void apply(Model &m, double) override {
    const World &w = m.get_world();
    RealField &f = m.get_real_field(get_field_name());
    const Box3i in = m.get_fft().get_inbox_bounds();
    std::vector<double> xline((size_t)w.size[0], std::numeric_limits<double>::lowest());
    size_t idx = 0;
    for (int k = in.low[2]; k <= in.high[2]; ++k)
      for (int j = in.low[1]; j <= in.high[1]; ++j)
        for (int i = in.low[0]; i <= in.high[0]; ++i, ++idx)
          xline[(size_t)i] = std::max(xline[(size_t)i], f[idx]);
    advance(w, xline);
    const double l = (double)w.size[0] * w.spacing[0];
    const double xpos = std::fmod(m_xpos, l);
    idx = 0;
    for (int k = in.low[2]; k <= in.high[2]; ++k)
      for (int j = in.low[1]; j <= in.high[1]; ++j)
        for (int i = in.low[0]; i <= in.high[0]; ++i, ++idx) {
          const double dist = domain::to_coords(w, {i, j, k})[0] - xpos;
          double d;
          if (std::abs(dist) < m_xwidth)
            d = dist;
          else if (xpos < m_xwidth && std::abs(dist - l) < m_xwidth)
            d = dist - l;
          else if (xpos > l - m_xwidth && std::abs(dist + l) < m_xwidth)
            d = dist + l;
          else
            continue;
          const double S = 1.0 / (1.0 + std::exp(-m_alpha * d));
          f[idx] = (m_rho_low * S) + (m_rho_high * (1.0 - S));
        }
  }

  bool apply_device(Model &m, double) override {
    DeviceField d;
    if (!m.get_device_field(get_field_name(), d)) return false;
    const World &w = m.get_world();
    std::vector<double> xline((size_t)w.size[0]);
    pfc::detail::check_abi(pfc_hip_column_max(d.plan, d.ptr, xline.data(), d.stream), "MovingBC");
    advance(w, xline);
    pfc::detail::check_abi(pfc_hip_bc_moving(d.plan, d.ptr, w.origin[0], w.spacing[0], m_xpos,
                                             m_rho_low, m_rho_high, m_xwidth, m_alpha, d.stream),
                           "MovingBC");
    m.device_field_modified(get_field_name());
    return true;
  }

private:
  void advance(const World &w, const std::vector<double> &xline) {
    const int Lx = w.size[0];
    if (m_first) {
      for (int i = Lx - 1; i >= 0; i--)
        if (xline[(size_t)i] > m_threshold) {
          m_idx = i;
          break;
        }
    } else {
      // at most one period, so a field above threshold everywhere ends
      int scanned = 0;
      while (xline[(size_t)(m_idx % Lx)] > m_threshold && scanned < Lx) {
        m_idx += 1;
        scanned += 1;
      }
    }
    m_xpos = std::max(w.origin[0] + (m_idx * w.spacing[0]) + m_disp, m_xpos);
    m_first = false;
  }

  double m_rho_low = 0.0, m_rho_high = 0.0;
  double m_xwidth = 15.0, m_alpha = 1.0, m_xpos = 0.0, m_threshold = 0.1, m_disp = 40.0;
  int m_idx = 0;
  bool m_first = true;
};

} // namespace pfc
