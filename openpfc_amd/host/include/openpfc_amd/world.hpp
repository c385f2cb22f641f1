// world.hpp -- grid geometry: size, origin and spacing with the reference's
// coordinate rule x = origin + i*spacing (include/openpfc/kernel/data/domain.hpp:185-189)
// and the k-space helpers of include/openpfc/kernel/fft/kspace.hpp:97-131.
#pragma once

#include <array>
#include <stdexcept>
#include <string>

namespace pfc {

using Int3 = std::array<int, 3>;
using Real3 = std::array<double, 3>;

struct GridSize { Int3 value; explicit GridSize(Int3 v) : value(v) {} };
struct PhysicalOrigin { Real3 value; explicit PhysicalOrigin(Real3 v) : value(v) {} };
struct GridSpacing { Real3 value; explicit GridSpacing(Real3 v) : value(v) {} };

// Index box {low, high, size} (inclusive bounds), the reference's Box3i
// (include/openpfc/kernel/data/box3i.hpp).
struct Box3i {
  Int3 low{0, 0, 0};
  Int3 high{0, 0, 0};
  Int3 size{0, 0, 0};
};

struct World {
  Int3 size{1, 1, 1};
  Real3 origin{0, 0, 0};
  Real3 spacing{1, 1, 1};
};
using Domain = World;

namespace domain {
inline World create(GridSize n, PhysicalOrigin o, GridSpacing d) {
  for (int i = 0; i < 3; ++i) {
    if (n.value[i] < 1) throw std::invalid_argument("domain::create: size must be >= 1");
    if (!(d.value[i] > 0)) throw std::invalid_argument("domain::create: spacing must be > 0");
  }
  return World{n.value, o.value, d.value};
}
// origin "center" of the JSON loader (from_json_world_time.hpp:180-184)
inline World create_centered(GridSize n, GridSpacing d) {
  Real3 o{};
  for (int i = 0; i < 3; ++i) o[i] = -0.5 * d.value[i] * n.value[i];
  return create(n, PhysicalOrigin(o), d);
}
inline Real3 to_coords(const World &w, const Int3 &idx) noexcept {
  Real3 x;
  for (int i = 0; i < 3; ++i) x[i] = w.origin[i] + idx[i] * w.spacing[i];
  return x;
}
inline const Int3 &get_size(const World &w) noexcept { return w.size; }
inline const Real3 &get_spacing(const World &w) noexcept { return w.spacing; }
} // namespace domain

namespace world = domain;

namespace fft::kspace {
constexpr double two_pi = 6.283185307179586;
inline std::array<double, 3> k_frequency_scaling(const World &w) noexcept {
  return {two_pi / (w.spacing[0] * w.size[0]), two_pi / (w.spacing[1] * w.size[1]),
          two_pi / (w.spacing[2] * w.size[2])};
}
inline double k_component(int index, int size, double f) noexcept {
  return (index <= size / 2) ? index * f : (index - size) * f;
}
inline double k_laplacian_value(double ki, double kj, double kk) noexcept {
  return -((ki * ki) + (kj * kj) + (kk * kk));
}
} // namespace fft::kspace

} // namespace pfc
