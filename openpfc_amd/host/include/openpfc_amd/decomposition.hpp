// decomposition.hpp -- the domain decomposition and FFT layout types the
// reference's FFT factories take, at one rank.
//
// Decomposition (include/openpfc/kernel/decomposition/decomposition.hpp:149-
// 353): the global World and a grid of parts with their index boxes.  This
// library plans one rank per GPU, so a decomposition of one part is the one a
// plan accepts; create(world, nparts) splits along z for nparts > 1 (slabs,
// the layout a multi-GPU step would shard in) so that such a decomposition
// can be built and inspected, and the factories refuse it.
//
// FFTLayout (kernel/fft/fft_layout.hpp:22-42; src/openpfc/runtime/cpu/
// fft.cpp:58-81): the real and complex boxes per part for r2c_direction 0
// (x halved: the complex box of a part spans kx in [0, nx/2]).
//
// plan_options stands in for heffte::plan_options (the HeFFTe type of the
// reference's create_with_backend): its fields choose the distributed
// reshape algorithm, which a one-rank plan does not use.
#pragma once

#include "comm.hpp"
#include "world.hpp"

#include <array>
#include <stdexcept>
#include <string>
#include <vector>

namespace pfc {

namespace decomposition {

struct Decomposition {
  World m_global_world;
  std::array<int, 3> m_grid;
  std::vector<Box3i> m_local_boxes;

  Decomposition(const World &world, const Int3 &grid)
      : m_global_world(world), m_grid(grid), m_local_boxes(split(world.size, grid)) {}

  // The index boxes of an extent split by a grid of parts, x fastest over
  // the parts; part i of n points over p parts starts at floor(n*i/p).
  static std::vector<Box3i> split(const Int3 &size, const Int3 &grid) {
    for (int d = 0; d < 3; ++d)
      if (grid[d] < 1 || grid[d] > size[d])
        throw std::invalid_argument("Decomposition: grid must be in [1, size] per axis");
    auto at = [](int n, int p, int i) { return (int)((long)n * i / p); };
    std::vector<Box3i> out;
    for (int k = 0; k < grid[2]; ++k)
      for (int j = 0; j < grid[1]; ++j)
        for (int i = 0; i < grid[0]; ++i) {
          Box3i b;
          const int idx[3] = {i, j, k};
          for (int d = 0; d < 3; ++d) {
            b.low[d] = at(size[d], grid[d], idx[d]);
            b.high[d] = at(size[d], grid[d], idx[d] + 1) - 1;
            b.size[d] = b.high[d] - b.low[d] + 1;
          }
          out.push_back(b);
        }
    return out;
  }
};

[[nodiscard]] inline Decomposition create(const World &world, const Int3 &grid) {
  return Decomposition(world, grid);
}
// nparts parts: one part is the whole grid; more parts are z slabs.
[[nodiscard]] inline Decomposition create(const World &world, const int &nparts) {
  if (nparts < 1) throw std::invalid_argument("decomposition::create: nparts must be >= 1");
  return Decomposition(world, Int3{1, 1, nparts});
}
[[nodiscard]] inline const World &get_world(const Decomposition &d) noexcept {
  return d.m_global_world;
}
[[nodiscard]] inline const std::array<int, 3> &get_grid(const Decomposition &d) noexcept {
  return d.m_grid;
}
[[nodiscard]] inline int get_num_domains(const Decomposition &d) noexcept {
  return d.m_grid[0] * d.m_grid[1] * d.m_grid[2];
}
[[nodiscard]] inline const Box3i &get_local_box(const Decomposition &d, int i) {
  return d.m_local_boxes.at((size_t)i);
}

} // namespace decomposition

using Decomposition = decomposition::Decomposition;

namespace fft {

// heffte::plan_options: the distributed reshape choices (ignored at one rank).
struct plan_options {
  bool use_reorder = true;
  bool use_pencils = true;
  bool use_gpu_aware = true;
  int algorithm = 0; // 0 alltoallv, 1 alltoall, 2 p2p_plined, 3 p2p
};
[[nodiscard]] inline plan_options default_options() { return {}; }

namespace layout {

struct FFTLayout {
  const Decomposition m_decomposition;
  const int m_r2c_direction = 0;
  const std::vector<Box3i> m_real_boxes;
  const std::vector<Box3i> m_complex_boxes;
};

// layout::create (src/openpfc/runtime/cpu/fft.cpp:58-81): the real index
// box [0, N-1]^3 and the complex box [0, N1/2] x [0, N2-1] x [0, N3-1] are
// each split by the decomposition's grid.  Only r2c_direction 0 (x halved)
// is accepted: the device transforms halve x.
[[nodiscard]] inline FFTLayout create(const Decomposition &d, int r2c_direction) {
  if (r2c_direction != 0)
    throw std::invalid_argument("fft::layout::create: r2c_direction must be 0 (x halved)");
  const Int3 &n = d.m_global_world.size;
  return FFTLayout{d, r2c_direction, Decomposition::split(n, d.m_grid),
                   Decomposition::split(Int3{n[0] / 2 + 1, n[1], n[2]}, d.m_grid)};
}
[[nodiscard]] inline const Box3i &get_real_box(const FFTLayout &l, int i) {
  return l.m_real_boxes.at((size_t)i);
}
[[nodiscard]] inline const Box3i &get_complex_box(const FFTLayout &l, int i) {
  return l.m_complex_boxes.at((size_t)i);
}
[[nodiscard]] inline int get_r2c_direction(const FFTLayout &l) { return l.m_r2c_direction; }

} // namespace layout
} // namespace fft

using FFTLayout = fft::layout::FFTLayout;

} // namespace pfc
