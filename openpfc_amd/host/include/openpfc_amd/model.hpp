// model.hpp -- pfc::Model: the base class spectral models derive from.
//
// Mirrors include/openpfc/kernel/simulation/model.hpp:151-613 for the parts
// the spectral step uses: construction from an FFT and a World (the FFT must
// outlive the model, model.hpp:100-109), virtual initialize(dt) / step(t),
// the field-modifier residency hooks (model.hpp:265,272), a registry of
// named real fields held by reference, and get_allocated_memory_bytes().
#pragma once

#include "fft.hpp"

#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace pfc {

using RealField = std::vector<double>;

// A field held in device memory by a device-resident model, with the plan
// and stream that work on it (field modifiers and writers use it to stay on
// the device).
struct DeviceField {
  void *ptr = nullptr;
  int precision = PFC_HIP_F64;
  pfc_hip_plan_t plan = nullptr;
  hipStream_t stream = nullptr;
};

class Model {
public:
  // comm: the reference's defaulted MPI_Comm (model.hpp:151); one rank here.
  Model(fft::IFFT &fft, const World &world, MPI_Comm comm = MPI_COMM_WORLD)
      : m_fft(fft), m_world(world), m_comm(comm) {
    pfc::detail::require_one_rank(comm, "Model");
  }
  virtual ~Model() = default;
  Model(const Model &) = delete;
  Model &operator=(const Model &) = delete;

  virtual void initialize(double dt) = 0;
  virtual void step(double t) = 0;
  // n steps ending at times[0..n-1] with nothing in between (no modifier,
  // no writer): a device model may keep the field on the device and fuse
  // the step boundaries.  Default: n calls of step().
  virtual void step_many(const double *times, int n) {
    for (int i = 0; i < n; ++i) step(times[i]);
  }
  // Device-resident models copy device -> host here so modifiers and
  // writers see current data, and host -> device in finalize.
  virtual void prepare_for_field_modifiers() {}
  virtual void finalize_after_field_modifiers() {}
  virtual size_t get_allocated_memory_bytes() const { return 0; }
  // Device-resident models: the device copy of a named field (false when the
  // field lives on the host only).  After editing it on the device, callers
  // report the edit so the host mirror is refreshed before host access.
  virtual bool get_device_field(const std::string &, DeviceField &) { return false; }
  virtual void device_field_modified(const std::string &) {}

  [[nodiscard]] MPI_Comm mpi_comm() const noexcept { return m_comm; }
  fft::IFFT &get_fft() noexcept { return m_fft; }
  const fft::IFFT &get_fft() const noexcept { return m_fft; }
  const World &get_world() const noexcept { return m_world; }

  void add_real_field(const std::string &name, RealField &f) { m_real[name] = &f; }
  bool has_real_field(const std::string &name) const { return m_real.count(name) > 0; }
  RealField &get_real_field(const std::string &name) {
    auto it = m_real.find(name);
    if (it == m_real.end())
      throw std::invalid_argument("Model: no real field named '" + name + "'");
    return *it->second;
  }

private:
  fft::IFFT &m_fft;
  World m_world;
  MPI_Comm m_comm;
  std::map<std::string, RealField *> m_real;
};

// Free-function API of the reference (model_free_functions.hpp).
inline void initialize(Model &m, double dt) { m.initialize(dt); }
inline void step(Model &m, double t) { m.step(t); }
inline void add_real_field(Model &m, const std::string &n, RealField &f) { m.add_real_field(n, f); }
inline RealField &get_real_field(Model &m, const std::string &n) { return m.get_real_field(n); }
inline fft::IFFT &get_fft(Model &m) { return m.get_fft(); }
inline const World &get_world(const Model &m) { return m.get_world(); }
inline const World &get_domain(const Model &m) { return m.get_world(); }

namespace field {
// field::apply (include/openpfc/kernel/field/operations.hpp:81-105): fn(x)
// over the inbox, k outer, i inner, x = origin + i*spacing.
template <typename Fn> void apply(RealField &f, const World &w, const fft::IFFT &fft, Fn &&fn) {
  const Box3i in = fft.get_inbox_bounds();
  const size_t expected = (size_t)in.size[0] * in.size[1] * in.size[2];
  if (f.size() != expected)
    throw std::invalid_argument("field::apply: field size does not match FFT inbox size");
  size_t idx = 0;
  for (int k = in.low[2]; k <= in.high[2]; ++k)
    for (int j = in.low[1]; j <= in.high[1]; ++j)
      for (int i = in.low[0]; i <= in.high[0]; ++i)
        f[idx++] = static_cast<double>(fn(domain::to_coords(w, {i, j, k})));
}
} // namespace field

} // namespace pfc
