// simulator.hpp -- the time loop (include/openpfc/kernel/simulation/simulator.hpp:549-553,
// simulator_integrator.hpp:43-61): on the first call initial conditions and
// boundary conditions are applied, then every call does time.next(),
// boundary conditions, model.step(time.get_current()).  A field modifier
// with a device path runs on the model's device field; the others run
// between the model's prepare/finalize residency hooks, so device-resident
// models see host edits.  Results writers get the result counter
// (simulator_integrator.hpp:31-36: file number, incremented per save).
#pragma once

#include "field_modifier.hpp"
#include "model.hpp"
#include "time.hpp"

#include <functional>
#include <memory>
#include <vector>

namespace pfc {

class Simulator {
public:
  // comm: the reference's defaulted MPI_Comm (simulator.hpp:152); one rank here.
  Simulator(Model &m, Time &t, MPI_Comm comm = MPI_COMM_WORLD)
      : m_model(m), m_time(t), m_comm(comm) {
    pfc::detail::require_one_rank(comm, "Simulator");
  }
  void set_mpi_comm(MPI_Comm comm) {
    pfc::detail::require_one_rank(comm, "Simulator");
    m_comm = comm;
  }
  [[nodiscard]] MPI_Comm mpi_comm() const noexcept { return m_comm; }

  void add_initial_conditions(std::unique_ptr<FieldModifier> ic) { m_ics.push_back(std::move(ic)); }
  void add_boundary_conditions(std::unique_ptr<FieldModifier> bc) { m_bcs.push_back(std::move(bc)); }
  // Called with the model whenever Time::do_save() holds (results writers).
  void add_results_writer(std::function<void(Model &, const Time &)> w) {
    m_writers.push_back([w = std::move(w)](Model &m, const Time &t, int) { w(m, t); });
  }
  void add_results_writer(std::function<void(Model &, const Time &, int)> w) {
    m_writers.push_back(std::move(w));
  }
  void set_result_counter(int n) { m_result_counter = n; }
  int get_result_counter() const { return m_result_counter; }
  void apply_initial_conditions() { apply(m_ics); }

  void step() {
    if (m_time.get_increment() == 0) {
      apply(m_ics);
      apply(m_bcs);
      if (m_time.do_save()) write();
    }
    m_time.next();
    apply(m_bcs);
    m_model.step(m_time.get_current());
    if (m_time.do_save()) write();
  }

  // Up to max_steps steps with the effect of that many step() calls (no
  // reference counterpart: the reference loops over step()).  Without
  // boundary conditions nothing happens between two steps up to the next
  // save point, so those steps go to Model::step_many in one call; returns
  // the number of steps taken.
  int step_block(int max_steps) {
    if (max_steps <= 1 || m_time.get_increment() == 0 || !m_bcs.empty()) {
      step();
      return 1;
    }
    Time look = m_time;
    std::vector<double> times;
    while ((int)times.size() < max_steps && !look.done()) {
      look.next();
      times.push_back(look.get_current());
      if (look.do_save()) break;
    }
    if (times.empty()) return 0;
    m_model.step_many(times.data(), (int)times.size());
    for (size_t i = 0; i < times.size(); ++i) m_time.next();
    if (m_time.do_save()) write();
    return (int)times.size();
  }

  bool done() const { return m_time.done(); }
  Model &get_model() { return m_model; }
  Time &get_time() { return m_time; }

private:
  void apply(std::vector<std::unique_ptr<FieldModifier>> &mods) {
    bool host_ready = false, host_dirty = false;
    for (auto &f : mods) {
      if (host_dirty) {
        // host edits go to the device before any device-side modifier
        DeviceField d;
        if (m_model.get_device_field(f->get_field_name(), d)) {
          m_model.finalize_after_field_modifiers();
          host_dirty = false;
          host_ready = false;
        }
      }
      if (!host_dirty && f->apply_device(m_model, m_time.get_current())) {
        host_ready = false;
        continue;
      }
      if (!host_ready) {
        m_model.prepare_for_field_modifiers();
        host_ready = true;
      }
      f->apply(m_model, m_time.get_current());
      host_dirty = true;
    }
    if (host_dirty) m_model.finalize_after_field_modifiers();
  }
  void write() {
    if (m_writers.empty()) return;
    for (auto &w : m_writers) w(m_model, m_time, m_result_counter);
    ++m_result_counter;
  }

  Model &m_model;
  Time &m_time;
  MPI_Comm m_comm;
  std::vector<std::unique_ptr<FieldModifier>> m_ics, m_bcs;
  std::vector<std::function<void(Model &, const Time &, int)>> m_writers;
  int m_result_counter = 0;
};

inline bool done(const Simulator &s) { return const_cast<Simulator &>(s).done(); }
inline void step(Simulator &s) { s.step(); }
inline Time &get_time(Simulator &s) { return s.get_time(); }

} // namespace pfc
