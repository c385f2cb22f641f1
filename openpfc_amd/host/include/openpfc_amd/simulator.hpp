// simulator.hpp -- the time loop (include/openpfc/kernel/simulation/simulator.hpp:549-553,
// simulator_integrator.hpp:43-61): on the first call initial conditions and
// boundary conditions are applied, then every call does time.next(),
// boundary conditions, model.step(time.get_current()).  Field modifiers run
// between the model's prepare/finalize residency hooks, so device-resident
// models see host edits.
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
  Simulator(Model &m, Time &t) : m_model(m), m_time(t) {}

  void add_initial_conditions(std::unique_ptr<FieldModifier> ic) { m_ics.push_back(std::move(ic)); }
  void add_boundary_conditions(std::unique_ptr<FieldModifier> bc) { m_bcs.push_back(std::move(bc)); }
  // Called with the model whenever Time::do_save() holds (results writers).
  void add_results_writer(std::function<void(Model &, const Time &)> w) {
    m_writers.push_back(std::move(w));
  }

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

  bool done() const { return m_time.done(); }
  Model &get_model() { return m_model; }
  Time &get_time() { return m_time; }

private:
  void apply(std::vector<std::unique_ptr<FieldModifier>> &mods) {
    if (mods.empty()) return;
    m_model.prepare_for_field_modifiers();
    for (auto &f : mods) f->apply(m_model, m_time.get_current());
    m_model.finalize_after_field_modifiers();
  }
  void write() {
    if (m_writers.empty()) return;
    m_model.prepare_for_field_modifiers();
    for (auto &w : m_writers) w(m_model, m_time);
  }

  Model &m_model;
  Time &m_time;
  std::vector<std::unique_ptr<FieldModifier>> m_ics, m_bcs;
  std::vector<std::function<void(Model &, const Time &)>> m_writers;
};

inline bool done(const Simulator &s) { return const_cast<Simulator &>(s).done(); }
inline void step(Simulator &s) { s.step(); }
inline Time &get_time(Simulator &s) { return s.get_time(); }

} // namespace pfc
