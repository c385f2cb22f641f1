// time.hpp -- fixed-step simulation clock with the reference's semantics
// (include/openpfc/kernel/simulation/time.hpp:366-395 validation,
// :618-619 get_current, :732-734 done with tolerance 1e-9, :751-754 next,
// :972-978 do_save).
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <stdexcept>
#include <string>

namespace pfc {

class Time {
public:
  Time(const std::array<double, 3> &time, double saveat)
      : m_t0(time[0]), m_t1(time[1]), m_dt(time[2]), m_accepted(time[0]), m_saveat(saveat) {
    if (m_t0 < 0) throw std::invalid_argument("Time: t0 must be >= 0, got " + std::to_string(m_t0));
    if (!(m_dt > 0)) throw std::invalid_argument("Time: dt must be > 0");
    if (!(m_t1 > m_t0))
      throw std::invalid_argument("Time: t1 must be > t0, got t0=" + std::to_string(m_t0) +
                                  ", t1=" + std::to_string(m_t1));
    if ((m_t1 - m_t0) < 1e-9) throw std::invalid_argument("Time: interval too small");
    if (m_saveat > m_t1)
      throw std::invalid_argument("Time: saveat " + std::to_string(m_saveat) + " > " +
                                  std::to_string(m_t1));
  }
  explicit Time(const std::array<double, 3> &time) : Time(time, time[2]) {}

  double get_t0() const { return m_t0; }
  double get_t1() const { return m_t1; }
  double get_dt() const { return m_dt; }
  double get_saveat() const { return m_saveat; }
  int get_increment() const { return m_increment; }
  double get_current() const { return (m_accepted > m_t1) ? m_t1 : m_accepted; }
  bool done() const { return get_current() >= m_t1 - 1e-9; }
  void next() {
    m_accepted = std::min(m_accepted + m_dt, m_t1);
    m_increment += 1;
  }
  bool do_save() const {
    if (m_saveat <= 0) return false;
    return (std::fmod(get_current() + 1.0e-9, m_saveat) < 1.e-6) || done() || (m_increment == 0);
  }
  operator double() const { return get_current(); }
  // time.hpp:654-661: restart at a given step count
  void set_increment(int increment) {
    if (increment < 0)
      throw std::invalid_argument("Time increment cannot be negative: " +
                                  std::to_string(increment));
    m_increment = increment;
    m_accepted = std::min(m_t0 + static_cast<double>(increment) * m_dt, m_t1);
  }

private:
  double m_t0, m_t1, m_dt;
  int m_increment = 0;
  double m_accepted;
  double m_saveat;
};

} // namespace pfc
