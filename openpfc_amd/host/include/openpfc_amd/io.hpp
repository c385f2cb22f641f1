// io.hpp -- raw binary field files: BinaryWriter, BinaryReader, FileReader.
//
// Drop-ins for the reference's BinaryWriter
// (include/openpfc/frontend/io/binary_writer.hpp:76-205), BinaryReader
// (include/openpfc/kernel/simulation/binary_reader.hpp:79-200) and the
// "from_file" initial condition FileReader
// (include/openpfc/kernel/simulation/initial_conditions/file_reader.hpp:48-92)
// for one process: set_domain(global, local, offset) describes the brick,
// write(increment, field) writes it into format_with_number(filename,
// increment), read(filename, field) fills it.  The file is the global array,
// x fastest, native f64 (complex: re, im), no header.  Host vectors
// (RealField, ComplexField) and device DataBuffers (staged through pinned
// chunks by the library) are both accepted.  Bad domains throw
// std::invalid_argument, size mismatches and file errors std::runtime_error.
#pragma once

#include "field_modifier.hpp"

#include <array>
#include <complex>
#include <regex>
#include <sstream>
#include <string>
#include <vector>

namespace pfc {

using ComplexField = std::vector<std::complex<double>>;

namespace utils {
// frontend/utils/utils.hpp:96-139: at most one %d conversion (flags and a
// numeric width allowed); no pattern leaves the name unchanged.
inline std::string format_with_number(const std::string &filename, int increment) {
  if (filename.find('%') == std::string::npos) return filename;
  static const std::regex valid(R"(^[^%]*%[-+ 0]*[0-9]*d[^%]*$)");
  if (!std::regex_match(filename, valid))
    throw std::invalid_argument("format_with_number: invalid filename pattern '" + filename +
                                "'. Must contain at most one % integer specifier "
                                "(%d with optional flags/width).");
  const int n = std::snprintf(nullptr, 0, filename.c_str(), increment);
  std::string out((size_t)n + 1, '\0');
  std::snprintf(out.data(), out.size(), filename.c_str(), increment);
  out.resize((size_t)n);
  return out;
}
} // namespace utils

namespace detail {
template <typename T> struct IoType;
template <> struct IoType<double> { static constexpr int prec = PFC_HIP_F64, kind = PFC_HIP_IO_REAL; };
template <> struct IoType<float> { static constexpr int prec = PFC_HIP_F32, kind = PFC_HIP_IO_REAL; };
template <> struct IoType<std::complex<double>> {
  static constexpr int prec = PFC_HIP_F64, kind = PFC_HIP_IO_COMPLEX;
};
template <> struct IoType<std::complex<float>> {
  static constexpr int prec = PFC_HIP_F32, kind = PFC_HIP_IO_COMPLEX;
};

class BrickIo {
public:
  void set_domain(const std::array<int, 3> &global, const std::array<int, 3> &local,
                  const std::array<int, 3> &offset) {
    // domain_geometry.hpp:36-60
    const std::string prefix = std::string(m_what) + "::set_domain: ";
    for (int i = 0; i < 3; ++i) {
      if (global[i] <= 0 || local[i] <= 0)
        throw std::invalid_argument(prefix + "global/local dimensions must be positive");
      if (offset[i] < 0) throw std::invalid_argument(prefix + "offsets must be non-negative");
      if ((long long)offset[i] + local[i] > global[i])
        throw std::invalid_argument(prefix + "piece does not lie inside global domain "
                                             "(check offset + local_size vs global)");
    }
    m_global = global;
    m_local = local;
    m_offset = offset;
    m_valid = true;
  }

protected:
  explicit BrickIo(const char *what) : m_what(what) {}
  void require(size_t n, const char *op) const {
    if (!m_valid)
      throw std::runtime_error(std::string(m_what) + "::" + op +
                               ": set_domain() was not called");
    const size_t expected = (size_t)m_local[0] * m_local[1] * m_local[2];
    if (n != expected) {
      std::ostringstream o;
      o << m_what << "::" << op << ": buffer size mismatch (expected " << expected
        << " elements from set_domain, got " << n << ")";
      throw std::runtime_error(o.str());
    }
  }
  const char *m_what;
  std::array<int, 3> m_global{}, m_local{}, m_offset{};
  bool m_valid = false;
};
} // namespace detail

class BinaryWriter : public detail::BrickIo {
public:
  explicit BinaryWriter(std::string filename)
      : BrickIo("BinaryWriter"), m_filename(std::move(filename)) {}

  template <typename T> std::string write(int increment, const std::vector<T> &data) {
    require(data.size(), "write");
    const std::string path = utils::format_with_number(m_filename, increment);
    pfc::detail::check_abi(
        pfc_hip_field_write_host(path.c_str(), data.data(), detail::IoType<T>::prec,
                                 detail::IoType<T>::kind, m_global.data(), m_local.data(),
                                 m_offset.data(), 1),
        "BinaryWriter::write");
    return path;
  }
  // device fast path: no host mirror of the field is needed
  template <typename T>
  std::string write(int increment, const core::DataBuffer<backend::HipTag, T> &data,
                    hipStream_t stream = nullptr) {
    require(data.size(), "write");
    const std::string path = utils::format_with_number(m_filename, increment);
    pfc::detail::check_abi(
        pfc_hip_field_write(path.c_str(), data.data(), detail::IoType<T>::prec,
                            detail::IoType<T>::kind, m_global.data(), m_local.data(),
                            m_offset.data(), 1, stream),
        "BinaryWriter::write");
    return path;
  }
  const std::string &get_filename() const { return m_filename; }

private:
  std::string m_filename;
};

class BinaryReader : public detail::BrickIo {
public:
  BinaryReader() : BrickIo("BinaryReader") {}

  template <typename T> void read(const std::string &filename, std::vector<T> &data) {
    require(data.size(), "read");
    pfc::detail::check_abi(
        pfc_hip_field_read_host(filename.c_str(), data.data(), detail::IoType<T>::prec,
                                detail::IoType<T>::kind, m_global.data(), m_local.data(),
                                m_offset.data()),
        "BinaryReader::read");
  }
  template <typename T>
  void read(const std::string &filename, core::DataBuffer<backend::HipTag, T> &data,
            hipStream_t stream = nullptr) {
    require(data.size(), "read");
    pfc::detail::check_abi(
        pfc_hip_field_read(filename.c_str(), data.data(), detail::IoType<T>::prec,
                           detail::IoType<T>::kind, m_global.data(), m_local.data(),
                           m_offset.data(), stream),
        "BinaryReader::read");
  }
};

// The "from_file" initial condition: reads the model's inbox brick of the
// global array into the named real field.
class FileReader : public FieldModifier {
public:
  FileReader() = default;
  explicit FileReader(std::string filename) : m_filename(std::move(filename)) {}
  void set_filename(std::string f) { m_filename = std::move(f); }
  const std::string &get_filename() const { return m_filename; }

  void apply(Model &m, double) override {
    const World &w = m.get_world();
    const auto inbox = fft::get_inbox(m.get_fft());
    RealField &f = m.get_real_field(get_field_name());
    try {
      BinaryReader r;
      r.set_domain({w.size[0], w.size[1], w.size[2]}, {inbox.size[0], inbox.size[1], inbox.size[2]},
                   {inbox.low[0], inbox.low[1], inbox.low[2]});
      r.read(m_filename, f);
    } catch (const std::exception &ex) {
      throw std::runtime_error("FileReader failed to read \"" + m_filename + "\" into field \"" +
                               get_field_name() + "\": " + ex.what());
    }
  }

  // device path: the inbox brick goes straight from the file to the model's
  // device field through the library's pinned staging
  bool apply_device(Model &m, double) override {
    DeviceField d;
    if (!m.get_device_field(get_field_name(), d)) return false;
    const World &w = m.get_world();
    const auto inbox = fft::get_inbox(m.get_fft());
    const int g[3] = {w.size[0], w.size[1], w.size[2]};
    const int l[3] = {inbox.size[0], inbox.size[1], inbox.size[2]};
    const int o[3] = {inbox.low[0], inbox.low[1], inbox.low[2]};
    const int rc = pfc_hip_field_read(m_filename.c_str(), d.ptr, d.precision, PFC_HIP_IO_REAL,
                                      g, l, o, d.stream);
    if (rc != PFC_HIP_OK)
      throw std::runtime_error("FileReader failed to read \"" + m_filename + "\" into field \"" +
                               get_field_name() + "\": " + pfc_hip_last_error());
    m.device_field_modified(get_field_name());
    return true;
  }

private:
  std::string m_filename;
};

} // namespace pfc
