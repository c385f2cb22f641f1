// core.hpp -- error mapping, backend tags and the device DataBuffer.
//
// DataBuffer<HipTag, T> mirrors the reference's RAII move-only device buffer
// (include/openpfc/runtime/gpu/databuffer_gpu.hpp:82-229): size(), data(),
// copy_from_host(), to_host(); allocation failures and HIP errors become
// std::runtime_error carrying the HIP error string.
#pragma once

#include <pfc_hip.h>

#include <hip/hip_runtime.h>

#include <complex>
#include <cstddef>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace pfc {

namespace backend {
struct CpuTag {};
struct HipTag {};
} // namespace backend

namespace detail {

// Map a C-ABI status to the reference's exception types:
// bad arguments -> std::invalid_argument, everything else -> std::runtime_error.
inline void check_abi(int rc, const char *what) {
  if (rc == PFC_HIP_OK) return;
  const std::string msg = std::string(what) + ": " + pfc_hip_last_error();
  if (rc == PFC_HIP_EINVAL) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

inline void check_hip(hipError_t e, const char *what) {
  if (e != hipSuccess)
    throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

} // namespace detail

namespace core {

template <typename Tag, typename T> class DataBuffer;

template <typename T> class DataBuffer<backend::HipTag, T> {
public:
  DataBuffer() = default;
  explicit DataBuffer(std::size_t n) : m_size(n) {
    if (n) {
      hipError_t e = hipMalloc(reinterpret_cast<void **>(&m_ptr), n * sizeof(T));
      if (e != hipSuccess) {
        (void)hipGetLastError(); // clear the sticky allocation error
        m_ptr = nullptr;
        m_size = 0;
        throw std::runtime_error(std::string("DataBuffer<HipTag>: hipMalloc of ") +
                                 std::to_string(n * sizeof(T)) +
                                 " bytes failed: " + hipGetErrorString(e));
      }
    }
  }
  DataBuffer(const DataBuffer &) = delete;
  DataBuffer &operator=(const DataBuffer &) = delete;
  DataBuffer(DataBuffer &&o) noexcept : m_ptr(o.m_ptr), m_size(o.m_size) {
    o.m_ptr = nullptr;
    o.m_size = 0;
  }
  DataBuffer &operator=(DataBuffer &&o) noexcept {
    if (this != &o) {
      release();
      std::swap(m_ptr, o.m_ptr);
      std::swap(m_size, o.m_size);
    }
    return *this;
  }
  ~DataBuffer() { release(); }

  T *data() noexcept { return m_ptr; }
  const T *data() const noexcept { return m_ptr; }
  std::size_t size() const noexcept { return m_size; }

  void copy_from_host(const std::vector<T> &h) {
    if (h.size() != m_size)
      throw std::invalid_argument("DataBuffer::copy_from_host: size " +
                                  std::to_string(h.size()) + " != " + std::to_string(m_size));
    if (m_size)
      detail::check_hip(hipMemcpy(m_ptr, h.data(), m_size * sizeof(T), hipMemcpyHostToDevice),
                        "DataBuffer::copy_from_host");
  }
  std::vector<T> to_host() const {
    std::vector<T> h(m_size);
    if (m_size)
      detail::check_hip(hipMemcpy(h.data(), m_ptr, m_size * sizeof(T), hipMemcpyDeviceToHost),
                        "DataBuffer::to_host");
    return h;
  }

private:
  void release() noexcept {
    if (m_ptr) (void)hipFree(m_ptr);
    m_ptr = nullptr;
    m_size = 0;
  }
  T *m_ptr = nullptr;
  std::size_t m_size = 0;
};

} // namespace core
} // namespace pfc
