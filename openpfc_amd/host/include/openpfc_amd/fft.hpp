// fft.hpp -- the drop-in FFT boundary: pfc::fft::IFFT and its MI355X
// implementation HipFft over the C ABI of include/pfc_hip.h.
//
// IFFT is the reference's interface verbatim in meaning
// (include/openpfc/kernel/fft/fft_interface.hpp:40-83).  HipFft replaces
// FFT_Impl<heffte::backend::rocfft> (fft_heffte_backend.hpp:130-275,
// runtime/hip/fft_hip.hpp:56) at one rank:
//   - DataBuffer<HipTag, T> overloads (T = double or float) are the fast
//     path: device pointers go straight to the library;
//   - std::vector overloads stage through the plan's device buffers, so a
//     host-side caller of IFFT keeps working unchanged (the reference's GPU
//     FFT throws there; staging is the drop-in behaviour the CPU callers of
//     IFFT expect);
//   - buffer extents are checked first and a mismatch throws
//     std::invalid_argument with the reference's message format
//     ("FFT_Impl::forward: real buffer size A != size_inbox() B");
//   - every call is synchronous on return, like the reference (the
//     library enqueues on the plan's stream and HipFft waits for it);
//   - get_fft_time() accumulates wall time of the calls (the reference's
//     MPI_Wtime meter).
#pragma once

#include "core.hpp"
#include "decomposition.hpp"
#include "world.hpp"

#include <chrono>
#include <complex>
#include <cstdint>
#include <memory>
#include <type_traits>
#include <string>
#include <vector>

namespace pfc {

namespace fft {

using RealVector = std::vector<double>;
using ComplexVector = std::vector<std::complex<double>>;

enum class Backend : std::uint8_t { FFTW, CUDA, HIP };

struct IFFT {
  virtual ~IFFT() = default;
  virtual void forward(const RealVector &in, ComplexVector &out) = 0;
  virtual void backward(const ComplexVector &in, RealVector &out) = 0;
  virtual void reset_fft_time() = 0;
  virtual double get_fft_time() const = 0;
  virtual size_t size_inbox() const = 0;
  virtual size_t size_outbox() const = 0;
  virtual size_t size_workspace() const = 0;
  virtual size_t get_allocated_memory_bytes() const = 0;
  [[nodiscard]] virtual Box3i get_inbox_bounds() const = 0;
  [[nodiscard]] virtual Box3i get_outbox_bounds() const = 0;
};

[[nodiscard]] inline Box3i get_inbox(const IFFT &f) noexcept { return f.get_inbox_bounds(); }
[[nodiscard]] inline Box3i get_outbox(const IFFT &f) noexcept { return f.get_outbox_bounds(); }
inline void reset_fft_time(IFFT &f) noexcept { f.reset_fft_time(); }
[[nodiscard]] inline double get_fft_time(const IFFT &f) noexcept { return f.get_fft_time(); }

namespace detail {
inline void require_equal_size(std::size_t actual, std::size_t expected, const char *context,
                               const char *expected_label) {
  if (actual != expected)
    throw std::invalid_argument(std::string(context) + std::to_string(actual) + " != " +
                                expected_label + "() " + std::to_string(expected));
}
} // namespace detail

class HipFft final : public IFFT {
public:
  // precision: PFC_HIP_F64 or PFC_HIP_F32 (the DataBuffer element type must match)
  explicit HipFft(const World &w, int precision = PFC_HIP_F64, hipStream_t stream = nullptr)
      : m_world(w), m_stream(stream) {
    pfc_hip_plan_t p = nullptr;
    pfc::detail::check_abi(
        pfc_hip_plan_create(w.size[0], w.size[1], w.size[2], precision, 0, &p),
        "HipFft: pfc_hip_plan_create");
    m_plan.reset(p);
    m_prec = precision;
    size_t a = 0, b = 0, c = 0;
    pfc::detail::check_abi(pfc_hip_sizes(p, &a, &b, &c), "HipFft: pfc_hip_sizes");
    m_inbox = a;
    m_outbox = b;
    m_workspace = c;
  }

  // -- device-resident overloads (the fast path) ------------------------------
  template <typename T>
  void forward(const core::DataBuffer<backend::HipTag, T> &in,
               core::DataBuffer<backend::HipTag, std::complex<T>> &out) {
    check_prec<T>("forward");
    detail::require_equal_size(in.size(), size_inbox(), "FFT_Impl::forward: real buffer size ",
                               "size_inbox");
    detail::require_equal_size(out.size(), size_outbox(),
                               "FFT_Impl::forward: complex buffer size ", "size_outbox");
    timed([&] {
      pfc::detail::check_abi(pfc_hip_r2c(m_plan.get(), in.data(), out.data(), m_stream),
                             "HipFft::forward");
    });
  }
  template <typename T>
  void backward(const core::DataBuffer<backend::HipTag, std::complex<T>> &in,
                core::DataBuffer<backend::HipTag, T> &out) {
    check_prec<T>("backward");
    detail::require_equal_size(in.size(), size_outbox(),
                               "FFT_Impl::backward: complex buffer size ", "size_outbox");
    detail::require_equal_size(out.size(), size_inbox(), "FFT_Impl::backward: real buffer size ",
                               "size_inbox");
    timed([&] {
      pfc::detail::check_abi(pfc_hip_c2r(m_plan.get(), in.data(), out.data(), m_stream),
                             "HipFft::backward");
    });
  }

  // -- host-vector overloads (IFFT): staged through device buffers -----------
  void forward(const RealVector &in, ComplexVector &out) override {
    detail::require_equal_size(in.size(), size_inbox(), "FFT_Impl::forward: real buffer size ",
                               "size_inbox");
    detail::require_equal_size(out.size(), size_outbox(),
                               "FFT_Impl::forward: complex buffer size ", "size_outbox");
    require_f64("forward(std::vector)");
    stage();
    m_real.copy_from_host(in);
    forward(m_real, m_cplx);
    out = m_cplx.to_host();
  }
  void backward(const ComplexVector &in, RealVector &out) override {
    detail::require_equal_size(in.size(), size_outbox(),
                               "FFT_Impl::backward: complex buffer size ", "size_outbox");
    detail::require_equal_size(out.size(), size_inbox(), "FFT_Impl::backward: real buffer size ",
                               "size_inbox");
    require_f64("backward(std::vector)");
    stage();
    m_cplx.copy_from_host(in);
    backward(m_cplx, m_real);
    out = m_real.to_host();
  }

  void reset_fft_time() override { m_fft_time = 0.0; }
  double get_fft_time() const override { return m_fft_time; }
  size_t size_inbox() const override { return m_inbox; }
  size_t size_outbox() const override { return m_outbox; }
  size_t size_workspace() const override { return m_workspace; }
  size_t get_allocated_memory_bytes() const override {
    size_t b = 0;
    pfc::detail::check_abi(pfc_hip_allocated_bytes(m_plan.get(), &b),
                           "HipFft::get_allocated_memory_bytes");
    return b + (m_real.size() * sizeof(double)) +
           (m_cplx.size() * sizeof(std::complex<double>));
  }
  Box3i get_inbox_bounds() const override {
    const Int3 n = m_world.size;
    return Box3i{{0, 0, 0}, {n[0] - 1, n[1] - 1, n[2] - 1}, n};
  }
  Box3i get_outbox_bounds() const override {
    const Int3 n = m_world.size;
    return Box3i{{0, 0, 0}, {n[0] / 2, n[1] - 1, n[2] - 1}, {n[0] / 2 + 1, n[1], n[2]}};
  }

  pfc_hip_plan_t plan() const noexcept { return m_plan.get(); }
  hipStream_t stream() const noexcept { return m_stream; }
  int precision() const noexcept { return m_prec; }
  const World &world() const noexcept { return m_world; }

private:
  struct PlanDeleter {
    void operator()(pfc_hip_plan_t p) const noexcept { pfc_hip_plan_destroy(p); }
  };

  template <typename T> void check_prec(const char *what) const {
    const int want = std::is_same_v<T, double> ? PFC_HIP_F64 : PFC_HIP_F32;
    if (want != m_prec)
      throw std::runtime_error(std::string("HipFft::") + what +
                               ": buffer precision does not match the plan precision");
  }
  void require_f64(const char *what) const {
    if (m_prec != PFC_HIP_F64)
      throw std::runtime_error(std::string("HipFft::") + what +
                               ": host vectors are double; the plan is single precision");
  }
  void stage() {
    if (m_real.size() != m_inbox) m_real = core::DataBuffer<backend::HipTag, double>(m_inbox);
    if (m_cplx.size() != m_outbox)
      m_cplx = core::DataBuffer<backend::HipTag, std::complex<double>>(m_outbox);
  }
  template <typename F> void timed(F &&f) {
    const auto t0 = std::chrono::steady_clock::now();
    f();
    pfc::detail::check_hip(hipStreamSynchronize(m_stream), "HipFft: stream synchronize");
    m_fft_time += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }

  World m_world;
  hipStream_t m_stream = nullptr;
  int m_prec = PFC_HIP_F64;
  std::unique_ptr<pfc_hip_plan_s, PlanDeleter> m_plan;
  size_t m_inbox = 0, m_outbox = 0, m_workspace = 0;
  double m_fft_time = 0.0;
  core::DataBuffer<backend::HipTag, double> m_real;
  core::DataBuffer<backend::HipTag, std::complex<double>> m_cplx;
};

// Factory of the HIP backend (runtime/hip/fft_hip.cpp:19-37 create_hip) at
// one rank.
inline HipFft create_hip(const World &w, int precision = PFC_HIP_F64) {
  return HipFft(w, precision);
}

namespace detail {
// The HipFft behind an IFFT& (device-resident models need the library plan).
inline HipFft &as_hip_fft(IFFT &f, const char *who) {
  if (auto *h = dynamic_cast<HipFft *>(&f)) return *h;
  throw std::invalid_argument(std::string(who) + ": the FFT must be a pfc::fft::HipFft "
                              "(fft::create_hip or create_with_backend(..., Backend::HIP))");
}

// The one part a plan of this library covers: part rank_id of a one-part
// decomposition on a one-rank communicator.
inline void require_single_part(const Decomposition &d, int rank_id, Comm comm,
                                const char *who) {
  pfc::detail::require_one_rank(comm, who);
  if (decomposition::get_num_domains(d) != 1)
    throw std::logic_error(std::string(who) + ": the decomposition has " +
                           std::to_string(decomposition::get_num_domains(d)) +
                           " parts; this library plans one part per GPU");
  if (rank_id != 0)
    throw std::invalid_argument(std::string(who) + ": rank_id " + std::to_string(rank_id) +
                                " outside a one-part decomposition");
}
} // namespace detail

// create_hip(const Decomposition&, int rank_id, MPI_Comm)
// (runtime/hip/fft_hip.hpp:65-66): the HIP FFT of part rank_id.
[[nodiscard]] inline HipFft create_hip(const Decomposition &d, int rank_id,
                                       MPI_Comm comm = MPI_COMM_WORLD) {
  detail::require_single_part(d, rank_id, comm, "fft::create_hip");
  return HipFft(decomposition::get_world(d));
}

// create_hip(const Decomposition&, MPI_Comm) (fft_hip.hpp:77-78): the rank
// is taken from the communicator, whose size must match the decomposition.
[[nodiscard]] inline HipFft create_hip(const Decomposition &d, MPI_Comm comm = MPI_COMM_WORLD) {
  return create_hip(d, 0, comm);
}

// create_with_backend (src/openpfc/runtime/cpu/fft.cpp:131-191): this library
// provides the HIP backend; FFTW and CUDA requests are rejected as the
// reference rejects backends it was not built with.
[[nodiscard]] inline std::unique_ptr<IFFT> create_with_backend(const World &w, Backend b) {
  if (b != Backend::HIP) throw std::runtime_error("Unsupported FFT backend requested");
  return std::make_unique<HipFft>(w);
}

// create_with_backend(const FFTLayout&, int rank_id, const plan_options&,
// Backend, MPI_Comm) (kernel/fft/fft.hpp:52-55; fft.cpp:131-166): the real and
// complex boxes of part rank_id must be the whole grid and its half spectrum.
[[nodiscard]] inline std::unique_ptr<IFFT>
create_with_backend(const FFTLayout &layout, int rank_id, const plan_options &options,
                    Backend backend, MPI_Comm comm = MPI_COMM_WORLD) {
  (void)options; // reshape choices of a distributed plan
  if (backend != Backend::HIP) throw std::runtime_error("Unsupported FFT backend requested");
  detail::require_single_part(layout.m_decomposition, rank_id, comm,
                              "fft::create_with_backend");
  return std::make_unique<HipFft>(decomposition::get_world(layout.m_decomposition));
}

// create_with_backend(const Decomposition&, int rank_id, Backend, MPI_Comm)
// (kernel/fft/fft.hpp:56-57; fft.cpp:168-191): layout with r2c_direction 0
// and the backend's default options.
[[nodiscard]] inline std::unique_ptr<IFFT>
create_with_backend(const Decomposition &d, int rank_id, Backend backend,
                    MPI_Comm comm = MPI_COMM_WORLD) {
  if (backend != Backend::HIP) throw std::runtime_error("Unsupported FFT backend requested");
  return create_with_backend(layout::create(d, 0), rank_id, default_options(), backend, comm);
}

// Backend names of the JSON configs ("hip" / "rocm"; runtime/common/backend_from_string.hpp:24-40).
inline Backend backend_from_string(const std::string &s) {
  if (s == "hip" || s == "rocm" || s == "HIP" || s == "ROCM") return Backend::HIP;
  if (s == "fftw" || s == "FFTW") return Backend::FFTW;
  if (s == "cuda" || s == "CUDA") return Backend::CUDA;
  throw std::invalid_argument("unknown FFT backend '" + s + "'");
}

} // namespace fft
} // namespace pfc
