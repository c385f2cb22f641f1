// comm.hpp -- the communicator argument of the reference's constructors and
// factories: Model(fft, world, MPI_Comm = MPI_COMM_WORLD)
// (include/openpfc/kernel/simulation/model.hpp:151), Simulator(model, time,
// MPI_Comm) (simulator.hpp:152), create_hip(decomposition, rank, MPI_Comm)
// (runtime/hip/fft_hip.hpp:65-78), create_with_backend(..., MPI_Comm)
// (kernel/fft/fft.hpp:52-57).
//
// This library runs one rank per GPU.  When <mpi.h> was included before this
// header (MPI_VERSION is defined), the real MPI_Comm is taken and a
// communicator of more than one rank is refused where a plan is made.
// Otherwise MPI_Comm / MPI_COMM_WORLD / MPI_COMM_SELF name a one-rank
// stand-in, so reference call sites compile unchanged without MPI (define
// OPENPFC_AMD_NO_MPI_NAMES to keep those global names free).
#pragma once

#include <stdexcept>
#include <string>

#if defined(MPI_VERSION)
namespace pfc {
using Comm = MPI_Comm;
// ranks of c (1 while MPI is not initialised)
inline int comm_size(Comm c) {
  int init = 0;
  MPI_Initialized(&init);
  if (!init) return 1;
  int n = 1;
  MPI_Comm_size(c, &n);
  return n;
}
} // namespace pfc
#else
namespace pfc {
// One-rank communicator stand-in (no MPI in this library).
struct Comm {
  int id = 0;
  friend bool operator==(Comm a, Comm b) { return a.id == b.id; }
};
inline int comm_size(Comm) { return 1; }
} // namespace pfc
#if !defined(OPENPFC_AMD_NO_MPI_NAMES)
using MPI_Comm = pfc::Comm;
inline constexpr pfc::Comm MPI_COMM_WORLD{0};
inline constexpr pfc::Comm MPI_COMM_SELF{1};
#endif
#endif

namespace pfc::detail {
// Plans and models of this library span one rank.
inline void require_one_rank(Comm c, const char *who) {
  const int n = comm_size(c);
  if (n != 1)
    throw std::logic_error(std::string(who) + ": this library runs one rank per GPU; the "
                           "communicator has " + std::to_string(n) + " ranks");
}
} // namespace pfc::detail
