// variants.hpp -- launch variants of every pass, selectable per plan.
//
// A variant fixes the compile-time shape of one pass: the radix plan of the
// line transform (radices and values per thread E), the tile width L of a
// column pass or the rows per block of a row pass, and whether the LDS
// exchange moves both planes at once or one plane at a time.  Index 0 of
// each list is the default; the autotuner times the others.
#pragma once

#include "fft_kernels.hpp"
#include "step_kernels_t.hpp"

#include <string>
#include <vector>

namespace pfc_amd {

enum Role {
  ROLE_X_R2C = 0, // x forward of the plain R2C (nx/2-point lines)
  ROLE_X_C2R = 1, // x inverse (nx/2-point lines), plain C2R and the step
  ROLE_COL_Y = 2, // y column passes (plain transforms and the step)
  ROLE_COL_Z = 3, // z column passes of the plain transforms
  ROLE_PFC_X = 4, // step forward x (nx-point lines, psi and psi^3)
  ROLE_PFC_Z = 5, // step z pass (forward, combine, inverse)
  ROLE_STEP_LAYOUT = 6, // 0: transposed spectra (T1/T2), 1: natural layout
  ROLE_T_X_FWD = 7,     // transposed step: psi -> T1
  ROLE_T_Y_FWD = 8,     // T1 -> T2
  ROLE_T_Z = 9,         // T2 z pass
  ROLE_T_Y_INV = 10,    // T2 -> T1
  ROLE_T_X_INV = 11,    // T1 -> psi
  ROLE_X_TUNG = 12,     // Tungsten natural step: fused inverse x, N, forward x
  ROLE_T_X_R2C = 13,    // Tungsten transposed step: psi -> T1 (psi only)
  ROLE_T_X_TUNG = 14,   // Tungsten transposed step: fused x pass on T1
  ROLE_T_X_FUSE = 15,   // PFC multi-step run: inverse x of one step fused with the next forward x
  ROLE_STEP_FUSE = 16,  // 0: multi-step runs as single steps, 1: fused step boundaries
  ROLE_COUNT = 17
};

// Kinds of transposed-step kernels (one variant list per kind and length).
enum TKind {
  TK_X_FWD = 0, TK_Y_FWD = 1, TK_Z = 2, TK_Y_INV = 3, TK_X_INV = 4,
  TK_X_R2C = 5, TK_X_TUNG = 6, TK_X_FUSE = 7, TK_COUNT = 8
};

template <typename T> struct ColV {
  std::string name;
  int plan_id; // index of the line plan (selects the twiddle table)
  hipError_t (*fn)(int dir, ColArgs<T> a, long n_outer, int nfields,
                   hipStream_t s);
};
template <typename T> struct ZV {
  std::string name;
  int plan_id;
  hipError_t (*fn)(PfcZArgs<T> a, long ny, hipStream_t s);
};
template <typename T> struct XV {
  std::string name;
  int plan_id;
  hipError_t (*fn)(XArgs<T> a, hipStream_t s);
};
template <typename T> struct PXV {
  std::string name;
  int plan_id;
  hipError_t (*fn)(PfcXArgs<T> a, hipStream_t s);
};

template <typename T> struct TXV {
  std::string name;
  int plan_id;
  hipError_t (*fn)(TungXArgs<T> a, hipStream_t s);
};

template <typename T> struct TV {
  std::string name;
  int plan_id;
  hipError_t (*fn)(TStepArgs<T> a, hipStream_t s);
  int half = 0; // 1: the kernel runs the n/2-point plan `plan_id` of a length-n role
  int plan_id2 = -1; // >= 0: also the 2n-point plan plan_id2 (its twiddles in tw_x)
};

// Variant lists for a line length n (empty when n is unsupported).
template <typename T> const std::vector<ColV<T>> &col_variants(int n);
template <typename T> const std::vector<ZV<T>> &pfcz_variants(int n);
template <typename T> const std::vector<XV<T>> &xr2c_variants(int m);
template <typename T> const std::vector<XV<T>> &xc2r_variants(int m);
template <typename T> const std::vector<PXV<T>> &pfcx_variants(int n);
template <typename T> const std::vector<TXV<T>> &tungx_variants(int m);
template <typename T> const std::vector<TV<T>> &t_variants(int kind, int n);
#define PFC_AMD_DECLARE_LISTS(T)                                               \
  template <> const std::vector<ColV<T>> &col_variants<T>(int);                \
  template <> const std::vector<ZV<T>> &pfcz_variants<T>(int);                 \
  template <> const std::vector<XV<T>> &xr2c_variants<T>(int);                 \
  template <> const std::vector<XV<T>> &xc2r_variants<T>(int);                 \
  template <> const std::vector<PXV<T>> &pfcx_variants<T>(int);               \
  template <> const std::vector<TXV<T>> &tungx_variants<T>(int);               \
  template <> const std::vector<TV<T>> &t_variants<T>(int, int);
PFC_AMD_DECLARE_LISTS(double)
PFC_AMD_DECLARE_LISTS(float)

// Radices of line plan `plan_id` of length n (host twiddle tables are built
// per (length, plan) since the table layout follows the pass sequence).
// Plan ids from kXpPid on are the double-butterfly plans (their tables hold
// double entries whatever the field precision).
constexpr int kXpPid = 64;
std::vector<int> radices_for(int n, int plan_id);
int plan_count(int n);
int plan_count_xp(int n);
bool supported_len(int n);

// Transposed PFC passes (kinds TK_X_FWD .. TK_X_INV) of single-precision
// plans made with PFC_HIP_FLAG_F64_BUTTERFLIES (empty for other kinds and
// for lengths without such plans).
const std::vector<TV<float>> &t_variants_xp(int kind, int n);

} // namespace pfc_amd
