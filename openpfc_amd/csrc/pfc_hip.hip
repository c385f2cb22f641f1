// pfc_hip.hip -- C-ABI implementation: plans, twiddle tables, size dispatch
// and the launch sequences of the R2C/C2R transforms and the fused PFC step.
#include "../../include/pfc_hip.h"
#include "adjoint_kernels.hpp"
#include "energy_kernels.hpp"
#include "variants.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <map>
#include <mutex>
#include <type_traits>
#include <utility>
#include <vector>

using namespace pfc_amd;


namespace pfc_amd {
// message of the last failed ABI call on this thread (field_io.hip sets it too)
thread_local std::string g_last_error;

// Scratch of pfc_hip_column_max_noalloc: the partial maxima of up to 256
// rows of nx <= 2048 columns (4 MiB), one copy per GPU for the life of the
// process.  Its name is a host-side stand-in for the device array, so host
// code uses only the address col_scratch() obtains on each device.
constexpr long kColScratchDoubles = 256L * 2048;
__device__ __align__(256) double scratch[kColScratchDoubles];
} // namespace pfc_amd

namespace {

using pfc_amd::g_last_error;

int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

#define PFC_HIP_TRY(expr, what)                                                \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return fail(PFC_HIP_EDEVICE,                                             \
                  std::string(what) + ": " + hipGetErrorString(e_));           \
  } while (0)

// ---------------------------------------------------------------------------
// Host twiddle tables (long double, rounded once).
// ---------------------------------------------------------------------------
const long double kTwoPi = 6.283185307179586476925286766559005768L;

void unit_root_ld(long m, long n, long double &c, long double &s) {
  m %= n;
  if (m < 0) m += n;
  // exact values on the axes
  if (m == 0) { c = 1; s = 0; return; }
  if (2 * m == n) { c = -1; s = 0; return; }
  if (4 * m == n) { c = 0; s = 1; return; }
  if (4 * m == 3 * n) { c = 0; s = -1; return; }
  long double a = kTwoPi * (long double)m / (long double)n;
  c = cosl(a);
  s = sinl(a);
}

// Line twiddles for the Stockham passes of plan `pid`: exp(-2 pi i q k/(Ns R)).
std::vector<long double> line_twiddles(int n, int pid) {
  std::vector<long double> out; // re, im pairs
  int ns = 1;
  for (int r : radices_for(n, pid)) {
    if (ns > 1) {
      for (int k = 0; k < ns; ++k)
        for (int q = 1; q < r; ++q) {
          long double c, s;
          unit_root_ld((long)q * k, (long)ns * r, c, s);
          out.push_back(c);
          out.push_back(-s);
        }
    }
    ns *= r;
  }
  return out;
}

// exp(-2 pi i k/nx), k = 0..nx/2.
std::vector<long double> real_twiddles(int nx) {
  std::vector<long double> out;
  for (int k = 0; k <= nx / 2; ++k) {
    long double c, s;
    unit_root_ld(k, nx, c, s);
    out.push_back(c);
    out.push_back(-s);
  }
  return out;
}

} // namespace

// ---------------------------------------------------------------------------
// Plan object.
// ---------------------------------------------------------------------------
struct pfc_hip_plan_s {
  int nx = 0, ny = 0, nz = 0, prec = 0, flags = 0;
  int W = 0;        // nx/2 + 1
  int P = 0;        // padded pitch of the plan's own spectrum buffers
  size_t esize = 8; // sizeof(real)
  int device = 0;
  // line twiddles per (length, plan id) and real-FFT twiddles of nx
  std::map<std::pair<int, int>, void *> tw;
  void *wn = nullptr;
  void *work = nullptr; // c2r workspace: W*ny*nz complex
  // fused-step state: natural layout (A, B pitched by P) or transposed
  // layouts T1 (A1, B1) and T2 (A2, B2), W*ny*nz complex each
  void *A = nullptr, *B = nullptr;
  void *A1 = nullptr, *B1 = nullptr, *A2 = nullptr, *B2 = nullptr;
  // adjoint workspace: three outbox spectra, two real fields, block partials
  void *adj_lam = nullptr, *adj_P = nullptr, *adj_N = nullptr;
  void *adj_a = nullptr, *adj_b = nullptr, *adj_part = nullptr;
  // column maxima (MovingBC): kColParts partial rows of nx, then the result
  void *colmax = nullptr;
  // exp(-2 pi i k/nx) in double for the C2R merge of FP64-butterfly plans
  void *wn64 = nullptr;
  bool ops_ready = false;
  int model_kind = -1;
  double dt = 0;
  PfcOp op{};    // k-grid, dt, 1/(nx ny nz); eps of the PFC model
  double spacing[3] = {}; // dx, dy, dz of the operators
  TungOp top{};  // Tungsten operator scalars
  double tq[5] = {}; // Tungsten nonlinearity: p3_bar, p4_bar, q3_bar, q4_bar, stabP
  bool mean_split = false; // field holds psi - mean; mean kept in double here
  double mean = 0.0;
  int var[ROLE_COUNT] = {};
  size_t bytes = 0;
  // Batched plans step `batch` fields of these extents stored one after
  // another.  The natural layout steps them all in each launch, reading
  // per-field eps and (3m, 3m^2) from m_eps (double) and m_nl (real type)
  // once `members` is set; the transposed layout steps them one after
  // another (every pass of field 0, then of field 1, ...: a field's spectra
  // stay in the last-level cache between its passes), through one set of
  // T1/T2 buffers, with each field's eps and mean (h_eps, h_mean) as launch
  // arguments.
  int batch = 1;
  void *m_eps = nullptr, *m_nl = nullptr;
  bool members = false;
  std::vector<double> h_eps, h_mean;
  // Bumped by every call that changes a value the step's kernels receive as
  // a launch argument (operators, eps or mean held as scalars, the
  // mean-split mode, variants): a graph captured before such a change
  // replays stale arguments and must be captured again.
  unsigned gen = 0;

  size_t inbox() const { return (size_t)nx * ny * nz; }
  size_t outbox() const { return (size_t)W * ny * nz; }
  size_t padded() const { return (size_t)P * ny * nz; }
  int role_len(int role) const {
    switch (role) {
    case ROLE_X_R2C: case ROLE_X_C2R: return nx / 2;
    case ROLE_COL_Y: return ny;
    case ROLE_COL_Z: case ROLE_PFC_Z: return nz;
    case ROLE_PFC_X: return nx;
    case ROLE_T_X_FWD: return nx;
    case ROLE_T_Y_FWD: case ROLE_T_Y_INV: return ny;
    case ROLE_T_Z: return nz;
    case ROLE_T_X_INV: return nx / 2;
    case ROLE_X_TUNG: case ROLE_T_X_R2C: case ROLE_T_X_TUNG: return nx / 2;
    case ROLE_T_X_FUSE: return nx / 2;
    default: return 0;
    }
  }
  bool transposed() const { return var[ROLE_STEP_LAYOUT] == 0 && t_ok; }
  bool tungsten() const { return model_kind == PFC_HIP_MODEL_TUNGSTEN; }
  int step_passes() const { return tungsten() ? 9 : 5; }
  bool t_ok = false; // transposed step supported for these extents
  // PFC_HIP_FLAG_F64_BUTTERFLIES: single-precision fields whose transposed
  // PFC passes run their butterflies in double (t_variants_xp lists)
  bool xp = false;
};

namespace {

int dalloc(pfc_hip_plan_s *p, void **ptr, size_t bytes, const char *what) {
  if (*ptr) return 0;
  hipError_t e = hipMalloc(ptr, bytes ? bytes : 16);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *ptr = nullptr;
    return fail(PFC_HIP_ENOMEM, std::string("hipMalloc(") + what + ", " +
                                    std::to_string(bytes) +
                                    " bytes): " + hipGetErrorString(e));
  }
  p->bytes += bytes ? bytes : 16;
  return 0;
}

int upload_table(pfc_hip_plan_s *p, void **dst, const std::vector<long double> &v,
                 const char *what) {
  const size_t n = v.size();
  const size_t bytes = n * p->esize;
  if (int rc = dalloc(p, dst, bytes, what)) return rc;
  if (n == 0) return 0;
  if (p->prec == PFC_HIP_F64) {
    std::vector<double> h(n);
    for (size_t i = 0; i < n; ++i) h[i] = (double)v[i];
    PFC_HIP_TRY(hipMemcpy(*dst, h.data(), bytes, hipMemcpyHostToDevice), what);
  } else {
    std::vector<float> h(n);
    for (size_t i = 0; i < n; ++i) h[i] = (float)v[i];
    PFC_HIP_TRY(hipMemcpy(*dst, h.data(), bytes, hipMemcpyHostToDevice), what);
  }
  return 0;
}

// Twiddles of every plan of length n.
int upload_line_tables(pfc_hip_plan_s *p, int n) {
  for (int pid = 0; pid < plan_count(n); ++pid) {
    auto key = std::make_pair(n, pid);
    if (p->tw.count(key)) continue;
    void *d = nullptr;
    if (int rc = upload_table(p, &d, line_twiddles(n, pid), "line twiddles")) return rc;
    p->tw[key] = d;
  }
  return 0;
}

// Twiddles of the double-butterfly plans of length n: double entries at
// plan ids kXpPid + i, whatever the plan's precision.
int upload_line_tables_xp(pfc_hip_plan_s *p, int n) {
  for (int i = 0; i < plan_count_xp(n); ++i) {
    const int pid = kXpPid + i;
    auto key = std::make_pair(n, pid);
    if (p->tw.count(key)) continue;
    const std::vector<long double> v = line_twiddles(n, pid);
    void *d = nullptr;
    if (int rc = dalloc(p, &d, v.size() * sizeof(double), "line twiddles (double)")) return rc;
    p->tw[key] = d;
    if (!v.empty()) {
      std::vector<double> h(v.begin(), v.end());
      PFC_HIP_TRY(hipMemcpy(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice),
                  "line twiddles (double)");
    }
  }
  return 0;
}

template <typename T> const cplx<T> *twp(pfc_hip_plan_s *p, int n, int pid) {
  auto it = p->tw.find(std::make_pair(n, pid));
  return it == p->tw.end() ? nullptr : (const cplx<T> *)it->second;
}

template <typename T> TungNl<T> tung_nonlin(const pfc_hip_plan_s *p) {
  return TungNl<T>{(T)p->tq[0], (T)p->tq[1], (T)p->tq[2], (T)p->tq[3], (T)p->tq[4]};
}

// Transposed-step kernel kind of a role, or -1.
int t_kind_of_role(int role) {
  if (role >= ROLE_T_X_FWD && role <= ROLE_T_X_INV) return role - ROLE_T_X_FWD;
  if (role == ROLE_T_X_R2C) return TK_X_R2C;
  if (role == ROLE_T_X_TUNG) return TK_X_TUNG;
  if (role == ROLE_T_X_FUSE) return TK_X_FUSE;
  return -1;
}

template <typename T> Nonlin<T> nonlin_of(bool split, double mean) {
  Nonlin<T> n{};
  n.split = split ? 1 : 0;
  n.c1 = (T)(3.0 * mean);
  n.c2 = (T)(3.0 * mean * mean);
  n.c1d = 3.0 * mean;
  n.c2d = 3.0 * mean * mean;
  return n;
}
template <typename T> Nonlin<T> nonlin(const pfc_hip_plan_s *p) {
  return nonlin_of<T>(p->mean_split, p->mean);
}

void free_all(pfc_hip_plan_s *p) {
  for (auto &kv : p->tw)
    if (kv.second) (void)hipFree(kv.second);
  p->tw.clear();
  void **bufs[] = {&p->wn,      &p->work,  &p->A,     &p->B,     &p->A1,
                   &p->B1,      &p->A2,    &p->B2,    &p->adj_lam, &p->adj_P,
                   &p->adj_N,   &p->adj_a, &p->adj_b, &p->adj_part, &p->m_eps,
                   &p->m_nl,    &p->colmax, &p->wn64};
  for (void **b : bufs) {
    if (*b) (void)hipFree(*b);
    *b = nullptr;
  }
  p->bytes = 0;
}

// Variant list of a transposed-step kind: the double-butterfly passes on a
// PFC_HIP_FLAG_F64_BUTTERFLIES plan (none for the Tungsten kinds and the fused
// boundary there), the plan precision's list otherwise.
template <typename T> const std::vector<TV<T>> &t_list(const pfc_hip_plan_s *p, int kind, int n) {
  if constexpr (std::is_same_v<T, float>) {
    if (p->xp) return t_variants_xp(kind, n);
  }
  return t_variants<T>(kind, n);
}

template <typename T> int variant_count(pfc_hip_plan_s *p, int role) {
  const int n = p->role_len(role);
  if (role == ROLE_STEP_LAYOUT) return p->t_ok && !p->xp ? 2 : 1;
  if (role == ROLE_STEP_FUSE)
    return p->t_ok && !t_list<T>(p, TK_X_FUSE, p->role_len(ROLE_T_X_FUSE)).empty() ? 2 : 1;
  if (t_kind_of_role(role) >= 0) return (int)t_list<T>(p, t_kind_of_role(role), n).size();
  switch (role) {
  case ROLE_X_R2C: return (int)xr2c_variants<T>(n).size();
  case ROLE_X_C2R: return (int)xc2r_variants<T>(n).size();
  case ROLE_COL_Y: case ROLE_COL_Z: return (int)col_variants<T>(n).size();
  case ROLE_PFC_X: return (int)pfcx_variants<T>(n).size();
  case ROLE_PFC_Z: return (int)pfcz_variants<T>(n).size();
  case ROLE_X_TUNG: return (int)tungx_variants<T>(n).size();
  default: return 0;
  }
}

template <typename T> const std::string *variant_name(pfc_hip_plan_s *p, int role, int v) {
  static const std::string layouts[2] = {"transposed", "natural"};
  static const std::string fuse[2] = {"per_step", "fused"};
  const int n = p->role_len(role);
  if (v < 0 || v >= variant_count<T>(p, role)) return nullptr;
  if (role == ROLE_STEP_LAYOUT) return &layouts[p->t_ok ? v : 1];
  if (role == ROLE_STEP_FUSE) return &fuse[v];
  if (t_kind_of_role(role) >= 0) return &t_list<T>(p, t_kind_of_role(role), n)[v].name;
  switch (role) {
  case ROLE_X_R2C: return &xr2c_variants<T>(n)[v].name;
  case ROLE_X_C2R: return &xc2r_variants<T>(n)[v].name;
  case ROLE_COL_Y: case ROLE_COL_Z: return &col_variants<T>(n)[v].name;
  case ROLE_PFC_X: return &pfcx_variants<T>(n)[v].name;
  case ROLE_PFC_Z: return &pfcz_variants<T>(n)[v].name;
  case ROLE_X_TUNG: return &tungx_variants<T>(n)[v].name;
  default: return nullptr;
  }
}

// Column pass along y (axis 1) or z (axis 2) over a (W, ny, nz) array with
// row pitch `pitch`; fields 0/1 in src/dst.
template <typename T>
hipError_t col_pass(pfc_hip_plan_s *p, int axis, int dir, const void *src0,
                    void *dst0, const void *src1, void *dst1, int nfields,
                    int pitch, hipStream_t s) {
  const int role = axis == 1 ? ROLE_COL_Y : ROLE_COL_Z;
  const int n = axis == 1 ? p->ny : p->nz;
  const auto &vs = col_variants<T>(n);
  const auto &v = vs[p->var[role]];
  ColArgs<T> a{};
  a.src[0] = (const cplx<T> *)src0;
  a.dst[0] = (cplx<T> *)dst0;
  a.src[1] = (const cplx<T> *)src1;
  a.dst[1] = (cplx<T> *)dst1;
  a.tw = twp<T>(p, n, v.plan_id);
  a.W = p->W;
  a.s_pos = axis == 1 ? (long)pitch : (long)pitch * p->ny;
  a.s_outer = axis == 1 ? (long)pitch * p->ny : (long)pitch;
  // y columns of a batched plan run over the z planes of every field
  const long n_outer = axis == 1 ? (long)p->nz * p->batch : p->ny;
  return v.fn(dir, a, n_outer, nfields, s);
}

template <typename T>
hipError_t x_c2r_pass(pfc_hip_plan_s *p, const void *in, void *out, int pitch,
                      double scale, hipStream_t s) {
  const int m = p->nx / 2;
  const auto &v = xc2r_variants<T>(m)[p->var[ROLE_X_C2R]];
  XArgs<T> a{};
  a.in = (const T *)in;
  a.out = (T *)out;
  a.tw = twp<T>(p, m, v.plan_id);
  a.wn = (const cplx<T> *)p->wn;
  a.rows = (long)p->ny * p->nz * p->batch;
  a.pitch = pitch;
  a.scale = (T)scale;
  return v.fn(a, s);
}

template <typename T>
int r2c_impl(pfc_hip_plan_s *p, const void *in, void *out, hipStream_t s) {
  const int m = p->nx / 2;
  const auto &v = xr2c_variants<T>(m)[p->var[ROLE_X_R2C]];
  XArgs<T> a{};
  a.in = (const T *)in;
  a.out = (T *)out;
  a.tw = twp<T>(p, m, v.plan_id);
  a.wn = (const cplx<T> *)p->wn;
  a.rows = (long)p->ny * p->nz;
  a.pitch = p->W;
  a.scale = (T)1;
  hipError_t e = v.fn(a, s);
  PFC_HIP_TRY(e, "r2c x pass");
  if (p->ny > 1) {
    e = col_pass<T>(p, 1, -1, out, out, nullptr, nullptr, 1, p->W, s);
    PFC_HIP_TRY(e, "r2c y pass");
  }
  if (p->nz > 1) {
    e = col_pass<T>(p, 2, -1, out, out, nullptr, nullptr, 1, p->W, s);
    PFC_HIP_TRY(e, "r2c z pass");
  }
  return 0;
}

// C2R through the workspace `work` (outbox complex values; the plan's own,
// allocated on first use, when null).
template <typename T>
int c2r_impl(pfc_hip_plan_s *p, const void *in, void *out, hipStream_t s,
             void *work = nullptr) {
  const void *cur = in;
  hipError_t e;
  if (!work && (p->nz > 1 || p->ny > 1)) {
    if (int rc = dalloc(p, &p->work, p->outbox() * 2 * p->esize, "c2r workspace"))
      return rc;
    work = p->work;
  }
  if (p->nz > 1) {
    e = col_pass<T>(p, 2, +1, cur, work, nullptr, nullptr, 1, p->W, s);
    PFC_HIP_TRY(e, "c2r z pass");
    cur = work;
  }
  if (p->ny > 1) {
    e = col_pass<T>(p, 1, +1, cur, work, nullptr, nullptr, 1, p->W, s);
    PFC_HIP_TRY(e, "c2r y pass");
    cur = work;
  }
  e = x_c2r_pass<T>(p, cur, out, p->W,
                    1.0 / ((double)p->nx * (double)p->ny * (double)p->nz), s);
  PFC_HIP_TRY(e, "c2r x pass");
  return 0;
}

// z pass of the natural layout on the plan's spectra A and B (ZMode `mode`).
template <typename T> hipError_t z_natural(pfc_hip_plan_s *p, int mode, hipStream_t s) {
  const auto &v = pfcz_variants<T>(p->nz)[p->var[ROLE_PFC_Z]];
  PfcZArgs<T> a{};
  a.zero_k0 = (mode == ZM_PFC && p->mean_split) ? 1 : 0;
  a.mode = mode;
  a.A = (cplx<T> *)p->A;
  a.B = (cplx<T> *)p->B;
  a.tw = twp<T>(p, p->nz, v.plan_id);
  a.op = p->op;
  a.top = p->top;
  a.meps = p->members ? (const double *)p->m_eps : nullptr;
  a.mstride = (long)p->padded();
  a.s_pos = (long)p->P * p->ny;
  a.pitch = p->P;
  a.W = p->W;
  return v.fn(a, (long)p->ny * p->batch, s);
}

// One fused PFC step: 5 launches.  `ev` (optional) receives 6 events
// bracketing the 5 passes.
template <typename T>
int step_natural(pfc_hip_plan_s *p, void *psi, hipStream_t s, hipEvent_t *ev) {
  hipError_t e;
  auto mark = [&](int i) {
    if (ev) (void)hipEventRecord(ev[i], s);
  };
  mark(0);
  {
    const auto &v = pfcx_variants<T>(p->nx)[p->var[ROLE_PFC_X]];
    PfcXArgs<T> a{};
    a.nl = nonlin<T>(p);
    a.mnl = p->members ? (const T *)p->m_nl : nullptr;
    a.mrows = (long)p->ny * p->nz;
    a.psi = (const T *)psi;
    a.outA = (cplx<T> *)p->A;
    a.outB = (cplx<T> *)p->B;
    a.tw = twp<T>(p, p->nx, v.plan_id);
    a.rows = (long)p->ny * p->nz * p->batch;
    a.pitch = p->P;
    e = v.fn(a, s);
    PFC_HIP_TRY(e, "pfc forward x pass");
  }
  mark(1);
  if (p->ny > 1) {
    e = col_pass<T>(p, 1, -1, p->A, p->A, p->B, p->B, 2, p->P, s);
    PFC_HIP_TRY(e, "pfc forward y pass");
  }
  mark(2);
  PFC_HIP_TRY(z_natural<T>(p, ZM_PFC, s), "pfc z pass");
  mark(3);
  if (p->ny > 1) {
    e = col_pass<T>(p, 1, +1, p->A, p->A, nullptr, nullptr, 1, p->P, s);
    PFC_HIP_TRY(e, "pfc inverse y pass");
  }
  mark(4);
  // 1/(nx ny nz) is folded into opL/opN
  e = x_c2r_pass<T>(p, p->A, psi, p->P, 1.0, s);
  PFC_HIP_TRY(e, "pfc inverse x pass");
  mark(5);
  return 0;
}

// Step buffers of the active layout (allocated on first use).
int step_buffers(pfc_hip_plan_s *p) {
  int rc = 0;
  if (p->transposed()) {
    const size_t c = p->outbox() * 2 * p->esize;
    if (!rc) rc = dalloc(p, &p->A1, c, "step T1 psi");
    if (!rc) rc = dalloc(p, &p->B1, c, "step T1 N");
    if (!rc) rc = dalloc(p, &p->A2, c, "step T2 psi");
    if (!rc) rc = dalloc(p, &p->B2, c, "step T2 N");
  } else {
    const size_t c = p->padded() * 2 * p->esize * p->batch;
    if (!rc) rc = dalloc(p, &p->A, c, "step spectrum A");
    if (!rc) rc = dalloc(p, &p->B, c, "step spectrum B");
  }
  return rc;
}

// Launch arguments of the transposed PFC step of field m of the plan (its
// own eps and mean when the members are set).
template <typename T> TStepArgs<T> pfc_t_args(pfc_hip_plan_s *p, void *psi, int m = 0) {
  TStepArgs<T> a{};
  T *field = (T *)psi + (size_t)m * p->inbox();
  a.psi_in = field;
  a.psi_out = field;
  a.A1 = (cplx<T> *)p->A1;
  a.B1 = (cplx<T> *)p->B1;
  a.A2 = (cplx<T> *)p->A2;
  a.B2 = (cplx<T> *)p->B2;
  // FP64-butterfly kernels read their merge twiddles as double entries
  a.wn = (const cplx<T> *)(p->xp ? p->wn64 : p->wn);
  a.op = p->op;
  a.nl = nonlin<T>(p);
  if (p->members && p->batch > 1) {
    a.op.eps = p->h_eps[(size_t)m];
    a.nl = nonlin_of<T>(p->mean_split, p->h_mean[(size_t)m]);
  }
  a.zero_k0 = p->mean_split ? 1 : 0;
  a.mode = ZM_PFC;
  a.nfields = 2;
  a.nx = p->nx;
  a.ny = p->ny;
  a.nz = p->nz;
  a.W = p->W;
  return a;
}

// Pass k of the transposed PFC step: 0 forward x, 1 forward y, 2 z,
// 3 inverse y, 4 inverse x; 5 the fused step boundary (inverse x of one step
// with the forward x of the next).
constexpr int kTPassRoles[6] = {ROLE_T_X_FWD, ROLE_T_Y_FWD, ROLE_T_Z,
                                ROLE_T_Y_INV, ROLE_T_X_INV, ROLE_T_X_FUSE};
template <typename T>
int t_pass(pfc_hip_plan_s *p, TStepArgs<T> a, int k, hipStream_t s) {
  static const char *what[6] = {"step x forward", "step y forward", "step z pass",
                                "step y inverse", "step x inverse", "step boundary x"};
  const int role = kTPassRoles[k];
  const int n = p->role_len(role);
  const auto &v = t_list<T>(p, t_kind_of_role(role), n)[p->var[role]];
  const cplx<T> *tw = twp<T>(p, v.half ? n / 2 : n, v.plan_id);
  a.tw_x = a.tw_xm = a.tw_y = a.tw_z = tw;
  if (v.plan_id2 >= 0) a.tw_x = twp<T>(p, 2 * n, v.plan_id2);
  PFC_HIP_TRY(v.fn(a, s), what[k]);
  return 0;
}

// One transposed PFC step of every field of the plan, field after field.
// `ev` (optional) brackets the passes of field 0: the pass times of one
// field, as for a single plan.
template <typename T>
int step_transposed(pfc_hip_plan_s *p, void *psi, hipStream_t s, hipEvent_t *ev) {
  for (int m = 0; m < p->batch; ++m) {
    auto mark = [&](int i) {
      if (ev && m == 0) (void)hipEventRecord(ev[i], s);
    };
    const TStepArgs<T> a = pfc_t_args<T>(p, psi, m);
    mark(0);
    for (int k = 0; k < 5; ++k) {
      if (int rc = t_pass<T>(p, a, k, s)) return rc;
      mark(k + 1);
    }
  }
  return 0;
}

// Whether multi-step PFC runs fuse their step boundaries (role 16).
template <typename T> bool fused_steps(pfc_hip_plan_s *p) {
  return p->transposed() && !p->tungsten() && p->var[ROLE_STEP_FUSE] == 1 &&
         !t_list<T>(p, TK_X_FUSE, p->role_len(ROLE_T_X_FUSE)).empty();
}

// nsteps >= 1 PFC steps with fused boundaries: forward x once, then per
// step the y, z and inverse y passes, with the fused boundary pass between
// steps, and the inverse x once at the end (4*nsteps + 1 launches).  `ev`
// (optional) receives 4*nsteps + 2 events around the launches, in order.
// (Batched plans: the nsteps steps of field 0, then of field 1, ...; `ev`
// brackets those of field 0.)
template <typename T>
int steps_fused(pfc_hip_plan_s *p, void *psi, int nsteps, hipStream_t s, hipEvent_t *ev) {
  for (int m = 0; m < p->batch; ++m) {
    int e = 0;
    auto mark = [&]() {
      if (ev && m == 0) (void)hipEventRecord(ev[e++], s);
    };
    const TStepArgs<T> a = pfc_t_args<T>(p, psi, m);
    mark();
    for (int i = 0; i < nsteps; ++i) {
      if (int rc = t_pass<T>(p, a, i == 0 ? 0 : 5, s)) return rc;
      mark();
      for (int k = 1; k <= 3; ++k) {
        if (int rc = t_pass<T>(p, a, k, s)) return rc;
        mark();
      }
    }
    if (int rc = t_pass<T>(p, a, 4, s)) return rc;
    mark();
  }
  return 0;
}

// One Tungsten mean-field step (cpu/tungsten_model.hpp:306-366) in the
// natural layout: 9 launches, A holding Psi and B the mean field, then N.
// `mf` (optional) receives psi_MF of the input field.
template <typename T>
int step_tung_natural(pfc_hip_plan_s *p, void *psi, void *mf, hipStream_t s,
                      hipEvent_t *ev) {
  auto mark = [&](int i) {
    if (ev) (void)hipEventRecord(ev[i], s);
  };
  const int m = p->nx / 2;
  const long rows = (long)p->ny * p->nz;
  hipError_t e;
  mark(0);
  {
    const auto &v = xr2c_variants<T>(m)[p->var[ROLE_X_R2C]];
    XArgs<T> a{};
    a.in = (const T *)psi;
    a.out = (T *)p->A;
    a.tw = twp<T>(p, m, v.plan_id);
    a.wn = (const cplx<T> *)p->wn;
    a.rows = rows;
    a.pitch = p->P;
    a.scale = (T)1;
    PFC_HIP_TRY(v.fn(a, s), "tungsten forward x");
  }
  mark(1);
  if (p->ny > 1) {
    e = col_pass<T>(p, 1, -1, p->A, p->A, nullptr, nullptr, 1, p->P, s);
    PFC_HIP_TRY(e, "tungsten forward y");
  }
  mark(2);
  PFC_HIP_TRY(z_natural<T>(p, ZM_TMF, s), "tungsten mean-field z pass");
  mark(3);
  if (p->ny > 1) {
    e = col_pass<T>(p, 1, +1, p->B, p->B, nullptr, nullptr, 1, p->P, s);
    PFC_HIP_TRY(e, "tungsten mean-field inverse y");
  }
  mark(4);
  {
    const auto &v = tungx_variants<T>(m)[p->var[ROLE_X_TUNG]];
    TungXArgs<T> a{};
    a.nl = tung_nonlin<T>(p);
    a.psi = (const T *)psi;
    a.spec = (cplx<T> *)p->B;
    a.mf_out = (T *)mf;
    a.tw = twp<T>(p, m, v.plan_id);
    a.wn = (const cplx<T> *)p->wn;
    a.rows = rows;
    a.pitch = p->P;
    PFC_HIP_TRY(v.fn(a, s), "tungsten fused x pass");
  }
  mark(5);
  if (p->ny > 1) {
    e = col_pass<T>(p, 1, -1, p->B, p->B, nullptr, nullptr, 1, p->P, s);
    PFC_HIP_TRY(e, "tungsten forward y (N)");
  }
  mark(6);
  PFC_HIP_TRY(z_natural<T>(p, ZM_TCOMB, s), "tungsten combine z pass");
  mark(7);
  if (p->ny > 1) {
    e = col_pass<T>(p, 1, +1, p->A, p->A, nullptr, nullptr, 1, p->P, s);
    PFC_HIP_TRY(e, "tungsten inverse y");
  }
  mark(8);
  PFC_HIP_TRY(x_c2r_pass<T>(p, p->A, psi, p->P, 1.0, s), "tungsten inverse x");
  mark(9);
  return 0;
}

// The Tungsten step on the transposed layouts: 9 launches (T1/T2 buffers A
// for Psi, B for the mean field and then N).
template <typename T>
int step_tung_transposed(pfc_hip_plan_s *p, void *psi, void *mf, hipStream_t s,
                         hipEvent_t *ev) {
  auto mark = [&](int i) {
    if (ev) (void)hipEventRecord(ev[i], s);
  };
  TStepArgs<T> a{};
  a.mode = ZM_PFC;
  a.nfields = 1;
  a.top = p->top;
  a.tnl = tung_nonlin<T>(p);
  a.mf_out = (T *)mf;
  a.psi_in = (const T *)psi;
  a.psi_out = (T *)psi;
  a.A1 = (cplx<T> *)p->A1;
  a.B1 = (cplx<T> *)p->B1;
  a.A2 = (cplx<T> *)p->A2;
  a.B2 = (cplx<T> *)p->B2;
  a.wn = (const cplx<T> *)p->wn;
  a.op = p->op;
  a.nx = p->nx;
  a.ny = p->ny;
  a.nz = p->nz;
  a.W = p->W;
  struct TPass { int role, mode; bool on_b; const char *what; };
  // on_b: the pass reads and writes the B buffers (passed as A1/A2)
  static const TPass passes[9] = {
      {ROLE_T_X_R2C, ZM_PFC, false, "tungsten x forward"},
      {ROLE_T_Y_FWD, ZM_PFC, false, "tungsten y forward"},
      {ROLE_T_Z, ZM_TMF, false, "tungsten mean-field z pass"},
      {ROLE_T_Y_INV, ZM_PFC, true, "tungsten mean-field y inverse"},
      {ROLE_T_X_TUNG, ZM_PFC, false, "tungsten fused x pass"},
      {ROLE_T_Y_FWD, ZM_PFC, true, "tungsten y forward (N)"},
      {ROLE_T_Z, ZM_TCOMB, false, "tungsten combine z pass"},
      {ROLE_T_Y_INV, ZM_PFC, false, "tungsten y inverse"},
      {ROLE_T_X_INV, ZM_PFC, false, "tungsten x inverse"}};
  mark(0);
  for (int k = 0; k < 9; ++k) {
    const TPass &ps = passes[k];
    const int n = p->role_len(ps.role);
    const auto &v = t_list<T>(p, t_kind_of_role(ps.role), n)[p->var[ps.role]];
    TStepArgs<T> b = a;
    b.mode = ps.mode;
    if (ps.on_b) {
      b.A1 = a.B1;
      b.A2 = a.B2;
    }
    b.tw_x = b.tw_xm = b.tw_y = b.tw_z = twp<T>(p, v.half ? n / 2 : n, v.plan_id);
    PFC_HIP_TRY(v.fn(b, s), ps.what);
    mark(k + 1);
  }
  return 0;
}

// One step of the plan's model; `ev` (optional) receives step_passes()+1
// events bracketing the passes; `mf` (Tungsten, optional) receives psi_MF.
template <typename T>
int step_impl(pfc_hip_plan_s *p, void *psi, hipStream_t s, hipEvent_t *ev,
              void *mf = nullptr) {
  if (int rc = step_buffers(p)) return rc;
  if (p->tungsten())
    return p->transposed() ? step_tung_transposed<T>(p, psi, mf, s, ev)
                           : step_tung_natural<T>(p, psi, mf, s, ev);
  return p->transposed() ? step_transposed<T>(p, psi, s, ev)
                         : step_natural<T>(p, psi, s, ev);
}

constexpr unsigned kAdjBlocks = 2048; // blocks of the reduction kernels
// partials of two reductions plus their two sums
constexpr size_t kPartials = 2 * kAdjBlocks + 2;

// Adjoint of one plain PFC step (see adjoint_kernels.hpp).
template <typename T>
int adjoint_impl(pfc_hip_plan_s *p, const void *psi, void *lambda, double *dr, hipStream_t s);

// partial rows of the column maxima (pfc_hip_column_max)
constexpr long kColParts = 256;

constexpr long kColScratch = pfc_amd::kColScratchDoubles;
using pfc_amd::scratch;

// Device address of `scratch` on device `dev` (current), looked up on the
// first call on that device and kept for the process.
constexpr int kMaxDevices = 64;
int col_scratch(int dev, double **out) {
  static std::mutex mu;
  static double *addr[kMaxDevices] = {};
  if (dev < 0 || dev >= kMaxDevices)
    return fail(PFC_HIP_EUNSUPPORTED, "column-max scratch: device index " + std::to_string(dev));
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(PFC_HIP_EDEVICE, std::string("hipGetSymbolAddress(column-max scratch): ") +
                                       hipGetErrorString(e));
    addr[dev] = (double *)p;
  }
  *out = addr[dev];
  return 0;
}

// second stage of the column maxima: the maximum of the nparts partials
__global__ void __launch_bounds__(256) k_column_max_final(const double *__restrict__ part,
                                                          int nx, int nparts,
                                                          double *__restrict__ out) {
  const int x = blockIdx.x * 256 + threadIdx.x;
  if (x >= nx) return;
  double m = part[x];
  for (int p = 1; p < nparts; ++p) {
    const double v = part[(long)p * nx + x];
    m = v > m ? v : m;
  }
  out[x] = m;
}


unsigned ic_blocks(long n) {
  return (unsigned)std::min<long>((n + 255) / 256, 1L << 20);
}

template <typename T>
int adjoint_impl(pfc_hip_plan_s *p, const void *psi, void *lambda, double *dr, hipStream_t s) {
  const size_t cb = p->outbox() * 2 * p->esize, rb = p->inbox() * p->esize;
  int rc = 0;
  if (!rc) rc = dalloc(p, &p->adj_lam, cb, "adjoint spectrum");
  if (!rc) rc = dalloc(p, &p->adj_P, cb, "adjoint spectrum");
  if (!rc) rc = dalloc(p, &p->adj_N, cb, "adjoint spectrum");
  if (!rc) rc = dalloc(p, &p->adj_a, rb, "adjoint field");
  if (!rc) rc = dalloc(p, &p->adj_b, rb, "adjoint field");
  if (!rc) rc = dalloc(p, &p->adj_part, kPartials * sizeof(double), "reduction partials");
  if (rc) return rc;
  const long n = (long)p->inbox();
  PfcOp o = p->op;
  o.scale = 1.0; // the C2R below applies 1/(nx ny nz)
  if (dr) {
    if ((rc = r2c_impl<T>(p, psi, p->adj_P, s))) return rc;
    hipLaunchKernelGGL(k_cube<T>, dim3(ic_blocks(n)), dim3(256), 0, s, (const T *)psi,
                       (T *)p->adj_a, n, nonlin<T>(p));
    PFC_HIP_TRY(hipGetLastError(), "adjoint cube");
    if ((rc = r2c_impl<T>(p, p->adj_a, p->adj_N, s))) return rc;
  }
  if ((rc = r2c_impl<T>(p, lambda, p->adj_lam, s))) return rc;
  hipLaunchKernelGGL(k_adj_spectral<T>, dim3(kAdjBlocks), dim3(256), 0, s, o, p->W,
                     (const cplx<T> *)p->adj_lam, (cplx<T> *)p->adj_P, (cplx<T> *)p->adj_N,
                     (double *)p->adj_part, dr ? 1 : 0);
  PFC_HIP_TRY(hipGetLastError(), "adjoint spectral kernel");
  if (dr) {
    double *part = (double *)p->adj_part;
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, (const double *)part,
                       (int)kAdjBlocks, part + kAdjBlocks);
    PFC_HIP_TRY(hipGetLastError(), "adjoint reduction");
    double h = 0;
    PFC_HIP_TRY(hipMemcpyAsync(&h, part + kAdjBlocks, sizeof(double), hipMemcpyDeviceToHost, s),
                "adjoint reduction copy");
    PFC_HIP_TRY(hipStreamSynchronize(s), "adjoint sync");
    *dr += h / ((double)p->nx * (double)p->ny * (double)p->nz);
  }
  if ((rc = c2r_impl<T>(p, p->adj_P, p->adj_a, s))) return rc;
  if ((rc = c2r_impl<T>(p, p->adj_N, p->adj_b, s))) return rc;
  hipLaunchKernelGGL(k_adj_combine<T>, dim3(ic_blocks(n)), dim3(256), 0, s, (const T *)psi,
                     (const T *)p->adj_a, (const T *)p->adj_b, (T *)lambda, n,
                     (T)(p->mean_split ? p->mean : 0.0));
  PFC_HIP_TRY(hipGetLastError(), "adjoint combine");
  return 0;
}

// Free energy of the plain PFC model (energy_kernels.hpp), summed in double.
template <typename T>
int free_energy_impl(pfc_hip_plan_s *p, const void *psi, double *F, hipStream_t s) {
  int rc = dalloc(p, &p->adj_P, p->outbox() * 2 * p->esize, "free-energy spectrum");
  if (!rc) rc = dalloc(p, &p->adj_part, kPartials * sizeof(double), "reduction partials");
  if (rc) return rc;
  if ((rc = r2c_impl<T>(p, psi, p->adj_P, s))) return rc;
  const double npts = (double)p->nx * (double)p->ny * (double)p->nz;
  const double m = p->mean_split ? p->mean : 0.0;
  double *part = (double *)p->adj_part;
  hipLaunchKernelGGL(k_fe_spectral<T>, dim3(kAdjBlocks), dim3(256), 0, s, p->op, p->W,
                     (const cplx<T> *)p->adj_P, npts * m, part);
  PFC_HIP_TRY(hipGetLastError(), "free-energy spectral sum");
  hipLaunchKernelGGL(k_fe_quartic<T>, dim3(kAdjBlocks), dim3(256), 0, s, (const T *)psi,
                     (long)p->inbox(), m, part + kAdjBlocks);
  PFC_HIP_TRY(hipGetLastError(), "free-energy quartic sum");
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, (const double *)part,
                     (int)kAdjBlocks, part + 2 * kAdjBlocks);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, (const double *)part + kAdjBlocks,
                     (int)kAdjBlocks, part + 2 * kAdjBlocks + 1);
  PFC_HIP_TRY(hipGetLastError(), "free-energy reduction");
  double h[2] = {0, 0};
  PFC_HIP_TRY(hipMemcpyAsync(h, part + 2 * kAdjBlocks, sizeof(h), hipMemcpyDeviceToHost, s),
              "free-energy copy");
  PFC_HIP_TRY(hipStreamSynchronize(s), "free-energy sync");
  const double dV = p->spacing[0] * p->spacing[1] * p->spacing[2];
  *F = dV * (h[0] / (2.0 * npts) + 0.25 * h[1]);
  return 0;
}

// One 16-byte value per thread: the plain streaming copy of
// pfc_hip_stream_copy.
__global__ void __launch_bounds__(256) k_stream_copy(const uint4 *__restrict__ src,
                                                     uint4 *__restrict__ dst, size_t n) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

int check_plan(pfc_hip_plan_t p) {
  if (!p) return fail(PFC_HIP_EINVAL, "null plan");
  return 0;
}

// Entry points that act on one field refuse batched plans.
int check_single(pfc_hip_plan_t p, const char *fn) {
  if (int rc = check_plan(p)) return rc;
  if (p->batch > 1)
    return fail(PFC_HIP_EUNSUPPORTED, std::string(fn) + ": not available on a batched plan");
  return 0;
}

int set_device(pfc_hip_plan_t p) {
  int cur = -1;
  if (hipGetDevice(&cur) == hipSuccess && cur == p->device) return 0;
  PFC_HIP_TRY(hipSetDevice(p->device), "hipSetDevice");
  return 0;
}

} // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int pfc_hip_abi_version(void) { return PFC_HIP_ABI_VERSION; }

const char *pfc_hip_last_error(void) { return g_last_error.c_str(); }

int pfc_hip_plan_create(int nx, int ny, int nz, int precision, int flags,
                        pfc_hip_plan_t *plan) {
  if (!plan) return fail(PFC_HIP_EINVAL, "pfc_hip_plan_create: plan is null");
  *plan = nullptr;
  if (nx < 1 || ny < 1 || nz < 1)
    return fail(PFC_HIP_EINVAL, "pfc_hip_plan_create: extents must be >= 1");
  if (precision != PFC_HIP_F64 && precision != PFC_HIP_F32)
    return fail(PFC_HIP_EINVAL, "pfc_hip_plan_create: unknown precision " +
                                    std::to_string(precision));
  if (flags & ~PFC_HIP_FLAG_F64_BUTTERFLIES)
    return fail(PFC_HIP_EINVAL, "pfc_hip_plan_create: unknown flags " + std::to_string(flags));
  if ((flags & PFC_HIP_FLAG_F64_BUTTERFLIES) && precision != PFC_HIP_F32)
    return fail(PFC_HIP_EINVAL,
                "pfc_hip_plan_create: PFC_HIP_FLAG_F64_BUTTERFLIES needs PFC_HIP_F32");
  if (nx % 2 != 0 || !supported_len(nx / 2) || !supported_len(ny) ||
      !supported_len(nz))
    return fail(PFC_HIP_EUNSUPPORTED,
                "pfc_hip_plan_create: unsupported extents " + std::to_string(nx) +
                    "x" + std::to_string(ny) + "x" + std::to_string(nz) +
                    " (each axis in {1,2,4,...,1024} or {12,24,48,96,192}; nx even with nx/2 "
                    "in the same set)");
  auto *p = new (std::nothrow) pfc_hip_plan_s();
  if (!p) return fail(PFC_HIP_ENOMEM, "pfc_hip_plan_create: host allocation");
  p->nx = nx;
  p->ny = ny;
  p->nz = nz;
  p->prec = precision;
  p->flags = flags;
  p->W = nx / 2 + 1;
  p->esize = precision == PFC_HIP_F64 ? 8 : 4;
  // 128-byte aligned rows for the plan-owned spectra
  const int per128 = (int)(128 / (2 * p->esize));
  p->P = (p->W + per128 - 1) / per128 * per128;
  if (hipGetDevice(&p->device) != hipSuccess) {
    delete p;
    return fail(PFC_HIP_EDEVICE, "pfc_hip_plan_create: no HIP device");
  }
  int rc = 0;
  if (!rc) rc = upload_line_tables(p, nx / 2);
  if (!rc && supported_len(nx)) rc = upload_line_tables(p, nx);
  if (!rc) rc = upload_line_tables(p, ny);
  if (!rc) rc = upload_line_tables(p, nz);
  if (!rc) rc = upload_table(p, &p->wn, real_twiddles(nx), "real twiddles");
  if (!rc) {
    bool ok = true;
    for (int r = 0; r < ROLE_COUNT; ++r) {
      if (t_kind_of_role(r) < 0) continue;
      int c = precision == PFC_HIP_F64 ? variant_count<double>(p, r)
                                       : variant_count<float>(p, r);
      ok = ok && c > 0;
    }
    p->t_ok = ok;
  }
  if (!rc && (flags & PFC_HIP_FLAG_F64_BUTTERFLIES)) {
    p->xp = true;
    bool ok = p->t_ok;
    for (int r = ROLE_T_X_FWD; r <= ROLE_T_X_INV && ok; ++r)
      ok = variant_count<float>(p, r) > 0;
    if (!ok) {
      free_all(p);
      delete p;
      return fail(PFC_HIP_EUNSUPPORTED,
                  "pfc_hip_plan_create: PFC_HIP_FLAG_F64_BUTTERFLIES needs the transposed "
                  "step (power-of-two extents, every axis >= 128, nx >= 256), got " +
                      std::to_string(nx) + "x" + std::to_string(ny) + "x" + std::to_string(nz));
    }
    for (int n : {nx, nx / 2, ny, nz})
      if (!rc) rc = upload_line_tables_xp(p, n);
    if (!rc) {
      const std::vector<long double> w = real_twiddles(nx);
      rc = dalloc(p, &p->wn64, w.size() * sizeof(double), "real twiddles (double)");
      if (!rc) {
        std::vector<double> h(w.begin(), w.end());
        hipError_t e = hipMemcpy(p->wn64, h.data(), h.size() * sizeof(double),
                                 hipMemcpyHostToDevice);
        if (e != hipSuccess)
          rc = fail(PFC_HIP_EDEVICE, std::string("real twiddles (double): ") +
                                         hipGetErrorString(e));
      }
    }
  }
  if (rc) {
    free_all(p);
    delete p;
    return rc;
  }
  *plan = p;
  return 0;
}

int pfc_hip_plan_create_batched(int nx, int ny, int nz, int batch, int precision, int flags,
                                pfc_hip_plan_t *plan) {
  if (!plan) return fail(PFC_HIP_EINVAL, "pfc_hip_plan_create_batched: plan is null");
  *plan = nullptr;
  if (batch < 1 || batch > (1 << 16))
    return fail(PFC_HIP_EINVAL, "pfc_hip_plan_create_batched: batch must be in [1, 65536]");
  if ((double)nx * ny * nz * batch > 9.0e18 / 16)
    return fail(PFC_HIP_EINVAL, "pfc_hip_plan_create_batched: too many points");
  if (int rc = pfc_hip_plan_create(nx, ny, nz, precision, flags, plan)) return rc;
  pfc_hip_plan_s *p = *plan;
  p->batch = batch;
  return 0;
}

int pfc_hip_plan_destroy(pfc_hip_plan_t plan) {
  if (!plan) return 0;
  (void)set_device(plan);
  free_all(plan);
  delete plan;
  return 0;
}

int pfc_hip_sizes(pfc_hip_plan_t plan, size_t *inbox, size_t *outbox,
                  size_t *workspace) {
  if (int rc = check_plan(plan)) return rc;
  if (inbox) *inbox = plan->inbox();
  if (outbox) *outbox = plan->outbox();
  if (workspace) *workspace = plan->outbox();
  return 0;
}

int pfc_hip_allocated_bytes(pfc_hip_plan_t plan, size_t *bytes) {
  if (int rc = check_plan(plan)) return rc;
  if (!bytes) return fail(PFC_HIP_EINVAL, "pfc_hip_allocated_bytes: null");
  *bytes = plan->bytes;
  return 0;
}

int pfc_hip_r2c(pfc_hip_plan_t plan, const void *in, void *out, void *stream) {
  if (int rc = check_single(plan, "pfc_hip_r2c")) return rc;
  if (!in || !out) return fail(PFC_HIP_EINVAL, "pfc_hip_r2c: null buffer");
  if (int rc = set_device(plan)) return rc;
  hipStream_t s = (hipStream_t)stream;
  return plan->prec == PFC_HIP_F64 ? r2c_impl<double>(plan, in, out, s)
                                   : r2c_impl<float>(plan, in, out, s);
}

int pfc_hip_c2r(pfc_hip_plan_t plan, const void *in, void *out, void *stream) {
  if (int rc = check_single(plan, "pfc_hip_c2r")) return rc;
  if (!in || !out) return fail(PFC_HIP_EINVAL, "pfc_hip_c2r: null buffer");
  if (int rc = set_device(plan)) return rc;
  hipStream_t s = (hipStream_t)stream;
  return plan->prec == PFC_HIP_F64 ? c2r_impl<double>(plan, in, out, s)
                                   : c2r_impl<float>(plan, in, out, s);
}

// Host-side scalars of the Tungsten operators, in the reference's order
// (tungsten_spectral.hpp:91-101), without contraction into FMAs.
static TungOp tungsten_scalars(const double *q) {
#pragma clang fp contract(off)
  // q = {alpha2, lambda2, farTol, highOrd, Bx, T, T0, stabP, p2_bar, q2_bar}
  TungOp t{};
  t.alpha2 = q[0];
  t.lambda2 = q[1];
  t.rTol = ((-q[0]) * std::log(q[2])) - 1.0;
  t.highOrd = (int)q[3];
  t.c_peak = q[4] * std::exp((-q[5]) / q[6]);
  t.sp2 = q[7] + q[8];
  t.q2 = q[9];
  return t;
}

int pfc_hip_pfc_set_operators(pfc_hip_plan_t plan, int model_kind,
                              const double *params, int nparams, double dt) {
  if (int rc = check_plan(plan)) return rc;
  if (model_kind != PFC_HIP_MODEL_PFC && model_kind != PFC_HIP_MODEL_TUNGSTEN)
    return fail(PFC_HIP_EUNSUPPORTED, "pfc_hip_pfc_set_operators: unknown model " +
                                          std::to_string(model_kind));
  if (model_kind == PFC_HIP_MODEL_PFC && (!params || nparams != 4))
    return fail(PFC_HIP_EINVAL,
                "pfc_hip_pfc_set_operators: PFC needs params {dx, dy, dz, eps}");
  if (model_kind == PFC_HIP_MODEL_TUNGSTEN && plan->batch > 1)
    return fail(PFC_HIP_EUNSUPPORTED,
                "pfc_hip_pfc_set_operators: batched plans run the PFC model only");
  if (model_kind == PFC_HIP_MODEL_TUNGSTEN && plan->xp)
    return fail(PFC_HIP_EUNSUPPORTED, "pfc_hip_pfc_set_operators: PFC_HIP_FLAG_F64_BUTTERFLIES "
                                      "plans run the PFC model only");
  if (model_kind == PFC_HIP_MODEL_TUNGSTEN && (!params || nparams != PFC_HIP_TUNGSTEN_NPARAMS))
    return fail(PFC_HIP_EINVAL,
                "pfc_hip_pfc_set_operators: Tungsten needs " +
                    std::to_string(PFC_HIP_TUNGSTEN_NPARAMS) + " params");
  if (model_kind == PFC_HIP_MODEL_PFC && !supported_len(plan->nx))
    return fail(PFC_HIP_EUNSUPPORTED,
                "pfc_hip_pfc_set_operators: nx unsupported for the fused step");
  const double dx = params[0], dy = params[1], dz = params[2];
  if (!(dx > 0) || !(dy > 0) || !(dz > 0) || !std::isfinite(dt))
    return fail(PFC_HIP_EINVAL, "pfc_hip_pfc_set_operators: bad spacing or dt");
  for (int i = 3; i < nparams; ++i)
    if (!std::isfinite(params[i]))
      return fail(PFC_HIP_EINVAL, "pfc_hip_pfc_set_operators: parameter " +
                                      std::to_string(i) + " is not finite");
  if (model_kind == PFC_HIP_MODEL_TUNGSTEN) {
    const double *q = params + 3;
    if (!(q[0] > 0) || !(q[1] > 0) || !(q[2] > 0) || q[6] == 0.0 || q[3] < 0 ||
        q[3] != std::floor(q[3]) || q[3] > 64)
      return fail(PFC_HIP_EINVAL,
                  "pfc_hip_pfc_set_operators: Tungsten needs alpha2 > 0, lambda2 > 0, "
                  "alpha_farTol > 0, T0 != 0 and an integer alpha_highOrd in [0, 64]");
  }
  if (int rc = set_device(plan)) return rc;
  // kspace.hpp:97-103: f = 2*pi/(d*n); the multipliers are evaluated per
  // mode in the z passes, scaled by the 1/(nx*ny*nz) of the backward transform.
  const double two_pi = 6.283185307179586;
  PfcOp o{};
  o.fx = two_pi / (dx * plan->nx);
  o.fy = two_pi / (dy * plan->ny);
  o.fz = two_pi / (dz * plan->nz);
  o.eps = model_kind == PFC_HIP_MODEL_PFC ? params[3] : 0.0;
  o.dt = dt;
  o.scale = 1.0 / ((double)plan->nx * (double)plan->ny * (double)plan->nz);
  o.nx = plan->nx;
  o.ny = plan->ny;
  o.nz = plan->nz;
  plan->op = o;
  plan->spacing[0] = dx;
  plan->spacing[1] = dy;
  plan->spacing[2] = dz;
  if (model_kind == PFC_HIP_MODEL_TUNGSTEN) {
    plan->top = tungsten_scalars(params + 3);
    for (int i = 0; i < 4; ++i) plan->tq[i] = params[13 + i]; // p3, p4, q3, q4
    plan->tq[4] = params[10];                                  // stabP
  }
  plan->model_kind = model_kind;
  if (int rc = step_buffers(plan)) return rc;
  plan->dt = dt;
  plan->ops_ready = true;
  ++plan->gen;
  return 0;
}

int pfc_hip_tungsten_operators(pfc_hip_plan_t plan, void *filterMF, void *expL,
                               void *nW, void *stream) {
  if (int rc = check_single(plan, "pfc_hip_tungsten_operators")) return rc;
  if (!filterMF || !expL || !nW)
    return fail(PFC_HIP_EINVAL, "pfc_hip_tungsten_operators: null buffer");
  if (!plan->ops_ready || !plan->tungsten())
    return fail(PFC_HIP_ESTATE, "pfc_hip_tungsten_operators: Tungsten operators not set");
  if (int rc = set_device(plan)) return rc;
  hipStream_t s = (hipStream_t)stream;
  const unsigned blocks = ic_blocks((long)plan->outbox());
  if (plan->prec == PFC_HIP_F64)
    hipLaunchKernelGGL(k_tung_operators<double>, dim3(blocks), dim3(256), 0, s,
                       plan->op, plan->top, plan->W, (double *)filterMF,
                       (double *)expL, (double *)nW);
  else
    hipLaunchKernelGGL(k_tung_operators<float>, dim3(blocks), dim3(256), 0, s,
                       plan->op, plan->top, plan->W, (float *)filterMF,
                       (float *)expL, (float *)nW);
  PFC_HIP_TRY(hipGetLastError(), "tungsten operator kernel launch");
  return 0;
}

int pfc_hip_pfc_operators(pfc_hip_plan_t plan, void *opL, void *opN,
                          void *stream) {
  if (int rc = check_single(plan, "pfc_hip_pfc_operators")) return rc;
  if (!opL || !opN) return fail(PFC_HIP_EINVAL, "pfc_hip_pfc_operators: null buffer");
  if (!plan->ops_ready || plan->tungsten())
    return fail(PFC_HIP_ESTATE, "pfc_hip_pfc_operators: PFC operators not set");
  if (int rc = set_device(plan)) return rc;
  hipStream_t s = (hipStream_t)stream;
  const unsigned blocks = ic_blocks((long)plan->outbox());
  if (plan->prec == PFC_HIP_F64)
    hipLaunchKernelGGL(k_pfc_operators<double>, dim3(blocks), dim3(256), 0, s,
                       plan->op, plan->W, (double *)opL, (double *)opN);
  else
    hipLaunchKernelGGL(k_pfc_operators<float>, dim3(blocks), dim3(256), 0, s,
                       plan->op, plan->W, (float *)opL, (float *)opN);
  PFC_HIP_TRY(hipGetLastError(), "operator kernel launch");
  return 0;
}

static int step_checks(pfc_hip_plan_t plan, void *psi, int nsteps, const char *fn) {
  if (int rc = check_plan(plan)) return rc;
  if (!psi) return fail(PFC_HIP_EINVAL, std::string(fn) + ": null psi");
  if (nsteps < 0) return fail(PFC_HIP_EINVAL, std::string(fn) + ": nsteps < 0");
  if (!plan->ops_ready)
    return fail(PFC_HIP_ESTATE, std::string(fn) +
                                    ": operators not set (call pfc_hip_pfc_set_operators)");
  if (plan->tungsten() && plan->mean_split)
    return fail(PFC_HIP_EUNSUPPORTED,
                std::string(fn) + ": the mean-split mode applies to the PFC model only");
  return set_device(plan);
}

int pfc_hip_pfc_step(pfc_hip_plan_t plan, void *psi, int nsteps, void *stream) {
  if (int rc = step_checks(plan, psi, nsteps, "pfc_hip_pfc_step")) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (nsteps >= 2) {
    const bool f64 = plan->prec == PFC_HIP_F64;
    if (f64 ? fused_steps<double>(plan) : fused_steps<float>(plan)) {
      if (int rc = step_buffers(plan)) return rc;
      return f64 ? steps_fused<double>(plan, psi, nsteps, s, nullptr)
                 : steps_fused<float>(plan, psi, nsteps, s, nullptr);
    }
  }
  for (int i = 0; i < nsteps; ++i) {
    int rc = plan->prec == PFC_HIP_F64 ? step_impl<double>(plan, psi, s, nullptr)
                                       : step_impl<float>(plan, psi, s, nullptr);
    if (rc) return rc;
  }
  return 0;
}

int pfc_hip_tungsten_step(pfc_hip_plan_t plan, void *psi, int nsteps, void *psiMF,
                          void *stream) {
  if (int rc = step_checks(plan, psi, nsteps, "pfc_hip_tungsten_step")) return rc;
  if (!plan->tungsten())
    return fail(PFC_HIP_ESTATE, "pfc_hip_tungsten_step: Tungsten operators not set");
  if (psiMF && psiMF == psi)
    return fail(PFC_HIP_EINVAL, "pfc_hip_tungsten_step: psiMF must not alias psi");
  hipStream_t s = (hipStream_t)stream;
  for (int i = 0; i < nsteps; ++i) {
    void *mf = (i == nsteps - 1) ? psiMF : nullptr;
    int rc = plan->prec == PFC_HIP_F64 ? step_impl<double>(plan, psi, s, nullptr, mf)
                                       : step_impl<float>(plan, psi, s, nullptr, mf);
    if (rc) return rc;
  }
  return 0;
}

int pfc_hip_pfc_adjoint_step(pfc_hip_plan_t plan, const void *psi_n, void *lambda,
                             double *dr, void *stream) {
  if (int rc = check_single(plan, "pfc_hip_pfc_adjoint_step")) return rc;
  if (!psi_n || !lambda) return fail(PFC_HIP_EINVAL, "pfc_hip_pfc_adjoint_step: null field");
  if (psi_n == lambda)
    return fail(PFC_HIP_EINVAL, "pfc_hip_pfc_adjoint_step: lambda must not alias psi_n");
  if (!plan->ops_ready || plan->model_kind != PFC_HIP_MODEL_PFC)
    return fail(PFC_HIP_ESTATE, "pfc_hip_pfc_adjoint_step: PFC operators not set");
  if (int rc = set_device(plan)) return rc;
  hipStream_t s = (hipStream_t)stream;
  return plan->prec == PFC_HIP_F64 ? adjoint_impl<double>(plan, psi_n, lambda, dr, s)
                                   : adjoint_impl<float>(plan, psi_n, lambda, dr, s);
}

int pfc_hip_pfc_free_energy(pfc_hip_plan_t plan, const void *psi, double *F, void *stream) {
  if (int rc = check_single(plan, "pfc_hip_pfc_free_energy")) return rc;
  if (!psi || !F) return fail(PFC_HIP_EINVAL, "pfc_hip_pfc_free_energy: null argument");
  if (!plan->ops_ready || plan->model_kind != PFC_HIP_MODEL_PFC)
    return fail(PFC_HIP_ESTATE, "pfc_hip_pfc_free_energy: PFC operators not set");
  if (int rc = set_device(plan)) return rc;
  hipStream_t s = (hipStream_t)stream;
  return plan->prec == PFC_HIP_F64 ? free_energy_impl<double>(plan, psi, F, s)
                                   : free_energy_impl<float>(plan, psi, F, s);
}

int pfc_hip_stream_copy(const void *src, void *dst, size_t bytes, void *stream) {
  if (!src || !dst) return fail(PFC_HIP_EINVAL, "pfc_hip_stream_copy: null pointer");
  if (bytes % 16 || (uintptr_t)src % 16 || (uintptr_t)dst % 16)
    return fail(PFC_HIP_EINVAL, "pfc_hip_stream_copy: bytes and pointers must be 16-byte multiples");
  const char *a = (const char *)src, *b = (const char *)dst;
  if (a < b + bytes && b < a + bytes)
    return fail(PFC_HIP_EINVAL, "pfc_hip_stream_copy: src and dst overlap");
  const size_t n = bytes / 16;
  if (n == 0) return 0;
  const size_t blocks = (n + 255) / 256;
  if (blocks > 0x7fffffffu) return fail(PFC_HIP_EINVAL, "pfc_hip_stream_copy: too large");
  hipLaunchKernelGGL(k_stream_copy, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                     (const uint4 *)src, (uint4 *)dst, n);
  PFC_HIP_TRY(hipGetLastError(), "pfc_hip_stream_copy");
  return 0;
}

int pfc_hip_pfc_step_timed(pfc_hip_plan_t plan, void *psi, int nsteps,
                           void *stream, double *kernel_ms, int nkernels) {
  if (int rc = step_checks(plan, psi, nsteps, "pfc_hip_pfc_step_timed")) return rc;
  const int np = plan->step_passes();
  if (!kernel_ms || nkernels < np)
    return fail(PFC_HIP_EINVAL, "pfc_hip_pfc_step_timed: kernel_ms needs " +
                                    std::to_string(np) + " entries");
  if (nsteps < 1) return fail(PFC_HIP_EINVAL, "pfc_hip_pfc_step_timed: nsteps < 1");
  hipStream_t s = (hipStream_t)stream;
  const bool f64 = plan->prec == PFC_HIP_F64;
  // with a sixth entry and a fused run: the launches of pfc_hip_pfc_step
  const bool fused = nkernels >= 6 && nsteps >= 2 &&
                     (f64 ? fused_steps<double>(plan) : fused_steps<float>(plan));
  const int ne = fused ? 4 * nsteps + 2 : (np + 1) * nsteps;
  std::vector<hipEvent_t> ev((size_t)ne, nullptr);
  int rc = 0;
  for (auto &e : ev) {
    hipError_t he = hipEventCreate(&e);
    if (he != hipSuccess) {
      e = nullptr;
      rc = fail(PFC_HIP_EDEVICE, std::string("hipEventCreate: ") + hipGetErrorString(he));
      break;
    }
  }
  if (!rc && fused) {
    rc = step_buffers(plan);
    if (!rc)
      rc = f64 ? steps_fused<double>(plan, psi, nsteps, s, ev.data())
               : steps_fused<float>(plan, psi, nsteps, s, ev.data());
  } else if (!rc) {
    const int pe = np + 1;
    for (int i = 0; i < nsteps && !rc; ++i)
      rc = f64 ? step_impl<double>(plan, psi, s, &ev[(size_t)pe * i])
               : step_impl<float>(plan, psi, s, &ev[(size_t)pe * i]);
  }
  if (!rc) {
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = fail(PFC_HIP_EDEVICE, hipGetErrorString(e));
  }
  auto span = [&](size_t i) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ev[i], ev[i + 1]);
    return (double)ms;
  };
  if (!rc && fused) {
    // launch j of step i is event interval 4*i + j; the exit is 4*nsteps
    for (int k = 0; k < 6; ++k) kernel_ms[k] = 0;
    kernel_ms[0] = span(0);
    for (int i = 0; i < nsteps; ++i) {
      if (i > 0) kernel_ms[5] += span((size_t)4 * i) / (nsteps - 1);
      for (int k = 1; k <= 3; ++k) kernel_ms[k] += span((size_t)4 * i + k) / nsteps;
    }
    kernel_ms[4] = span((size_t)4 * nsteps);
  } else if (!rc) {
    const int pe = np + 1;
    for (int k = 0; k < np; ++k) kernel_ms[k] = 0;
    for (int i = 0; i < nsteps; ++i)
      for (int k = 0; k < np; ++k) kernel_ms[k] += span((size_t)pe * i + k) / nsteps;
    if (nkernels > np) kernel_ms[np] = 0;
  }
  for (auto &e : ev)
    if (e) (void)hipEventDestroy(e);
  return rc;
}

int pfc_hip_ic_constant(pfc_hip_plan_t plan, void *psi, double value,
                        void *stream) {
  if (int rc = check_single(plan, "pfc_hip_ic_constant")) return rc;
  if (!psi) return fail(PFC_HIP_EINVAL, "pfc_hip_ic_constant: null psi");
  if (int rc = set_device(plan)) return rc;
  const long n = (long)plan->inbox();
  hipStream_t s = (hipStream_t)stream;
  if (plan->prec == PFC_HIP_F64)
    hipLaunchKernelGGL(k_ic_constant<double>, dim3(ic_blocks(n)), dim3(256),
                       0, s, (double *)psi, n, value);
  else
    hipLaunchKernelGGL(k_ic_constant<float>, dim3(ic_blocks(n)), dim3(256), 0,
                       s, (float *)psi, n, value);
  PFC_HIP_TRY(hipGetLastError(), "ic_constant launch");
  return 0;
}

int pfc_hip_ic_single_seed(pfc_hip_plan_t plan, void *psi, const double *origin3,
                           const double *spacing3, double rho, double amp,
                           double radius, int keep_outside, void *stream) {
  if (int rc = check_single(plan, "pfc_hip_ic_single_seed")) return rc;
  if (!psi || !origin3 || !spacing3)
    return fail(PFC_HIP_EINVAL, "pfc_hip_ic_single_seed: null argument");
  if (int rc = set_device(plan)) return rc;
  SeedArgs a{};
  a.x0 = origin3[0];
  a.y0 = origin3[1];
  a.z0 = origin3[2];
  a.dx = spacing3[0];
  a.dy = spacing3[1];
  a.dz = spacing3[2];
  a.rho = rho;
  a.amp = amp;
  a.r2 = radius * radius;
  a.nx = plan->nx;
  a.ny = plan->ny;
  a.nz = plan->nz;
  a.keep_outside = keep_outside ? 1 : 0;
  const long n = (long)plan->inbox();
  hipStream_t s = (hipStream_t)stream;
  if (plan->prec == PFC_HIP_F64)
    hipLaunchKernelGGL(k_ic_single_seed<double>, dim3(ic_blocks(n)),
                       dim3(256), 0, s, (double *)psi, a);
  else
    hipLaunchKernelGGL(k_ic_single_seed<float>, dim3(ic_blocks(n)), dim3(256),
                       0, s, (float *)psi, a);
  PFC_HIP_TRY(hipGetLastError(), "ic_single_seed launch");
  return 0;
}

int pfc_hip_ic_add_noise(pfc_hip_plan_t plan, void *psi, unsigned long long seed,
                         double lo, double hi, void *stream) {
  if (int rc = check_single(plan, "pfc_hip_ic_add_noise")) return rc;
  if (!psi) return fail(PFC_HIP_EINVAL, "pfc_hip_ic_add_noise: null psi");
  if (int rc = set_device(plan)) return rc;
  const long n = (long)plan->inbox();
  hipStream_t s = (hipStream_t)stream;
  if (plan->prec == PFC_HIP_F64)
    hipLaunchKernelGGL(k_ic_add_noise<double>, dim3(ic_blocks(n)), dim3(256),
                       0, s, (double *)psi, n, seed, lo, hi);
  else
    hipLaunchKernelGGL(k_ic_add_noise<float>, dim3(ic_blocks(n)), dim3(256), 0,
                       s, (float *)psi, n, seed, lo, hi);
  PFC_HIP_TRY(hipGetLastError(), "ic_add_noise launch");
  return 0;
}

int pfc_hip_bc_fixed(pfc_hip_plan_t plan, void *psi, double x0, double dx, double rho_low,
                     double rho_high, double xwidth, double alpha, void *stream) {
  if (int rc = check_single(plan, "pfc_hip_bc_fixed")) return rc;
  if (!psi) return fail(PFC_HIP_EINVAL, "pfc_hip_bc_fixed: null psi");
  if (!std::isfinite(x0) || !std::isfinite(dx) || !std::isfinite(xwidth) ||
      !std::isfinite(alpha))
    return fail(PFC_HIP_EINVAL, "pfc_hip_bc_fixed: non-finite geometry");
  if (int rc = set_device(plan)) return rc;
  FixedBcArgs a{};
  a.x0 = x0;
  a.dx = dx;
  a.xwidth = xwidth;
  a.alpha = alpha;
  a.rho_low = rho_low;
  a.rho_high = rho_high;
  {
#pragma clang fp contract(off)
    a.xpos = ((double)plan->nx * dx) - xwidth; // fixed_bc.hpp: (Lx * dx) - xwidth
  }
  a.nx = plan->nx;
  a.rows = (long)plan->ny * plan->nz;
  const long n = (long)plan->inbox();
  hipStream_t s = (hipStream_t)stream;
  if (plan->prec == PFC_HIP_F64)
    hipLaunchKernelGGL(k_bc_fixed<double>, dim3(ic_blocks(n)), dim3(256), 0, s,
                       (double *)psi, a);
  else
    hipLaunchKernelGGL(k_bc_fixed<float>, dim3(ic_blocks(n)), dim3(256), 0, s,
                       (float *)psi, a);
  PFC_HIP_TRY(hipGetLastError(), "bc_fixed launch");
  return 0;
}

int pfc_hip_ic_seeds(pfc_hip_plan_t plan, void *psi, const double *origin3,
                     const double *spacing3, const double *seeds, int nseeds, int keep_outside,
                     void *stream) {
  if (int rc = check_single(plan, "pfc_hip_ic_seeds")) return rc;
  if (!psi || !origin3 || !spacing3 || (nseeds > 0 && !seeds))
    return fail(PFC_HIP_EINVAL, "pfc_hip_ic_seeds: null argument");
  if (nseeds < 0) return fail(PFC_HIP_EINVAL, "pfc_hip_ic_seeds: nseeds < 0");
  for (int d = 0; d < 3; ++d)
    if (!std::isfinite(origin3[d]) || !(spacing3[d] > 0) || !std::isfinite(spacing3[d]))
      return fail(PFC_HIP_EINVAL, "pfc_hip_ic_seeds: non-finite origin or spacing");
  for (long v = 0; v < 24L * nseeds; ++v)
    if (!std::isfinite(seeds[v]))
      return fail(PFC_HIP_EINVAL, "pfc_hip_ic_seeds: non-finite seed value");
  if (int rc = set_device(plan)) return rc;
  hipStream_t s = (hipStream_t)stream;
  const bool f64 = plan->prec == PFC_HIP_F64;
  if (!keep_outside) {
    const long n = (long)plan->inbox();
    if (f64)
      hipLaunchKernelGGL(k_ic_constant<double>, dim3(ic_blocks(n)), dim3(256), 0, s,
                         (double *)psi, n, 0.0);
    else
      hipLaunchKernelGGL(k_ic_constant<float>, dim3(ic_blocks(n)), dim3(256), 0, s,
                         (float *)psi, n, 0.0);
    PFC_HIP_TRY(hipGetLastError(), "ic_seeds clear");
  }
  const int ext[3] = {plan->nx, plan->ny, plan->nz};
  // last seed first: the first seed holding a point writes it last
  for (int sd = nseeds - 1; sd >= 0; --sd) {
    const double *v = seeds + 24L * sd;
    SeedBoxArgs a{};
    a.x0 = origin3[0];
    a.y0 = origin3[1];
    a.z0 = origin3[2];
    a.dx = spacing3[0];
    a.dy = spacing3[1];
    a.dz = spacing3[2];
    for (int d = 0; d < 3; ++d) a.loc[d] = v[d];
    for (int q = 0; q < 18; ++q) a.q[q] = v[3 + q];
    a.rho = v[21];
    a.r2 = v[22] * v[22];
    a.amp = v[23];
    a.nx = plan->nx;
    a.ny = plan->ny;
    // index box of the sphere, one point of margin each side, clipped
    int lo[3], hi[3];
    bool empty = false;
    for (int d = 0; d < 3; ++d) {
      const double r = std::fabs(v[22]);
      const double f0 = std::floor((v[d] - r - origin3[d]) / spacing3[d]) - 1.0;
      const double f1 = std::ceil((v[d] + r - origin3[d]) / spacing3[d]) + 1.0;
      lo[d] = (int)std::max(0.0, std::min(f0, (double)ext[d]));
      hi[d] = (int)std::min((double)ext[d] - 1.0, std::max(f1, -1.0));
      if (hi[d] < lo[d]) empty = true;
    }
    if (empty) continue;
    a.i0 = lo[0];
    a.j0 = lo[1];
    a.k0 = lo[2];
    a.bx = hi[0] - lo[0] + 1;
    a.by = hi[1] - lo[1] + 1;
    a.nbox = (long)a.bx * a.by * (hi[2] - lo[2] + 1);
    if (f64)
      hipLaunchKernelGGL(k_ic_seed_box<double>, dim3(ic_blocks(a.nbox)), dim3(256), 0, s,
                         (double *)psi, a);
    else
      hipLaunchKernelGGL(k_ic_seed_box<float>, dim3(ic_blocks(a.nbox)), dim3(256), 0, s,
                         (float *)psi, a);
    PFC_HIP_TRY(hipGetLastError(), "ic_seeds launch");
  }
  return 0;
}

int pfc_hip_column_max(pfc_hip_plan_t plan, const void *psi, double *xline, void *stream) {
  if (int rc = check_single(plan, "pfc_hip_column_max")) return rc;
  if (!psi || !xline) return fail(PFC_HIP_EINVAL, "pfc_hip_column_max: null argument");
  if (int rc = set_device(plan)) return rc;
  const int nx = plan->nx;
  const long rows = (long)plan->ny * plan->nz;
  const int parts = (int)std::min<long>(rows, kColParts);
  if (int rc = dalloc(plan, &plan->colmax, ((size_t)kColParts + 1) * nx * sizeof(double),
                      "column maxima"))
    return rc;
  double *part = (double *)plan->colmax;
  double *out = part + (size_t)kColParts * nx;
  hipStream_t s = (hipStream_t)stream;
  const unsigned gx = (unsigned)((nx + 255) / 256);
  if (plan->prec == PFC_HIP_F64)
    hipLaunchKernelGGL(k_column_max_part<double>, dim3(gx, (unsigned)parts), dim3(256), 0, s,
                       (const double *)psi, nx, rows, part);
  else
    hipLaunchKernelGGL(k_column_max_part<float>, dim3(gx, (unsigned)parts), dim3(256), 0, s,
                       (const float *)psi, nx, rows, part);
  PFC_HIP_TRY(hipGetLastError(), "column max (partials)");
  hipLaunchKernelGGL(k_column_max_final, dim3(gx), dim3(256), 0, s, (const double *)part, nx,
                     parts, out);
  PFC_HIP_TRY(hipGetLastError(), "column max (final)");
  PFC_HIP_TRY(hipMemcpyAsync(xline, out, (size_t)nx * sizeof(double), hipMemcpyDeviceToHost, s),
              "column max copy");
  PFC_HIP_TRY(hipStreamSynchronize(s), "column max sync");
  return 0;
}

// The same maxima without allocating: partials in the library's static
// scratch, result in the caller's device buffer, nothing waited for.
int pfc_hip_column_max_noalloc(pfc_hip_plan_t plan, const void *psi, double *xline_dev,
                               void *stream) {
  if (int rc = check_single(plan, "pfc_hip_column_max_noalloc")) return rc;
  if (!psi || !xline_dev)
    return fail(PFC_HIP_EINVAL, "pfc_hip_column_max_noalloc: null argument");
  const int nx = plan->nx;
  const long rows = (long)plan->ny * plan->nz;
  const int parts = (int)std::min<long>(rows, kColParts);
  if ((long)parts * nx > kColScratch)
    return fail(PFC_HIP_EINVAL, "pfc_hip_column_max_noalloc: nx = " + std::to_string(nx) +
                                    " needs " + std::to_string((long)parts * nx) +
                                    " doubles of scratch, more than its " +
                                    std::to_string(kColScratch) + " (nx <= " +
                                    std::to_string(kColScratch / kColParts) + ")");
  if (int rc = set_device(plan)) return rc;
  double *part = nullptr;
  if (int rc = col_scratch(plan->device, &part)) return rc;
  hipStream_t s = (hipStream_t)stream;
  const unsigned gx = (unsigned)((nx + 255) / 256);
  if (plan->prec == PFC_HIP_F64)
    hipLaunchKernelGGL(k_column_max_part<double>, dim3(gx, (unsigned)parts), dim3(256), 0, s,
                       (const double *)psi, nx, rows, part);
  else
    hipLaunchKernelGGL(k_column_max_part<float>, dim3(gx, (unsigned)parts), dim3(256), 0, s,
                       (const float *)psi, nx, rows, part);
  PFC_HIP_TRY(hipGetLastError(), "column max (partials)");
  hipLaunchKernelGGL(k_column_max_final, dim3(gx), dim3(256), 0, s, (const double *)part, nx,
                     parts, xline_dev);
  PFC_HIP_TRY(hipGetLastError(), "column max (final)");
  return 0;
}

int pfc_hip_bc_moving(pfc_hip_plan_t plan, void *psi, double x0, double dx, double xpos,
                      double rho_low, double rho_high, double xwidth, double alpha,
                      void *stream) {
  if (int rc = check_single(plan, "pfc_hip_bc_moving")) return rc;
  if (!psi) return fail(PFC_HIP_EINVAL, "pfc_hip_bc_moving: null psi");
  if (!std::isfinite(x0) || !std::isfinite(dx) || !(dx > 0) || !std::isfinite(xpos) ||
      !std::isfinite(xwidth) || !std::isfinite(alpha))
    return fail(PFC_HIP_EINVAL, "pfc_hip_bc_moving: non-finite geometry");
  if (int rc = set_device(plan)) return rc;
  MovingBcArgs a{};
  a.x0 = x0;
  a.dx = dx;
  a.xwidth = xwidth;
  a.alpha = alpha;
  a.rho_low = rho_low;
  a.rho_high = rho_high;
  a.nx = plan->nx;
  {
#pragma clang fp contract(off)
    const double Lx = (double)plan->nx;
    a.l = Lx * dx;                 // moving_bc.hpp: l = Lx * dx
    a.xpos = std::fmod(xpos, a.l); // xpos = fmod(m_xpos, l)
  }
  // the columns in a band, by the kernel's own test (host and device round
  // these few operations identically)
  int first = -1, last = -1;
  for (int i = 0; i < plan->nx; ++i) {
#pragma clang fp contract(off)
    const double x = (x0 + ((double)i * dx));
    const double dist = x - a.xpos;
    const bool in = std::fabs(dist) < xwidth ||
                    (a.xpos < xwidth && std::fabs(dist - a.l) < xwidth) ||
                    (a.xpos > a.l - xwidth && std::fabs(dist + a.l) < xwidth);
    if (in) {
      if (first < 0) first = i;
      last = i;
    }
  }
  if (first < 0) return 0;
  a.i0 = first;
  a.bx = last - first + 1;
  a.nbox = (long)a.bx * plan->ny * plan->nz;
  hipStream_t s = (hipStream_t)stream;
  if (plan->prec == PFC_HIP_F64)
    hipLaunchKernelGGL(k_bc_moving<double>, dim3(ic_blocks(a.nbox)), dim3(256), 0, s,
                       (double *)psi, a);
  else
    hipLaunchKernelGGL(k_bc_moving<float>, dim3(ic_blocks(a.nbox)), dim3(256), 0, s,
                       (float *)psi, a);
  PFC_HIP_TRY(hipGetLastError(), "bc_moving launch");
  return 0;
}

// ---------------------------------------------------------------------------
// Launch variants and autotuning.
// ---------------------------------------------------------------------------
int pfc_hip_variant_count(pfc_hip_plan_t plan, int role, int *count) {
  if (int rc = check_plan(plan)) return rc;
  if (!count || role < 0 || role >= ROLE_COUNT)
    return fail(PFC_HIP_EINVAL, "pfc_hip_variant_count: bad role or null count");
  *count = plan->prec == PFC_HIP_F64 ? variant_count<double>(plan, role)
                                     : variant_count<float>(plan, role);
  return 0;
}

const char *pfc_hip_variant_name(pfc_hip_plan_t plan, int role, int index) {
  if (!plan || role < 0 || role >= ROLE_COUNT) return nullptr;
  const std::string *n = plan->prec == PFC_HIP_F64
                             ? variant_name<double>(plan, role, index)
                             : variant_name<float>(plan, role, index);
  return n ? n->c_str() : nullptr;
}

int pfc_hip_variant_get(pfc_hip_plan_t plan, int role, int *index) {
  if (int rc = check_plan(plan)) return rc;
  if (!index || role < 0 || role >= ROLE_COUNT)
    return fail(PFC_HIP_EINVAL, "pfc_hip_variant_get: bad role or null index");
  *index = plan->var[role];
  return 0;
}

int pfc_hip_variant_set(pfc_hip_plan_t plan, int role, int index) {
  if (int rc = check_plan(plan)) return rc;
  if (role < 0 || role >= ROLE_COUNT)
    return fail(PFC_HIP_EINVAL, "pfc_hip_variant_set: bad role");
  int count = 0;
  pfc_hip_variant_count(plan, role, &count);
  if (index < 0 || index >= count)
    return fail(PFC_HIP_EINVAL, "pfc_hip_variant_set: index " + std::to_string(index) +
                                    " out of range (" + std::to_string(count) +
                                    " variants)");
  if (role == ROLE_STEP_LAYOUT && index != plan->var[role]) {
    if (int rc = set_device(plan)) return rc;
    (void)hipDeviceSynchronize();
    void **drop[4] = {&plan->A1, &plan->B1, &plan->A2, &plan->B2};
    void **drop_nat[2] = {&plan->A, &plan->B};
    const size_t tb = plan->outbox() * 2 * plan->esize;
    const size_t nb = plan->padded() * 2 * plan->esize * plan->batch;
    if (index == 1) {
      for (void **b : drop)
        if (*b) { (void)hipFree(*b); *b = nullptr; plan->bytes -= tb; }
    } else {
      for (void **b : drop_nat)
        if (*b) { (void)hipFree(*b); *b = nullptr; plan->bytes -= nb; }
    }
  }
  if (plan->var[role] != index) ++plan->gen;
  plan->var[role] = index;
  if (role == ROLE_STEP_LAYOUT && plan->ops_ready) return step_buffers(plan);
  return 0;
}

int pfc_hip_pfc_autotune(pfc_hip_plan_t plan, void *psi, int reps, void *stream,
                         double *times_ms, int max_times) {
  if (int rc = check_plan(plan)) return rc;
  if (!psi || reps < 1)
    return fail(PFC_HIP_EINVAL, "pfc_hip_pfc_autotune: null psi or reps < 1");
  if (!plan->ops_ready)
    return fail(PFC_HIP_ESTATE, "pfc_hip_pfc_autotune: operators not set");
  if (int rc = set_device(plan)) return rc;
  hipStream_t s = (hipStream_t)stream;
  // psi is restored at the end from the c2r workspace (outbox complex >= inbox real)
  if (int rc = dalloc(plan, &plan->work, plan->outbox() * 2 * plan->esize * plan->batch,
                      "c2r workspace"))
    return rc;
  const size_t rbytes = plan->inbox() * plan->esize * plan->batch;
  PFC_HIP_TRY(hipMemcpyAsync(plan->work, psi, rbytes, hipMemcpyDeviceToDevice, s),
              "autotune backup");
  const int np = plan->step_passes();
  std::vector<hipEvent_t> ev((size_t)np + 1, nullptr);
  int rc = 0;
  for (auto &e : ev) {
    hipError_t he = hipEventCreate(&e);
    if (he != hipSuccess) {
      e = nullptr;
      rc = fail(PFC_HIP_EDEVICE, std::string("hipEventCreate: ") + hipGetErrorString(he));
      break;
    }
  }
  // Tuned roles in slot order, with the passes (bit k = pass k of the step)
  // their kernels occupy; the last two slots receive each layout's step time.
  struct RoleSpan { int role, layout; unsigned passes; };
  static const RoleSpan pfc_spans[] = {
      {ROLE_PFC_X, 1, 0x1},   {ROLE_COL_Y, 1, 0xA},    {ROLE_PFC_Z, 1, 0x4},
      {ROLE_X_C2R, 1, 0x10},  {ROLE_T_X_FWD, 0, 0x1},  {ROLE_T_Y_FWD, 0, 0x2},
      {ROLE_T_Z, 0, 0x4},     {ROLE_T_Y_INV, 0, 0x8},  {ROLE_T_X_INV, 0, 0x10}};
  static const RoleSpan tung_spans[] = {
      {ROLE_X_R2C, 1, 0x1},   {ROLE_COL_Y, 1, 0xAA},   {ROLE_PFC_Z, 1, 0x44},
      {ROLE_X_TUNG, 1, 0x10}, {ROLE_X_C2R, 1, 0x100},  {ROLE_T_X_R2C, 0, 0x1},
      {ROLE_T_Y_FWD, 0, 0x22}, {ROLE_T_Z, 0, 0x44},    {ROLE_T_Y_INV, 0, 0x88},
      {ROLE_T_X_TUNG, 0, 0x10}, {ROLE_T_X_INV, 0, 0x100}};
  const RoleSpan *spans = plan->tungsten() ? tung_spans : pfc_spans;
  const int nspans = plan->tungsten() ? (int)(sizeof(tung_spans) / sizeof(RoleSpan))
                                      : (int)(sizeof(pfc_spans) / sizeof(RoleSpan));
  auto run_step = [&](double *pass_ms) -> int {
    int r = plan->prec == PFC_HIP_F64 ? step_impl<double>(plan, psi, s, ev.data())
                                      : step_impl<float>(plan, psi, s, ev.data());
    if (r) return r;
    hipError_t he = hipEventSynchronize(ev[np]);
    if (he != hipSuccess) return fail(PFC_HIP_EDEVICE, hipGetErrorString(he));
    for (int k = 0; k < np; ++k) {
      float ms = 0;
      (void)hipEventElapsedTime(&ms, ev[k], ev[k + 1]);
      pass_ms[k] = ms;
    }
    return 0;
  };
  auto span_ms = [&](const double *pm, unsigned passes) {
    double t = 0;
    for (int k = 0; k < np; ++k)
      if (passes & (1u << k)) t += pm[k];
    return t;
  };
  int slot = 0;
  double layout_ms[2] = {0, 0};
  double role_best[ROLE_COUNT];
  for (double &b : role_best) b = 1e300;
  // both layouts when the plan has both (natural first, slot order); a
  // double-butterfly plan has the transposed one only
  const bool two = plan->t_ok && !plan->xp;
  const int layouts = two ? 2 : 1;
  for (int li = 0; li < layouts && !rc; ++li) {
    const int layout = two ? 1 - li : (plan->t_ok ? 0 : 1);
    rc = pfc_hip_variant_set(plan, ROLE_STEP_LAYOUT, two ? layout : 0);
    for (int si = 0; si < nspans && !rc; ++si) {
      const RoleSpan &sp = spans[si];
      if (sp.layout != layout) continue;
      int count = 0;
      pfc_hip_variant_count(plan, sp.role, &count);
      // Each variant is scored by its fastest rep (interference only adds
      // time); the kTop best are then timed again, round-robin, so that a
      // drift of the clock over the sweep cannot decide between close ones.
      std::vector<double> score((size_t)count, 1e300);
      auto time_variant = [&](int v) {
        plan->var[sp.role] = v;
        double pm[9];
        for (int r = 0; r <= reps && !rc; ++r) { // r == 0 warms up
          rc = run_step(pm);
          if (!rc && r > 0) score[(size_t)v] = std::min(score[(size_t)v], span_ms(pm, sp.passes));
        }
      };
      for (int v = 0; v < count && !rc; ++v) time_variant(v);
      constexpr int kTop = 3;
      std::vector<int> order((size_t)count);
      for (int v = 0; v < count; ++v) order[(size_t)v] = v;
      std::stable_sort(order.begin(), order.end(),
                       [&](int a, int b) { return score[(size_t)a] < score[(size_t)b]; });
      const int top = std::min(kTop, count);
      for (int round = 0; round < 2 && top > 1 && !rc; ++round)
        for (int i = 0; i < top && !rc; ++i) time_variant(order[(size_t)i]);
      int best_v = plan->var[sp.role];
      double best = 1e300;
      for (int v = 0; v < count; ++v) {
        const int idx = slot + v;
        if (times_ms && idx < max_times) times_ms[idx] = score[(size_t)v];
        if (score[(size_t)v] < best) {
          best = score[(size_t)v];
          best_v = v;
        }
      }
      plan->var[sp.role] = best_v;
      role_best[sp.role] = best;
      slot += count;
    }
    if (!rc && layout == 0 && !plan->tungsten()) {
      // the fused step boundary of multi-step runs, timed on its own: each
      // variant starts from the T1 spectrum a full step leaves (the pass
      // rewrites T1 in place), one warm-up launch, then `reps` timed ones
      int count = 0;
      pfc_hip_variant_count(plan, ROLE_T_X_FUSE, &count);
      double best = 1e300;
      int best_v = plan->var[ROLE_T_X_FUSE];
      for (int v = 0; v < count && !rc; ++v) {
        plan->var[ROLE_T_X_FUSE] = v;
        double pm[9], fastest = 1e300;
        rc = run_step(pm);
        for (int r = 0; r <= reps && !rc; ++r) {
          hipError_t he = hipEventRecord(ev[0], s);
          if (he != hipSuccess) {
            rc = fail(PFC_HIP_EDEVICE, std::string("autotune event: ") + hipGetErrorString(he));
            break;
          }
          rc = plan->prec == PFC_HIP_F64
                   ? t_pass<double>(plan, pfc_t_args<double>(plan, psi), 5, s)
                   : t_pass<float>(plan, pfc_t_args<float>(plan, psi), 5, s);
          if (rc) break;
          he = hipEventRecord(ev[1], s);
          if (he == hipSuccess) he = hipEventSynchronize(ev[1]);
          if (he != hipSuccess) {
            rc = fail(PFC_HIP_EDEVICE, std::string("autotune sync: ") + hipGetErrorString(he));
            break;
          }
          float ms = 0;
          (void)hipEventElapsedTime(&ms, ev[0], ev[1]);
          if (r > 0) fastest = std::min(fastest, (double)ms);
        }
        if (times_ms && slot + v < max_times) times_ms[slot + v] = fastest;
        if (fastest < best) {
          best = fastest;
          best_v = v;
        }
      }
      plan->var[ROLE_T_X_FUSE] = best_v;
      slot += count;
      // fuse step boundaries when one fused pass beats the two it replaces
      // by a clear margin (the two timings come from different launches)
      constexpr double kFuseMargin = 0.95;
      if (count > 0)
        plan->var[ROLE_STEP_FUSE] =
            best < kFuseMargin * (role_best[ROLE_T_X_FWD] + role_best[ROLE_T_X_INV]) ? 1 : 0;
    }
    double fastest = 1e300, pm[9];
    for (int r = 0; r <= reps && !rc; ++r) {
      rc = run_step(pm);
      if (!rc && r > 0) fastest = std::min(fastest, span_ms(pm, ~0u));
    }
    // a batched plan's transposed step times its field 0 (the fields run
    // one after another); the natural one times every field at once
    layout_ms[layout] = layout == 0 ? fastest * plan->batch : fastest;
  }
  if (!rc && two) {
    rc = pfc_hip_variant_set(plan, ROLE_STEP_LAYOUT,
                             layout_ms[0] <= layout_ms[1] ? 0 : 1);
  }
  if (times_ms && slot + 1 < max_times) {
    times_ms[slot] = layout_ms[0];
    times_ms[slot + 1] = layout_ms[1];
  }
  for (auto &e : ev)
    if (e) (void)hipEventDestroy(e);
  ++plan->gen; // the variants of every tuned role may have changed
  // psi is restored on every exit path once the backup was enqueued; a
  // failure of the sweep keeps its own code and message
  const std::string sweep_error = rc ? g_last_error : std::string();
  hipError_t he = hipMemcpyAsync(psi, plan->work, rbytes, hipMemcpyDeviceToDevice, s);
  if (he == hipSuccess) he = hipStreamSynchronize(s);
  if (rc) {
    g_last_error = sweep_error + (he == hipSuccess ? std::string(" (psi restored)")
                                                   : std::string(" (psi not restored: ") +
                                                         hipGetErrorString(he) + ")");
    return rc;
  }
  PFC_HIP_TRY(he, "autotune restore");
  return 0;
}

int pfc_hip_pfc_set_members(pfc_hip_plan_t plan, const double *eps, const double *mean,
                            int mean_split) {
  if (int rc = check_plan(plan)) return rc;
  if (!plan->ops_ready || plan->model_kind != PFC_HIP_MODEL_PFC)
    return fail(PFC_HIP_ESTATE, "pfc_hip_pfc_set_members: PFC operators not set");
  if (mean_split && !mean)
    return fail(PFC_HIP_EINVAL, "pfc_hip_pfc_set_members: mean_split needs the means");
  const int nb = plan->batch;
  std::vector<double> e((size_t)nb, plan->op.eps), m((size_t)nb, 0.0);
  for (int b = 0; b < nb; ++b) {
    if (eps) e[(size_t)b] = eps[b];
    if (mean_split) m[(size_t)b] = mean[b];
    if (!std::isfinite(e[(size_t)b]) || !std::isfinite(m[(size_t)b]))
      return fail(PFC_HIP_EINVAL, "pfc_hip_pfc_set_members: value of field " +
                                      std::to_string(b) + " not finite");
  }
  if (int rc = set_device(plan)) return rc;
  if (nb == 1) { // a plain plan: the scalars are the members
    plan->op.eps = e[0];
    plan->mean_split = mean_split != 0;
    plan->mean = m[0];
    ++plan->gen; // launch arguments changed
    return 0;
  }
  if (int rc = dalloc(plan, &plan->m_eps, (size_t)nb * sizeof(double), "field eps")) return rc;
  if (int rc = dalloc(plan, &plan->m_nl, (size_t)nb * 2 * plan->esize, "field means"))
    return rc;
  // (3m, 3m^2) rounded once to the field type, as for a single field
  std::vector<double> c2d((size_t)2 * nb);
  std::vector<float> c2f((size_t)2 * nb);
  for (int b = 0; b < nb; ++b) {
    const double mb = m[(size_t)b];
    c2d[(size_t)2 * b] = 3.0 * mb;
    c2d[(size_t)2 * b + 1] = 3.0 * mb * mb;
    c2f[(size_t)2 * b] = (float)c2d[(size_t)2 * b];
    c2f[(size_t)2 * b + 1] = (float)c2d[(size_t)2 * b + 1];
  }
  // the stream order of later steps is kept by the synchronous copies
  PFC_HIP_TRY(hipDeviceSynchronize(), "pfc_hip_pfc_set_members: sync");
  PFC_HIP_TRY(hipMemcpy(plan->m_eps, e.data(), (size_t)nb * sizeof(double),
                        hipMemcpyHostToDevice),
              "pfc_hip_pfc_set_members: eps upload");
  if (plan->prec == PFC_HIP_F64)
    PFC_HIP_TRY(hipMemcpy(plan->m_nl, c2d.data(), c2d.size() * sizeof(double),
                          hipMemcpyHostToDevice),
                "pfc_hip_pfc_set_members: mean upload");
  else
    PFC_HIP_TRY(hipMemcpy(plan->m_nl, c2f.data(), c2f.size() * sizeof(float),
                          hipMemcpyHostToDevice),
                "pfc_hip_pfc_set_members: mean upload");
  plan->h_eps = e;
  plan->h_mean = m;
  // the natural layout reads eps and the means from device memory; the
  // transposed one takes them, and the mean-split mode, as launch arguments
  ++plan->gen;
  plan->members = true;
  plan->mean_split = mean_split != 0;
  plan->mean = 0.0;
  return 0;
}

int pfc_hip_pfc_set_mean(pfc_hip_plan_t plan, int enable, double mean) {
  if (int rc = check_single(plan, "pfc_hip_pfc_set_mean")) return rc;
  if (enable && plan->tungsten())
    return fail(PFC_HIP_EUNSUPPORTED,
                "pfc_hip_pfc_set_mean: the mean-split mode applies to the PFC model only");
  if (!std::isfinite(mean)) return fail(PFC_HIP_EINVAL, "pfc_hip_pfc_set_mean: mean not finite");
  plan->mean_split = enable != 0;
  plan->mean = enable ? mean : 0.0;
  ++plan->gen;
  return 0;
}

} // extern "C"

namespace pfc_amd {
// Launch-argument generation of a plan (ensemble.hip re-captures its graph
// when a group plan's generation moved since the capture).
unsigned plan_generation(const pfc_hip_plan_s *p) { return p ? p->gen : 0u; }
} // namespace pfc_amd
