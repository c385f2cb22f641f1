// pfc_hip.hip -- C-ABI implementation: plans, twiddle tables, size dispatch
// and the launch sequences of the R2C/C2R transforms and the fused PFC step.
#include "../../include/pfc_hip.h"
#include "fft_kernels.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

using namespace pfc_amd;

namespace {

thread_local std::string g_last_error;

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
// Compile-time plans per line length.
// ---------------------------------------------------------------------------
template <int N> struct PlanFor;
template <> struct PlanFor<1> { using P = FftPlan<1, 1, RadixList<1>>; };
template <> struct PlanFor<2> { using P = FftPlan<2, 2, RadixList<2>>; };
template <> struct PlanFor<4> { using P = FftPlan<4, 4, RadixList<4>>; };
template <> struct PlanFor<8> { using P = FftPlan<8, 8, RadixList<8>>; };
template <> struct PlanFor<16> { using P = FftPlan<16, 16, RadixList<16>>; };
template <> struct PlanFor<32> { using P = FftPlan<32, 8, RadixList<8, 4>>; };
template <> struct PlanFor<64> { using P = FftPlan<64, 8, RadixList<8, 8>>; };
template <> struct PlanFor<128> { using P = FftPlan<128, 16, RadixList<16, 8>>; };
template <> struct PlanFor<256> { using P = FftPlan<256, 16, RadixList<16, 16>>; };
template <> struct PlanFor<512> { using P = FftPlan<512, 8, RadixList<8, 8, 8>>; };
template <> struct PlanFor<1024> {
  using P = FftPlan<1024, 16, RadixList<16, 16, 4>>;
};

// Radices of each plan, for the host twiddle builder.
template <typename List> struct RadixVec;
template <int... Rs> struct RadixVec<RadixList<Rs...>> {
  static std::vector<int> get() { return {Rs...}; }
};

bool supported_len(int n) {
  switch (n) {
  case 1: case 2: case 4: case 8: case 16: case 32: case 64: case 128:
  case 256: case 512: case 1024:
    return true;
  default:
    return false;
  }
}

std::vector<int> radices_of(int n) {
  switch (n) {
  case 1: return RadixVec<PlanFor<1>::P::List>::get();
  case 2: return RadixVec<PlanFor<2>::P::List>::get();
  case 4: return RadixVec<PlanFor<4>::P::List>::get();
  case 8: return RadixVec<PlanFor<8>::P::List>::get();
  case 16: return RadixVec<PlanFor<16>::P::List>::get();
  case 32: return RadixVec<PlanFor<32>::P::List>::get();
  case 64: return RadixVec<PlanFor<64>::P::List>::get();
  case 128: return RadixVec<PlanFor<128>::P::List>::get();
  case 256: return RadixVec<PlanFor<256>::P::List>::get();
  case 512: return RadixVec<PlanFor<512>::P::List>::get();
  case 1024: return RadixVec<PlanFor<1024>::P::List>::get();
  default: return {};
  }
}

// Tile width in columns: 128 contiguous bytes per column-tile row.
template <typename T, typename Plan> constexpr int col_tile() {
  constexpr int l = 128 / (2 * (int)sizeof(T));
  return (l * Plan::TL > 1024) ? 1024 / Plan::TL : l;
}
// Exchange one plane at a time when a full exchange would exceed 64 KiB.
template <typename T, typename Plan> constexpr bool col_split() {
  return 2 * col_tile<T, Plan>() * Plan::N * (int)sizeof(T) > 65536;
}
// Rows per block for row kernels: one wave per block.
template <typename Plan> constexpr int rows_per_block() {
  return Plan::TL >= 64 ? 1 : 64 / Plan::TL;
}

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

// Line twiddles for the Stockham passes: exp(-2 pi i q k/(Ns R)).
std::vector<long double> line_twiddles(int n) {
  std::vector<long double> out; // re, im pairs
  int ns = 1;
  for (int r : radices_of(n)) {
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
  // device tables: twiddles of the nx-, nx/2-, ny-, nz-point lines and the
  // real-FFT twiddles of nx (all in the plan's precision)
  void *tw_xn = nullptr, *tw_xm = nullptr, *tw_y = nullptr, *tw_z = nullptr;
  void *wn = nullptr;
  void *work = nullptr; // c2r workspace: W*ny*nz complex
  // fused-step state
  void *A = nullptr, *B = nullptr, *opL = nullptr, *opN = nullptr;
  bool ops_ready = false;
  int model_kind = -1;
  double dt = 0;
  size_t bytes = 0;

  size_t inbox() const { return (size_t)nx * ny * nz; }
  size_t outbox() const { return (size_t)W * ny * nz; }
  size_t padded() const { return (size_t)P * ny * nz; }
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

void free_all(pfc_hip_plan_s *p) {
  void **bufs[] = {&p->tw_xn, &p->tw_xm, &p->tw_y, &p->tw_z, &p->wn,
                   &p->work,  &p->A,     &p->B,    &p->opL,  &p->opN};
  for (void **b : bufs) {
    if (*b) (void)hipFree(*b);
    *b = nullptr;
  }
  p->bytes = 0;
}

// ---------------------------------------------------------------------------
// Launchers (templated on precision and line length).
// ---------------------------------------------------------------------------
template <typename T, int N> struct Launch {
  using Plan = typename PlanFor<N>::P;
  static constexpr int L = col_tile<T, Plan>();
  static constexpr bool SPLIT = col_split<T, Plan>();
  static constexpr int LPB = rows_per_block<Plan>();

  static hipError_t col(int dir, const ColArgs<T> &a, long n_outer, int nfields,
                        hipStream_t s) {
    dim3 grid((unsigned)(a.ntx * n_outer), (unsigned)nfields);
    dim3 block(L * Plan::TL);
    if (dir < 0)
      hipLaunchKernelGGL((k_col<T, Plan, -1, L, SPLIT>), grid, block, 0, s, a);
    else
      hipLaunchKernelGGL((k_col<T, Plan, +1, L, SPLIT>), grid, block, 0, s, a);
    return hipGetLastError();
  }
  static hipError_t x_r2c(const XArgs<T> &a, hipStream_t s) {
    dim3 grid((unsigned)((a.rows + LPB - 1) / LPB)), block(LPB * Plan::TL);
    hipLaunchKernelGGL((k_x_r2c<T, Plan, LPB>), grid, block, 0, s, a);
    return hipGetLastError();
  }
  static hipError_t x_c2r(const XArgs<T> &a, hipStream_t s) {
    dim3 grid((unsigned)((a.rows + LPB - 1) / LPB)), block(LPB * Plan::TL);
    hipLaunchKernelGGL((k_x_c2r<T, Plan, LPB>), grid, block, 0, s, a);
    return hipGetLastError();
  }
  static hipError_t pfc_x(const PfcXArgs<T> &a, hipStream_t s) {
    dim3 grid((unsigned)((a.rows + LPB - 1) / LPB)), block(LPB * Plan::TL);
    hipLaunchKernelGGL((k_pfc_x<T, Plan, LPB>), grid, block, 0, s, a);
    return hipGetLastError();
  }
  static hipError_t pfc_z(const PfcZArgs<T> &a, long ny, hipStream_t s) {
    dim3 grid((unsigned)(a.ntx * ny)), block(L * Plan::TL);
    hipLaunchKernelGGL((k_pfc_z<T, Plan, L, SPLIT>), grid, block, 0, s, a);
    return hipGetLastError();
  }
  static int tile() { return L; }
};

template <typename T, typename F> hipError_t with_len(int n, F &&f) {
  switch (n) {
  case 1: return f(Launch<T, 1>{});
  case 2: return f(Launch<T, 2>{});
  case 4: return f(Launch<T, 4>{});
  case 8: return f(Launch<T, 8>{});
  case 16: return f(Launch<T, 16>{});
  case 32: return f(Launch<T, 32>{});
  case 64: return f(Launch<T, 64>{});
  case 128: return f(Launch<T, 128>{});
  case 256: return f(Launch<T, 256>{});
  case 512: return f(Launch<T, 512>{});
  case 1024: return f(Launch<T, 1024>{});
  default: return hipErrorInvalidValue;
  }
}

// Column pass along y (axis 1) or z (axis 2) over a (W, ny, nz) array with
// row pitch `pitch`; fields 0/1 in src/dst.
template <typename T>
hipError_t col_pass(pfc_hip_plan_s *p, int axis, int dir, const void *src0,
                    void *dst0, const void *src1, void *dst1, int nfields,
                    int pitch, hipStream_t s) {
  const int n = axis == 1 ? p->ny : p->nz;
  ColArgs<T> a{};
  a.src[0] = (const cplx<T> *)src0;
  a.dst[0] = (cplx<T> *)dst0;
  a.src[1] = (const cplx<T> *)src1;
  a.dst[1] = (cplx<T> *)dst1;
  a.tw = (const cplx<T> *)(axis == 1 ? p->tw_y : p->tw_z);
  a.W = p->W;
  a.s_pos = axis == 1 ? (long)pitch : (long)pitch * p->ny;
  a.s_outer = axis == 1 ? (long)pitch * p->ny : (long)pitch;
  const long n_outer = axis == 1 ? p->nz : p->ny;
  return with_len<T>(n, [&](auto l) {
    using Lx = decltype(l);
    a.ntx = (p->W + Lx::tile() - 1) / Lx::tile();
    return Lx::col(dir, a, n_outer, nfields, s);
  });
}

template <typename T>
int r2c_impl(pfc_hip_plan_s *p, const void *in, void *out, hipStream_t s) {
  XArgs<T> a{};
  a.in = (const T *)in;
  a.out = (T *)out;
  a.tw = (const cplx<T> *)p->tw_xm;
  a.wn = (const cplx<T> *)p->wn;
  a.rows = (long)p->ny * p->nz;
  a.pitch = p->W;
  a.scale = (T)1;
  hipError_t e = with_len<T>(p->nx / 2, [&](auto l) { return l.x_r2c(a, s); });
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

template <typename T>
int c2r_impl(pfc_hip_plan_s *p, const void *in, void *out, hipStream_t s) {
  const void *cur = in;
  hipError_t e;
  if (p->nz > 1 || p->ny > 1) {
    if (int rc = dalloc(p, &p->work, p->outbox() * 2 * p->esize, "c2r workspace"))
      return rc;
  }
  if (p->nz > 1) {
    e = col_pass<T>(p, 2, +1, cur, p->work, nullptr, nullptr, 1, p->W, s);
    PFC_HIP_TRY(e, "c2r z pass");
    cur = p->work;
  }
  if (p->ny > 1) {
    e = col_pass<T>(p, 1, +1, cur, p->work, nullptr, nullptr, 1, p->W, s);
    PFC_HIP_TRY(e, "c2r y pass");
    cur = p->work;
  }
  XArgs<T> a{};
  a.in = (const T *)cur;
  a.out = (T *)out;
  a.tw = (const cplx<T> *)p->tw_xm;
  a.wn = (const cplx<T> *)p->wn;
  a.rows = (long)p->ny * p->nz;
  a.pitch = p->W;
  a.scale = (T)(1.0 / ((double)p->nx * (double)p->ny * (double)p->nz));
  e = with_len<T>(p->nx / 2, [&](auto l) { return l.x_c2r(a, s); });
  PFC_HIP_TRY(e, "c2r x pass");
  return 0;
}

// One fused PFC step: 5 launches.  `ev` (optional) receives 6 events
// bracketing the 5 passes.
template <typename T>
int step_impl(pfc_hip_plan_s *p, void *psi, hipStream_t s, hipEvent_t *ev) {
  hipError_t e;
  auto mark = [&](int i) {
    if (ev) (void)hipEventRecord(ev[i], s);
  };
  mark(0);
  {
    PfcXArgs<T> a{};
    a.psi = (const T *)psi;
    a.outA = (cplx<T> *)p->A;
    a.outB = (cplx<T> *)p->B;
    a.tw = (const cplx<T> *)p->tw_xn;
    a.rows = (long)p->ny * p->nz;
    a.pitch = p->P;
    e = with_len<T>(p->nx, [&](auto l) { return l.pfc_x(a, s); });
    PFC_HIP_TRY(e, "pfc forward x pass");
  }
  mark(1);
  if (p->ny > 1) {
    e = col_pass<T>(p, 1, -1, p->A, p->A, p->B, p->B, 2, p->P, s);
    PFC_HIP_TRY(e, "pfc forward y pass");
  }
  mark(2);
  {
    PfcZArgs<T> a{};
    a.A = (cplx<T> *)p->A;
    a.B = (const cplx<T> *)p->B;
    a.opL = (const T *)p->opL;
    a.opN = (const T *)p->opN;
    a.tw = (const cplx<T> *)p->tw_z;
    a.s_pos = (long)p->P * p->ny;
    a.pitch = p->P;
    a.W = p->W;
    e = with_len<T>(p->nz, [&](auto l) {
      using Lx = decltype(l);
      a.ntx = (p->W + Lx::tile() - 1) / Lx::tile();
      return Lx::pfc_z(a, p->ny, s);
    });
    PFC_HIP_TRY(e, "pfc z pass");
  }
  mark(3);
  if (p->ny > 1) {
    e = col_pass<T>(p, 1, +1, p->A, p->A, nullptr, nullptr, 1, p->P, s);
    PFC_HIP_TRY(e, "pfc inverse y pass");
  }
  mark(4);
  {
    XArgs<T> a{};
    a.in = (const T *)p->A;
    a.out = (T *)psi;
    a.tw = (const cplx<T> *)p->tw_xm;
    a.wn = (const cplx<T> *)p->wn;
    a.rows = (long)p->ny * p->nz;
    a.pitch = p->P;
    a.scale = (T)1; // 1/(nx ny nz) is folded into opL/opN
    e = with_len<T>(p->nx / 2, [&](auto l) { return l.x_c2r(a, s); });
    PFC_HIP_TRY(e, "pfc inverse x pass");
  }
  mark(5);
  return 0;
}

unsigned ic_blocks(long n) {
  return (unsigned)std::min<long>((n + 255) / 256, 1L << 20);
}

int check_plan(pfc_hip_plan_t p) {
  if (!p) return fail(PFC_HIP_EINVAL, "null plan");
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
  if (flags != 0)
    return fail(PFC_HIP_EINVAL, "pfc_hip_plan_create: flags must be 0");
  if (nx % 2 != 0 || !supported_len(nx / 2) || !supported_len(ny) ||
      !supported_len(nz))
    return fail(PFC_HIP_EUNSUPPORTED,
                "pfc_hip_plan_create: unsupported extents " + std::to_string(nx) +
                    "x" + std::to_string(ny) + "x" + std::to_string(nz) +
                    " (each axis a power of two <= 1024, nx even and >= 2)");
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
  if (!rc) rc = upload_table(p, &p->tw_xm, line_twiddles(nx / 2), "x twiddles");
  if (!rc && supported_len(nx))
    rc = upload_table(p, &p->tw_xn, line_twiddles(nx), "x twiddles (full)");
  if (!rc) rc = upload_table(p, &p->tw_y, line_twiddles(ny), "y twiddles");
  if (!rc) rc = upload_table(p, &p->tw_z, line_twiddles(nz), "z twiddles");
  if (!rc) rc = upload_table(p, &p->wn, real_twiddles(nx), "real twiddles");
  if (rc) {
    free_all(p);
    delete p;
    return rc;
  }
  *plan = p;
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
  if (int rc = check_plan(plan)) return rc;
  if (!in || !out) return fail(PFC_HIP_EINVAL, "pfc_hip_r2c: null buffer");
  if (int rc = set_device(plan)) return rc;
  hipStream_t s = (hipStream_t)stream;
  return plan->prec == PFC_HIP_F64 ? r2c_impl<double>(plan, in, out, s)
                                   : r2c_impl<float>(plan, in, out, s);
}

int pfc_hip_c2r(pfc_hip_plan_t plan, const void *in, void *out, void *stream) {
  if (int rc = check_plan(plan)) return rc;
  if (!in || !out) return fail(PFC_HIP_EINVAL, "pfc_hip_c2r: null buffer");
  if (int rc = set_device(plan)) return rc;
  hipStream_t s = (hipStream_t)stream;
  return plan->prec == PFC_HIP_F64 ? c2r_impl<double>(plan, in, out, s)
                                   : c2r_impl<float>(plan, in, out, s);
}

int pfc_hip_pfc_set_operators(pfc_hip_plan_t plan, int model_kind,
                              const double *params, int nparams, double dt) {
  if (int rc = check_plan(plan)) return rc;
  if (model_kind != PFC_HIP_MODEL_PFC)
    return fail(PFC_HIP_EUNSUPPORTED, "pfc_hip_pfc_set_operators: unknown model " +
                                          std::to_string(model_kind));
  if (!params || nparams != 4)
    return fail(PFC_HIP_EINVAL,
                "pfc_hip_pfc_set_operators: PFC needs params {dx, dy, dz, eps}");
  if (!supported_len(plan->nx))
    return fail(PFC_HIP_EUNSUPPORTED,
                "pfc_hip_pfc_set_operators: nx unsupported for the fused step");
  const double dx = params[0], dy = params[1], dz = params[2], eps = params[3];
  if (!(dx > 0) || !(dy > 0) || !(dz > 0) || !std::isfinite(eps) ||
      !std::isfinite(dt))
    return fail(PFC_HIP_EINVAL, "pfc_hip_pfc_set_operators: bad spacing/eps/dt");
  if (int rc = set_device(plan)) return rc;
  const size_t nc = plan->padded();
  int rc = 0;
  if (!rc) rc = dalloc(plan, &plan->A, nc * 2 * plan->esize, "step spectrum A");
  if (!rc) rc = dalloc(plan, &plan->B, nc * 2 * plan->esize, "step spectrum B");
  if (!rc) rc = dalloc(plan, &plan->opL, nc * plan->esize, "opL");
  if (!rc) rc = dalloc(plan, &plan->opN, nc * plan->esize, "opN");
  if (rc) return rc;
  const double two_pi = 6.283185307179586;
  PfcOpArgs a{};
  a.fx = two_pi / (dx * plan->nx);
  a.fy = two_pi / (dy * plan->ny);
  a.fz = two_pi / (dz * plan->nz);
  a.eps = eps;
  a.dt = dt;
  a.scale = 1.0 / ((double)plan->nx * (double)plan->ny * (double)plan->nz);
  a.nx = plan->nx;
  a.ny = plan->ny;
  a.nz = plan->nz;
  a.W = plan->W;
  a.pitch = plan->P;
  // zero the padding columns once so they stay finite
  PFC_HIP_TRY(hipMemset(plan->opL, 0, nc * plan->esize), "memset opL");
  PFC_HIP_TRY(hipMemset(plan->opN, 0, nc * plan->esize), "memset opN");
  const long total = (long)plan->outbox();
  const unsigned blocks = (unsigned)std::min<long>((total + 255) / 256, 1L << 20);
  if (plan->prec == PFC_HIP_F64)
    hipLaunchKernelGGL(k_pfc_operators<double>, dim3(blocks), dim3(256), 0, 0, a,
                       (double *)plan->opL, (double *)plan->opN);
  else
    hipLaunchKernelGGL(k_pfc_operators<float>, dim3(blocks), dim3(256), 0, 0, a,
                       (float *)plan->opL, (float *)plan->opN);
  PFC_HIP_TRY(hipGetLastError(), "operator kernel launch");
  PFC_HIP_TRY(hipDeviceSynchronize(), "operator kernel");
  plan->model_kind = model_kind;
  plan->dt = dt;
  plan->ops_ready = true;
  return 0;
}

int pfc_hip_pfc_step(pfc_hip_plan_t plan, void *psi, int nsteps, void *stream) {
  if (int rc = check_plan(plan)) return rc;
  if (!psi) return fail(PFC_HIP_EINVAL, "pfc_hip_pfc_step: null psi");
  if (nsteps < 0) return fail(PFC_HIP_EINVAL, "pfc_hip_pfc_step: nsteps < 0");
  if (!plan->ops_ready)
    return fail(PFC_HIP_ESTATE,
                "pfc_hip_pfc_step: operators not set (call pfc_hip_pfc_set_operators)");
  if (int rc = set_device(plan)) return rc;
  hipStream_t s = (hipStream_t)stream;
  for (int i = 0; i < nsteps; ++i) {
    int rc = plan->prec == PFC_HIP_F64 ? step_impl<double>(plan, psi, s, nullptr)
                                       : step_impl<float>(plan, psi, s, nullptr);
    if (rc) return rc;
  }
  return 0;
}

int pfc_hip_pfc_step_timed(pfc_hip_plan_t plan, void *psi, int nsteps,
                           void *stream, double *kernel_ms, int nkernels) {
  if (int rc = check_plan(plan)) return rc;
  if (!psi || !kernel_ms || nkernels < 5)
    return fail(PFC_HIP_EINVAL, "pfc_hip_pfc_step_timed: bad arguments");
  if (nsteps < 1) return fail(PFC_HIP_EINVAL, "pfc_hip_pfc_step_timed: nsteps < 1");
  if (!plan->ops_ready)
    return fail(PFC_HIP_ESTATE, "pfc_hip_pfc_step_timed: operators not set");
  if (int rc = set_device(plan)) return rc;
  hipStream_t s = (hipStream_t)stream;
  std::vector<hipEvent_t> ev(6 * (size_t)nsteps);
  for (auto &e : ev) PFC_HIP_TRY(hipEventCreate(&e), "hipEventCreate");
  int rc = 0;
  for (int i = 0; i < nsteps && !rc; ++i)
    rc = plan->prec == PFC_HIP_F64 ? step_impl<double>(plan, psi, s, &ev[6 * i])
                                   : step_impl<float>(plan, psi, s, &ev[6 * i]);
  if (!rc) {
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = fail(PFC_HIP_EDEVICE, hipGetErrorString(e));
  }
  if (!rc) {
    for (int k = 0; k < 5; ++k) kernel_ms[k] = 0;
    for (int i = 0; i < nsteps; ++i)
      for (int k = 0; k < 5; ++k) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, ev[6 * i + k], ev[6 * i + k + 1]);
        kernel_ms[k] += ms / nsteps;
      }
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  return rc;
}

int pfc_hip_ic_constant(pfc_hip_plan_t plan, void *psi, double value,
                        void *stream) {
  if (int rc = check_plan(plan)) return rc;
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
  if (int rc = check_plan(plan)) return rc;
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
  if (int rc = check_plan(plan)) return rc;
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

} // extern "C"
