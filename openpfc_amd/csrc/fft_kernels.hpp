// fft_kernels.hpp -- HBM-streaming passes of the 3-D R2C/C2R transform and
// of the fused PFC step, built on the line FFTs of fft_device.hpp.
//
// Layout (the reference's r2c_direction = 0 contract, fft.cpp:58-70):
//   real field     psi[z][y][x], x fastest, nx*ny*nz values
//   half spectrum  F[z][y][kx], kx in [0, nx/2], row pitch P >= nx/2+1
//                  (P = nx/2+1 for caller buffers, padded to a 128-byte
//                  multiple for the plan's own step buffers)
//
// Passes
//   k_x_r2c     one row per line: the real row is read as nx/2 complex
//               values, transformed, and split into the half spectrum.
//   k_x_c2r     the inverse: half spectrum row -> nx/2-point inverse ->
//               interleaved real row (optionally scaled).
//   k_col       complex transform of L adjacent columns along y or z; the
//               lanes of a wave cover (column, position) pairs so each load
//               and store moves L*sizeof(complex) contiguous bytes.
//   k_pfc_x     fused step, forward x: psi row -> (psi, psi^3) packed as one
//               complex row -> nx-point transform -> both half spectra.
//   k_pfc_z     fused step, z: forward z of both spectra, ETD combine with
//               opL/opN, inverse z, written back in place.
#pragma once

#include "fft_device.hpp"

namespace pfc_amd {

// XCD-aware block order.  A block's G lines (or L columns) are adjacent in
// the strided dimension, so when G values of cplx<T> fill less than a 128-byte line, the
// P = 128 / (G * sizeof(cplx<T>)) blocks of consecutive line groups share
// every line they touch there.  Workgroups go to the 8 XCDs round-robin
// (physical block b on XCD b mod 8) and each XCD has its own L2, so those P
// blocks are made to run on one XCD, back to back: physical block
// 8*(P*q + r) + x executes logical block P*(8*q + x) + r.  The last partial
// group of 8*P blocks keeps its order; the map is a bijection on [0, n).
template <typename T, int G> constexpr unsigned line_partners() {
  constexpr int bytes = G * 2 * (int)sizeof(T);
  return bytes >= 128 ? 1u : 128u / bytes;
}
template <typename T, int G>
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned n) {
  constexpr unsigned P = line_partners<T, G>(), X = 8, S = X * P;
  if constexpr (P <= 1) {
    return b;
  } else {
    if (b >= n / S * S) return b;
    const unsigned x = b % X, r = b / X;
    return ((r / P) * X + x) * P + (r % P);
  }
}

// ----------------------------------------------------------------------------
// Column transform along y or z.
//   column (x, outer) starts at outer*s_outer + x; position p at + p*s_pos.
//   gridDim.x = ntx * n_outer, gridDim.y = number of fields (1 or 2).
// ----------------------------------------------------------------------------
template <typename T> struct ColArgs {
  const cplx<T> *src[2];
  cplx<T> *dst[2];
  const cplx<T> *tw;
  long s_pos;
  long s_outer;
  int W;   // active columns in x
  int ntx; // tiles in x
};

template <typename T, typename Plan, int DIR, int L, bool SPLIT>
__global__ void __launch_bounds__(L *Plan::TL)
    k_col(ColArgs<T> a) {
  constexpr int TL = Plan::TL, E = Plan::E, N = Plan::N;
  // position-major, column-interleaved LDS with one pad position per 16:
  // the stride-16 writes of a radix-16 pass then spread over all banks
  constexpr int NP = row_padded(N);
  __shared__ T lds[(SPLIT ? 1 : 2) * L * NP];
  const int f = blockIdx.y;
  const cplx<T> *__restrict__ src = a.src[f];
  cplx<T> *__restrict__ dst = a.dst[f];
  const int c = threadIdx.x % L, t = threadIdx.x / L;
  const unsigned bid = xcd_block<T, L>(blockIdx.x, gridDim.x);
  const int tx = (int)(bid % (unsigned)a.ntx);
  const long outer = bid / (unsigned)a.ntx;
  const int x = tx * L + c;
  const bool active = x < a.W;
  const long base = outer * a.s_outer + x;

  cplx<T> v[E];
#pragma unroll
  for (int e = 0; e < E; ++e)
    v[e] = active ? src[base + (long)(t + TL * e) * a.s_pos] : cplx<T>{0, 0};
  const LdsLine<T> ll{lds, c, L, L * NP};
  LineFft<T, Plan, DIR, SPLIT, true>::run(v, t, ll, a.tw);
  if (active) {
#pragma unroll
    for (int e = 0; e < E; ++e) dst[base + (long)(t + TL * e) * a.s_pos] = v[e];
  }
}

// ----------------------------------------------------------------------------
// Row helpers.  Padded LDS slot of line position t + TL*e (0 <= t < TL) and of
// its mirror (L - k for k = t + TL*e != 0, 0 for k = 0; L = E*TL the line
// length), each as a per-lane part plus a per-e constant when TL is a
// multiple of 16, so that the accesses of all e share one address register:
// pad_pos(u - t) = u + u/16 - t - ceil(t/16) for u = L - TL*e, a multiple of
// 16 and larger than t.
// ----------------------------------------------------------------------------
template <int TL> __device__ __forceinline__ int row_slot(int t, int e) {
  if constexpr (TL % 16 == 0)
    return t + (t >> 4) + e * (TL + TL / 16);
  else
    return pad_pos<true>(t + TL * e);
}
template <int E, int TL> __device__ __forceinline__ int row_mirror_slot(int t, int e) {
  constexpr int L = E * TL;
  if constexpr (TL % 16 == 0) {
    const int u = L - TL * e;
    const int i = u + u / 16 - t - ((t + 15) >> 4);
    return e == 0 ? (t == 0 ? 0 : i) : i;
  } else {
    const int k = t + TL * e;
    return pad_pos<true>(k == 0 ? 0 : L - k);
  }
}

// Fetch, for every owned element k = t + TL*e, the value at the mirror
// position (L - k, or 0 for k = 0) of the same line, through LDS.
template <typename T, int E, int TL, bool WAVE = false>
__device__ __forceinline__ void row_fetch_partner(const cplx<T> (&v)[E],
                                                  cplx<T> (&p)[E], int t,
                                                  T *lds_re, T *lds_im) {
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int i = row_slot<TL>(t, e);
    lds_re[i] = v[e].x;
    lds_im[i] = v[e].y;
  }
  lds_sync<WAVE>();
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int i = row_mirror_slot<E, TL>(t, e);
    p[e] = cplx<T>{lds_re[i], lds_im[i]};
  }
  lds_sync<WAVE>();
}

// ----------------------------------------------------------------------------
// Half-length real transforms.  A real row x of length nx = 2M is read as the
// M complex values z[m] = x[2m] + i*x[2m+1] and transformed with an M-point
// complex FFT Z.  r2c_split forms the half spectrum
//   X[k] = (Z[k] + conj Z[M-k])/2 + w^k * (-i/2)(Z[k] - conj Z[M-k]),
// w = exp(-2 pi i/nx), from zk = Z[k] and zmk = Z[(M-k) mod M];
// X[M] = Re Z[0] - Im Z[0].  c2r_merge is its inverse: from X[k] and X[M-k]
// it forms Z'[k] = (X[k] + conj X[M-k]) + i*conj(w^k)*(X[k] - conj X[M-k]),
// whose M-point inverse is the unscaled real row (imaginary parts of X[0]
// and X[M] ignored, as a half-complex inverse does).
// ----------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ cplx<T> r2c_split(cplx<T> zk, cplx<T> zmk, cplx<T> w) {
  const cplx<T> zc = c_conj(zmk);
  const cplx<T> ev = c_scale(c_add(zk, zc), (T)0.5);
  const cplx<T> d = c_sub(zk, zc);
  const cplx<T> od = {(T)0.5 * d.y, (T)-0.5 * d.x}; // -i/2 * d
  return c_add(ev, c_mul(w, od));
}

template <typename T>
__device__ __forceinline__ cplx<T> c2r_merge(cplx<T> xa, cplx<T> xmk, cplx<T> w,
                                             bool k0) {
  cplx<T> xb = c_conj(xmk);
  if (k0) {
    xa.y = (T)0;
    xb.y = (T)0;
  }
  const cplx<T> ev = c_add(xa, xb);
  const cplx<T> od = c_mulc(c_sub(xa, xb), w); // * exp(+2 pi i k/nx)
  return cplx<T>{ev.x - od.y, ev.y + od.x};    // ev + i*od
}

// ----------------------------------------------------------------------------
// x-axis R2C: real rows of length nx = 2M -> half spectrum rows (M+1 values).
// Plan is the M-point complex transform; LPB rows per block.
// ----------------------------------------------------------------------------
template <typename T> struct XArgs {
  const T *in;       // real rows (r2c) / spectrum rows (c2r)
  T *out;            // spectrum rows (r2c) / real rows (c2r)
  const cplx<T> *tw; // line twiddles
  const cplx<T> *wn; // exp(-2 pi i k / nx), k = 0..M
  long rows;
  int pitch; // spectrum row pitch (complex elements)
  T scale;   // c2r output scale
};

template <typename T, typename Plan, int LPB>
__global__ void __launch_bounds__(LPB *Plan::TL) k_x_r2c(XArgs<T> a) {
  constexpr int TL = Plan::TL, E = Plan::E, M = Plan::N;
  constexpr int RP = row_padded(M);
  __shared__ T lds[2 * LPB * RP];
  const int t = threadIdx.x % TL, c = threadIdx.x / TL;
  const long row = (long)blockIdx.x * LPB + c;
  const bool active = row < a.rows;
  const long r = active ? row : 0;
  const cplx<T> *__restrict__ in =
      reinterpret_cast<const cplx<T> *>(a.in + r * (2L * M));
  cplx<T> *__restrict__ out = reinterpret_cast<cplx<T> *>(a.out) + r * a.pitch;

  cplx<T> v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = in[t + TL * e];
  T *re = lds + c * RP;
  T *im = lds + LPB * RP + c * RP;
  const LdsLine<T> ll{lds, c * RP, 1, LPB * RP};
  LineFft<T, Plan, -1, false, true, row_wave<Plan>()>::run(v, t, ll, a.tw);

  cplx<T> pz[E];
  row_fetch_partner<T, E, TL, row_wave<Plan>()>(v, pz, t, re, im);
  if (!active) return;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int k = t + TL * e;
    out[k] = r2c_split(v[e], pz[e], a.wn[k]);
  }
  if (t == 0) {
    // X[M] = Re Z0 - Im Z0 (v[0] holds Z0 on thread 0)
    out[M] = cplx<T>{v[0].x - v[0].y, (T)0};
  }
}

// ----------------------------------------------------------------------------
// x-axis C2R: half spectrum rows -> real rows (scaled).  The imaginary parts
// of X[0] and X[M] are ignored, as a half-complex inverse does.
// ----------------------------------------------------------------------------
template <typename T, typename Plan, int LPB>
__global__ void __launch_bounds__(LPB *Plan::TL) k_x_c2r(XArgs<T> a) {
  constexpr int TL = Plan::TL, E = Plan::E, M = Plan::N;
  constexpr int RP = row_padded(M);
  __shared__ T lds[2 * LPB * RP];
  const int t = threadIdx.x % TL, c = threadIdx.x / TL;
  const long row = (long)blockIdx.x * LPB + c;
  const bool active = row < a.rows;
  const long r = active ? row : 0;
  const cplx<T> *__restrict__ in =
      reinterpret_cast<const cplx<T> *>(a.in) + r * a.pitch;
  cplx<T> *__restrict__ out = reinterpret_cast<cplx<T> *>(a.out + r * (2L * M));

  cplx<T> v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int k = t + TL * e;
    v[e] = c2r_merge(in[k], in[M - k], a.wn[k], k == 0);
  }
  const LdsLine<T> ll{lds, c * RP, 1, LPB * RP};
  LineFft<T, Plan, +1, false, true, row_wave<Plan>()>::run(v, t, ll, a.tw);
  if (!active) return;
#pragma unroll
  for (int e = 0; e < E; ++e) out[t + TL * e] = c_scale(v[e], a.scale);
}

// ----------------------------------------------------------------------------
// Fused PFC step, forward x pass.  One row of psi (nx = N real values) is
// packed as z = psi + i*psi^3, transformed with an N-point complex FFT and
// split into Psi[k] = (Z[k] + conj Z[N-k])/2 and Nl[k] = (Z[k] - conj Z[N-k])/(2i)
// for k = 0..N/2.  The nonlinearity is psi*psi*psi, evaluated in that order
// (10_complete_pfc_model.cpp:131-133).
// ----------------------------------------------------------------------------
// Nonlinearity of the step.  Plain: psi*psi*psi in the reference's order
// (10_complete_pfc_model.cpp:131-133).  Mean-split (the field holds the
// deviation d = psi - m from an exactly conserved mean m): d*(3m^2 + (3m + d)*d),
// which equals psi^3 - m^3; the constant m^3 only reaches k = 0, where opN
// vanishes, so every other mode is unchanged.
template <typename T> struct Nonlin {
  T c1, c2; // 3m, 3m^2 (used when split != 0)
  int split;
  double c1d, c2d; // the same constants unrounded (the double evaluation)
  __device__ __forceinline__ T operator()(T u) const {
    return split ? u * (c2 + (c1 + u) * u) : u * u * u;
  }
  // in double, from the unrounded constants (FP64-butterfly kernels)
  __device__ __forceinline__ double exact(double u) const {
    return split ? u * (c2d + (c1d + u) * u) : u * u * u;
  }
};

template <typename T> struct PfcXArgs {
  Nonlin<T> nl;
  const T *mnl; // batched plans: per-field (3m, 3m^2) pairs overriding nl.c1/c2, or null
  long mrows;   // rows per field (ny*nz) when mnl is set
  const T *psi;
  cplx<T> *outA; // Psi half spectrum
  cplx<T> *outB; // N half spectrum
  const cplx<T> *tw;
  long rows;
  int pitch;
};

template <typename T, typename Plan, int LPB>
__global__ void __launch_bounds__(LPB *Plan::TL) k_pfc_x(PfcXArgs<T> a) {
  constexpr int TL = Plan::TL, E = Plan::E, N = Plan::N;
  constexpr int RP = row_padded(N);
  __shared__ T lds[2 * LPB * RP];
  const int t = threadIdx.x % TL, c = threadIdx.x / TL;
  const long row = (long)blockIdx.x * LPB + c;
  const bool active = row < a.rows;
  const long r = active ? row : 0;
  const T *__restrict__ in = a.psi + r * (long)N;
  Nonlin<T> nl = a.nl;
  if (a.mnl) {
    const long b = r / a.mrows;
    nl.c1 = a.mnl[2 * b];
    nl.c2 = a.mnl[2 * b + 1];
  }

  cplx<T> v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const T u = in[t + TL * e];
    v[e] = cplx<T>{u, nl(u)};
  }
  T *re = lds + c * RP;
  T *im = lds + LPB * RP + c * RP;
  const LdsLine<T> ll{lds, c * RP, 1, LPB * RP};
  LineFft<T, Plan, -1, false, true, row_wave<Plan>()>::run(v, t, ll, a.tw);

  cplx<T> pz[E];
  row_fetch_partner<T, E, TL, row_wave<Plan>()>(v, pz, t, re, im);
  if (!active) return;
  cplx<T> *__restrict__ oa = a.outA + r * a.pitch;
  cplx<T> *__restrict__ ob = a.outB + r * a.pitch;
#pragma unroll
  for (int e = 0; e < E / 2; ++e) {
    const int k = t + TL * e;
    const cplx<T> zk = v[e], zc = c_conj(pz[e]);
    const cplx<T> s = c_add(zk, zc), d = c_sub(zk, zc);
    oa[k] = c_scale(s, (T)0.5);
    ob[k] = cplx<T>{(T)0.5 * d.y, (T)-0.5 * d.x};
  }
  if (t == 0) {
    // k = N/2 is owned by thread 0 as element E/2 (for E == 1, N == 1)
    constexpr int eh = E / 2;
    const cplx<T> zk = v[eh], zc = c_conj(pz[eh]);
    const cplx<T> s = c_add(zk, zc), d = c_sub(zk, zc);
    oa[N / 2] = c_scale(s, (T)0.5);
    ob[N / 2] = cplx<T>{(T)0.5 * d.y, (T)-0.5 * d.x};
  }
}

// Functions that restate reference arithmetic use `#pragma clang fp
// contract(off)` and plain operators, so every product and sum rounds as the
// reference's expression does (the _rn intrinsics are plain operators defined
// outside the pragma's scope and would still be fused into FMAs).
// ----------------------------------------------------------------------------
// PFC operators (docs/api/examples/10_complete_pfc_model.cpp:166-213) for one
// mode, in the reference's formula and evaluation order:
//   k2 = kx*kx + ky*ky + kz*kz, L = -k2*(eps + (1-k2)^2),
//   opL = exp(L*dt), opN = (opL-1)/L*(-k2)  (dt*(-k2) when |L| <= 1e-14),
// each multiplied by `scale` (the 1/(nx*ny*nz) of the backward transform).
// kxy = kx*kx + ky*ky is passed in so the sum keeps the reference's order.
// ----------------------------------------------------------------------------
struct PfcOp {
  double fx, fy, fz; // 2*pi/(d*n) per axis
  double eps, dt, scale;
  int nx, ny, nz;
};

__device__ __forceinline__ double pfc_kc(int i, int n, double f) {
  return (i <= n / 2) ? i * f : (i - n) * f;
}

__device__ __forceinline__ void pfc_op_mode(const PfcOp &o, double kxy, double kz,
                                            double &opL, double &opN) {
#pragma clang fp contract(off)
  const double k2 = (kxy + (kz * kz));
  const double k2_lap = -k2;
  const double one_m = 1.0 - k2;
  const double L = (k2_lap * (o.eps + (one_m * one_m)));
  const double oL = exp((L * o.dt));
  double oN;
  if (fabs(L) > 1e-14)
    oN = ((oL - 1.0) / L * k2_lap);
  else
    oN = (o.dt * k2_lap);
  opL = (oL * o.scale);
  opN = (oN * o.scale);
}

// kx*kx + ky*ky rounded as written (no contraction into an FMA).
__device__ __forceinline__ double pfc_kxy(const PfcOp &o, int i, int j) {
#pragma clang fp contract(off)
  const double kx = pfc_kc(i, o.nx, o.fx), ky = pfc_kc(j, o.ny, o.fy);
  return kx * kx + ky * ky;
}

// Operator tables in the outbox layout (pitch W), for inspection and tests.
template <typename T>
__global__ void k_pfc_operators(PfcOp o, int W, T *opL, T *opN) {
#pragma clang fp contract(off)
  const long total = (long)W * o.ny * o.nz;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const int i = (int)(idx % W);
    const long jk = idx / W;
    const int j = (int)(jk % o.ny);
    const int k = (int)(jk / o.ny);
    const double kxy = pfc_kxy(o, i, j);
    double l, n;
    pfc_op_mode(o, kxy, pfc_kc(k, o.nz, o.fz), l, n);
    opL[idx] = (T)l;
    opN[idx] = (T)n;
  }
}

// ----------------------------------------------------------------------------
// Tungsten mean-field PFC operators for one mode, in the reference's formula
// and evaluation order (apps/tungsten/include/tungsten/common/
// tungsten_spectral.hpp:83-104; tungsten_etd_workspace.hpp:99-130;
// include/openpfc/kernel/integrator/spectral_exp_coefficients.hpp:68-82):
//   kLap = -(kx*kx + ky*ky + kz*kz)
//   fMF = exp(kLap/lambda2)                        (mean-field filter chi(k))
//   k_val = sqrt(-kLap) - 1, k2 = k_val^2
//   gf = k_val < 0 ? exp(-(k2 + rTol*k_val^highOrd)/alpha2)   (highOrd != 0)
//                  : 1 - (1/alpha2)*k2
//   opCk = ((stabP + p2_bar) - Bx*exp(-T/T0)*gf) + q2_bar*fMF
//   L = kLap*opCk, expL = exp(L*dt),
//   nW = kLap * (|L| < 1e-12 ? dt + 0.5*L*dt*dt : expm1(L*dt)/L)
// (for L*dt < -1, where most modes are, expm1(L*dt) is taken as expL - 1,
// within 2 ulp of expm1 there, so one exponential serves both multipliers)
// The mode-independent scalars rTol = -alpha2*log(farTol) - 1,
// c_peak = Bx*exp(-T/T0) and sp2 = stabP + p2_bar are formed on the host in
// the same order.  Step multipliers are scaled by PfcOp::scale.
// ----------------------------------------------------------------------------
struct TungOp {
  double alpha2, lambda2, rTol, c_peak, sp2, q2;
  int highOrd;
};

__device__ __forceinline__ double tung_filter(const TungOp &p, double kxy, double kz) {
#pragma clang fp contract(off)
  const double kLap = -(kxy + (kz * kz));
  return exp(kLap / p.lambda2);
}

__device__ __forceinline__ void tung_op_mode(const PfcOp &o, const TungOp &p,
                                             double kxy, double kz, double &fMF,
                                             double &expL, double &nW) {
#pragma clang fp contract(off)
  const double kLap = -(kxy + (kz * kz));
  fMF = exp(kLap / p.lambda2);
  const double k_val = sqrt(-kLap) - 1.0;
  const double k2 = k_val * k_val;
  double gf;
  if (k_val < 0.0) {
    if (p.highOrd == 0)
      gf = exp((-k2) / p.alpha2);
    else
      gf = exp((-(k2 + (p.rTol * pow(k_val, (double)p.highOrd)))) / p.alpha2);
  } else {
    gf = 1.0 - ((1.0 / p.alpha2) * k2);
  }
  const double opCk = (p.sp2 - (p.c_peak * gf)) + (p.q2 * fMF);
  const double L = kLap * opCk;
  const double arg = L * o.dt;
  expL = exp(arg);
  double phi1;
  if (fabs(L) < 1e-12)
    phi1 = o.dt + (((0.5 * L) * o.dt) * o.dt);
  else if (arg < -1.0)
    // expm1(arg) from the exp above: |exp(arg) - 1| > 0.63 here, so the
    // subtraction adds at most one rounding (within 2 ulp of expm1)
    phi1 = (expL - 1.0) / L;
  else
    phi1 = expm1(arg) / L;
  nW = kLap * phi1;
}

// Operator tables of the Tungsten step in the outbox layout (pitch W),
// scaled by o.scale, for inspection and tests.
template <typename T>
__global__ void k_tung_operators(PfcOp o, TungOp p, int W, T *fMF, T *expL, T *nW) {
#pragma clang fp contract(off)
  const long total = (long)W * o.ny * o.nz;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const int i = (int)(idx % W);
    const long jk = idx / W;
    const int j = (int)(jk % o.ny);
    const int k = (int)(jk / o.ny);
    double f, e, n;
    tung_op_mode(o, p, pfc_kxy(o, i, j), pfc_kc(k, o.nz, o.fz), f, e, n);
    fMF[idx] = (T)(f * o.scale);
    expL[idx] = (T)(e * o.scale);
    nW[idx] = (T)(n * o.scale);
  }
}

// Tungsten nonlinearity (apps/tungsten/include/tungsten/cpu/tungsten_model.hpp:
// 331-348) with u = psi and v = psi_MF:
//   N = p3*u^2 + p4*u^3 + q3*v^2 + q4*v^3, then N - stabP*u when stabP != 0,
// every power and sum rounded in the reference's order.
template <typename T> struct TungNl {
  T p3, p4, q3, q4, stabP;
  __device__ __forceinline__ T operator()(T u, T v) const {
#pragma clang fp contract(off)
    const T u2 = u * u, u3 = (u * u) * u, v2 = v * v, v3 = (v * v) * v;
    T n = (((p3 * u2) + (p4 * u3)) + (q3 * v2)) + (q4 * v3);
    if (stabP != (T)0) n = n - (stabP * u);
    return n;
  }
};

// z-pass kinds.  ZM_PFC: forward both spectra, ETD combine, inverse.
// ZM_TMF (Tungsten, first z pass): forward Psi, keep the full spectrum, then
// the inverse of fMF*Psi (the mean field).  ZM_TCOMB (Tungsten, second z
// pass): forward N, combine with the kept spectrum, inverse.
enum ZMode { ZM_PFC = 0, ZM_TMF = 1, ZM_TCOMB = 2 };

// Multipliers of one mode for a z pass of kind MODE, scaled by o.scale:
// (opL, opN) for ZM_PFC, (fMF, -) for ZM_TMF, (expL, nW) for ZM_TCOMB.
template <int MODE>
__device__ __forceinline__ void z_mode_ops(const PfcOp &o, const TungOp &p,
                                           double kxy, double kz, double &ma,
                                           double &mb) {
  if constexpr (MODE == ZM_PFC) {
    pfc_op_mode(o, kxy, kz, ma, mb);
  } else if constexpr (MODE == ZM_TMF) {
    ma = tung_filter(p, kxy, kz) * o.scale;
    mb = 0.0;
  } else {
    double f, e, n;
    tung_op_mode(o, p, kxy, kz, f, e, n);
    ma = e * o.scale;
    mb = n * o.scale;
  }
}

// ----------------------------------------------------------------------------
// Fused step, z pass (natural layout): for L adjacent kx columns at one ky,
//   ZM_PFC:   A = Fz^-1(opL*Fz(A) + opN*Fz(B))
//   ZM_TMF:   A = Fz(A); B = Fz^-1(fMF*A)
//   ZM_TCOMB: A = Fz^-1(expL*A + nW*Fz(B))
// with the multipliers evaluated per mode on the fly (no tables in HBM).
// ----------------------------------------------------------------------------
template <typename T> struct PfcZArgs {
  int zero_k0; // mean-split step: the k = 0 mode of the deviation stays 0
  int mode;    // ZMode
  cplx<T> *A;
  cplx<T> *B;
  const cplx<T> *tw;
  PfcOp op;
  TungOp top;
  const double *meps; // batched plans: per-field eps overriding op.eps, or null
  long mstride;       // distance between the spectra of consecutive fields (P*ny*nz)
  long s_pos; // P*ny
  int pitch;  // P
  int W;      // nx/2+1
  int ntx;
};

template <typename T, typename Plan, int L, bool SPLIT, int MODE>
__global__ void __launch_bounds__(L *Plan::TL) k_pfc_z(PfcZArgs<T> a) {
  constexpr int TL = Plan::TL, E = Plan::E, N = Plan::N;
  constexpr int NP = row_padded(N);
  __shared__ T lds[(SPLIT ? 1 : 2) * L * NP];
  const int c = threadIdx.x % L, t = threadIdx.x / L;
  const unsigned bid = xcd_block<T, L>(blockIdx.x, gridDim.x);
  const int tx = (int)(bid % (unsigned)a.ntx);
  // blocks run over (field, y, x tile); one field unless the plan is batched
  const long yb = bid / (unsigned)a.ntx;
  const int y = (int)(yb % a.op.ny);
  const long b = yb / a.op.ny;
  const int x = tx * L + c;
  const bool active = x < a.W;
  const long base = b * a.mstride + (long)y * a.pitch + x;
  const LdsLine<T> ll{lds, c, L, L * NP};
  PfcOp op = a.op;
  if (a.meps) op.eps = a.meps[b];
  const double kxy = pfc_kxy(op, x, y);

  cplx<T> acc[E];
  if constexpr (MODE == ZM_PFC) {
#pragma unroll
    for (int e = 0; e < E; ++e)
      acc[e] = active ? a.A[base + (long)(t + TL * e) * a.s_pos] : cplx<T>{0, 0};
    LineFft<T, Plan, -1, SPLIT, true>::run(acc, t, ll, a.tw);
    T opn[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      double l, n;
      pfc_op_mode(op, kxy, pfc_kc(t + TL * e, op.nz, op.fz), l, n);
      acc[e] = c_scale(acc[e], (T)l);
      opn[e] = (T)n;
    }
    {
      cplx<T> v[E];
#pragma unroll
      for (int e = 0; e < E; ++e)
        v[e] = active ? a.B[base + (long)(t + TL * e) * a.s_pos] : cplx<T>{0, 0};
      LineFft<T, Plan, -1, SPLIT, true>::run(v, t, ll, a.tw);
#pragma unroll
      for (int e = 0; e < E; ++e) acc[e] = c_add(acc[e], c_scale(v[e], opn[e]));
    }
    if (a.zero_k0 && x == 0 && y == 0 && t == 0) acc[0] = cplx<T>{0, 0};
  } else if constexpr (MODE == ZM_TMF) {
#pragma unroll
    for (int e = 0; e < E; ++e)
      acc[e] = active ? a.A[base + (long)(t + TL * e) * a.s_pos] : cplx<T>{0, 0};
    LineFft<T, Plan, -1, SPLIT, true>::run(acc, t, ll, a.tw);
    if (active) {
#pragma unroll
      for (int e = 0; e < E; ++e) a.A[base + (long)(t + TL * e) * a.s_pos] = acc[e];
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
      double f, unused;
      z_mode_ops<ZM_TMF>(op, a.top, kxy, pfc_kc(t + TL * e, op.nz, op.fz), f, unused);
      acc[e] = c_scale(acc[e], (T)f);
    }
  } else {
    // multipliers first, while nothing else is live
    T opl[E], opn[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      double l, n;
      z_mode_ops<ZM_TCOMB>(op, a.top, kxy, pfc_kc(t + TL * e, op.nz, op.fz), l, n);
      opl[e] = (T)l;
      opn[e] = (T)n;
    }
#pragma unroll
    for (int e = 0; e < E; ++e)
      acc[e] = active ? a.B[base + (long)(t + TL * e) * a.s_pos] : cplx<T>{0, 0};
    LineFft<T, Plan, -1, SPLIT, true>::run(acc, t, ll, a.tw);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const cplx<T> h = active ? a.A[base + (long)(t + TL * e) * a.s_pos] : cplx<T>{0, 0};
      acc[e] = c_add(c_scale(h, opl[e]), c_scale(acc[e], opn[e]));
    }
  }
  LineFft<T, Plan, +1, SPLIT, true>::run(acc, t, ll, a.tw);
  if (active) {
    cplx<T> *__restrict__ out = MODE == ZM_TMF ? a.B : a.A;
#pragma unroll
    for (int e = 0; e < E; ++e) out[base + (long)(t + TL * e) * a.s_pos] = acc[e];
  }
}

// ----------------------------------------------------------------------------
// Tungsten step, fused x pass (natural layout).  Per row: the mean-field half
// spectrum (inverse y and z already applied, 1/(nx ny nz) folded into fMF) is
// brought to real space with the nx/2-point inverse, combined with the row of
// psi into the nonlinearity N(psi, psi_MF), and N is transformed forward into
// the same spectrum row.  psi_MF never goes to HBM unless mf_out is set.
// ----------------------------------------------------------------------------
template <typename T> struct TungXArgs {
  TungNl<T> nl;
  const T *psi;      // real rows (nx = 2M values)
  cplx<T> *spec;     // in: mean-field half spectra; out: N half spectra
  T *mf_out;         // optional copy of psi_MF (real rows), or null
  const cplx<T> *tw; // M-point line twiddles
  const cplx<T> *wn; // exp(-2 pi i k/nx), k = 0..M
  long rows;
  int pitch;
};

template <typename T, typename Plan, int LPB>
__global__ void __launch_bounds__(LPB *Plan::TL) k_x_tung(TungXArgs<T> a) {
  constexpr int TL = Plan::TL, E = Plan::E, M = Plan::N;
  constexpr int RP = row_padded(M);
  __shared__ T lds[2 * LPB * RP];
  const int t = threadIdx.x % TL, c = threadIdx.x / TL;
  const long row = (long)blockIdx.x * LPB + c;
  const bool active = row < a.rows;
  const long r = active ? row : 0;
  cplx<T> *__restrict__ sp = a.spec + r * a.pitch;
  const LdsLine<T> ll{lds, c * RP, 1, LPB * RP};

  cplx<T> v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int k = t + TL * e;
    v[e] = c2r_merge(sp[k], sp[M - k], a.wn[k], k == 0);
  }
  LineFft<T, Plan, +1, false, true, row_wave<Plan>()>::run(v, t, ll, a.tw);
  // v[e] = (psi_MF[2k], psi_MF[2k+1]), k = t + TL*e
  const cplx<T> *__restrict__ pr =
      reinterpret_cast<const cplx<T> *>(a.psi + r * (2L * M));
  if (a.mf_out && active) {
    cplx<T> *__restrict__ mo = reinterpret_cast<cplx<T> *>(a.mf_out + r * (2L * M));
#pragma unroll
    for (int e = 0; e < E; ++e) mo[t + TL * e] = v[e];
  }
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const cplx<T> u = pr[t + TL * e];
    v[e] = cplx<T>{a.nl(u.x, v[e].x), a.nl(u.y, v[e].y)};
  }
  LineFft<T, Plan, -1, false, true, row_wave<Plan>()>::run(v, t, ll, a.tw);
  cplx<T> pz[E];
  row_fetch_partner<T, E, TL, row_wave<Plan>()>(v, pz, t, lds + c * RP, lds + LPB * RP + c * RP);
  if (!active) return;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int k = t + TL * e;
    sp[k] = r2c_split(v[e], pz[e], a.wn[k]);
  }
  if (t == 0) sp[M] = cplx<T>{v[0].x - v[0].y, (T)0};
}

// ----------------------------------------------------------------------------
// Initial conditions on the device, in the reference's coordinate convention
// x = origin + i*spacing (include/openpfc/kernel/data/domain.hpp:185-189) and
// x-fastest linear order (include/openpfc/kernel/field/operations.hpp:81-105).
// ----------------------------------------------------------------------------
template <typename T>
__global__ void k_ic_constant(T *psi, long n, double value) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    psi[i] = (T)value;
}

// splitmix64 finaliser (Steele, Lea, Flood 2014; public-domain algorithm).
__host__ __device__ inline unsigned long long splitmix64(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// psi[i] += lo + (hi-lo)*u_i, u_i = (splitmix64(splitmix64(seed) + i) >> 11)*2^-53:
// a counter-based uniform stream, identical on host and device.
template <typename T>
__global__ void k_ic_add_noise(T *psi, long n, unsigned long long seed,
                               double lo, double hi) {
#pragma clang fp contract(off)
  const unsigned long long s0 = splitmix64(seed);
  const double w = hi - lo;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const double u =
        (double)(splitmix64(s0 + (unsigned long long)i) >> 11) * 0x1.0p-53;
    psi[i] = (T)((double)psi[i] + (lo + (w * u)));
  }
}

// FixedBC (include/openpfc/kernel/simulation/boundary_conditions/fixed_bc.hpp):
// at x = x0 + i*dx within |x - xpos| < xwidth the field becomes
// rho_low*S + rho_high*(1 - S), S = 1/(1 + exp(-alpha*(x - xpos))); elsewhere
// it is left untouched.  Only the band's points are written.
struct FixedBcArgs {
  double x0, dx, xpos, xwidth, alpha, rho_low, rho_high;
  int nx;
  long rows; // ny*nz
};

template <typename T> __global__ void k_bc_fixed(T *psi, FixedBcArgs a) {
#pragma clang fp contract(off)
  const long n = (long)a.nx * a.rows;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < n;
       idx += (long)gridDim.x * blockDim.x) {
    const int i = (int)(idx % a.nx);
    const double x = (a.x0 + ((double)i * a.dx));
    if (fabs(x - a.xpos) < a.xwidth) {
      const double S = 1.0 / (1.0 + exp((-a.alpha) * (x - a.xpos)));
      psi[idx] = (T)((a.rho_low * S) + (a.rho_high * (1.0 - S)));
    }
  }
}

// SingleSeed (include/openpfc/kernel/simulation/initial_conditions/single_seed.hpp:51-83):
// inside x^2+y^2+z^2 < radius^2 the value is rho + sum_q 2*amp*cos(q.x) over
// the six <110>/sqrt(2) vectors; outside it writes 0.0 (keep_outside = 0, the
// reference behaviour) or leaves the field untouched (keep_outside = 1).
struct SeedArgs {
  double x0, y0, z0, dx, dy, dz, rho, amp, r2;
  int nx, ny, nz, keep_outside;
};

template <typename T> __global__ void k_ic_single_seed(T *psi, SeedArgs a) {
#pragma clang fp contract(off)
  const long n = (long)a.nx * a.ny * a.nz;
  const double s = 1.0 / sqrt(2.0);
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < n;
       idx += (long)gridDim.x * blockDim.x) {
    const int i = (int)(idx % a.nx);
    const long jk = idx / a.nx;
    const int j = (int)(jk % a.ny);
    const int k = (int)(jk / a.ny);
    const double x = (a.x0 + ((double)i * a.dx));
    const double y = (a.y0 + ((double)j * a.dy));
    const double z = (a.z0 + ((double)k * a.dz));
    const double rr =
        (((x * x) + (y * y)) + (z * z));
    if (rr >= a.r2) {
      if (!a.keep_outside) psi[idx] = (T)0;
      continue;
    }
    const double sx = (s * x), sy = (s * y), sz = (s * z);
    // q.x for q1..q6 = (s,s,0) (s,0,s) (0,s,s) (s,0,-s) (s,-s,0) (0,s,-s),
    // summed as (qx*x + qy*y) + qz*z with the zero terms kept
    const double zero = 0.0;
    const double arg[6] = {
        ((sx + sy) + (zero * z)),
        ((sx + (zero * y)) + sz),
        (((zero * x) + sy) + sz),
        ((sx + (zero * y)) + -sz),
        ((sx + -sy) + (zero * z)),
        (((zero * x) + sy) + -sz)};
    double u = a.rho;
    const double two_amp = (2.0 * a.amp);
#pragma unroll
    for (int q = 0; q < 6; ++q) u = (u + (two_amp * cos(arg[q])));
    psi[idx] = (T)u;
  }
}

// One rotated seed (initial_conditions/seed.hpp:124-141) painted over the
// box [i0, i0+bx) x [j0, j0+by) x [k0, k0+bz) that holds its sphere: where
// (X - location)^2 < radius^2 the value is rho + sum_i 2*amp*cos(q_i . X),
// X the absolute coordinate.  A seed list is painted last seed first, so a
// point inside several seeds ends with the first one's value, as the
// reference's first-match loop gives.
struct SeedBoxArgs {
  double x0, y0, z0, dx, dy, dz;
  double loc[3], q[18], rho, r2, amp;
  int nx, ny, i0, j0, k0, bx, by;
  long nbox; // bx*by*bz
};

template <typename T> __global__ void k_ic_seed_box(T *psi, SeedBoxArgs a) {
#pragma clang fp contract(off)
  for (long b = blockIdx.x * (long)blockDim.x + threadIdx.x; b < a.nbox;
       b += (long)gridDim.x * blockDim.x) {
    const int i = a.i0 + (int)(b % a.bx);
    const long jk = b / a.bx;
    const int j = a.j0 + (int)(jk % a.by);
    const int k = a.k0 + (int)(jk / a.by);
    const double X = (a.x0 + ((double)i * a.dx));
    const double Y = (a.y0 + ((double)j * a.dy));
    const double Z = (a.z0 + ((double)k * a.dz));
    const double x = X - a.loc[0], y = Y - a.loc[1], z = Z - a.loc[2];
    if (!((((x * x) + (y * y)) + (z * z)) < a.r2)) continue;
    double u = a.rho;
    const double two_amp = (2.0 * a.amp);
#pragma unroll
    for (int q = 0; q < 6; ++q)
      u = (u + (two_amp * cos((((a.q[3 * q] * X) + (a.q[3 * q + 1] * Y)) +
                               (a.q[3 * q + 2] * Z)))));
    psi[((long)k * a.ny + j) * a.nx + i] = (T)u;
  }
}

// MovingBC's column profile (boundary_conditions/moving_bc.hpp:120-127):
// the maximum of the field over (y, z) for every x.  Stage one: block
// (bx, by) reduces rows by, by + gridDim.y, ... for the 256 columns of bx
// into part[by*nx + x]; stage two (k_column_max_final, next to its launch)
// reduces the gridDim.y partials per column.  max is exact, so the result
// does not depend on the split.
template <typename T>
__global__ void __launch_bounds__(256) k_column_max_part(const T *__restrict__ psi, int nx,
                                                         long rows, double *__restrict__ part) {
  const int x = blockIdx.x * 256 + threadIdx.x;
  if (x >= nx) return;
  double m = -1.7976931348623157e308; // numeric_limits<double>::lowest()
  for (long r = blockIdx.y; r < rows; r += gridDim.y) {
    const double v = (double)psi[r * nx + x];
    m = v > m ? v : m;
  }
  part[(long)blockIdx.y * nx + x] = m;
}

// MovingBC::fill_bc (moving_bc.hpp:187-219): with l = nx*dx and
// p = fmod(xpos, l), columns whose x = x0 + i*dx lies within xwidth of p (or
// of p -+ l where the band wraps round the periodic box) become
// rho_low*S + rho_high*(1-S), S = 1/(1 + exp(-alpha*d)); the launch covers
// the columns [i0, i0+bx) that can be in a band, every other value is kept.
struct MovingBcArgs {
  double x0, dx, xpos, l, xwidth, alpha, rho_low, rho_high;
  int nx, i0, bx;
  long nbox; // bx*ny*nz
};

template <typename T> __global__ void k_bc_moving(T *psi, MovingBcArgs a) {
#pragma clang fp contract(off)
  for (long b = blockIdx.x * (long)blockDim.x + threadIdx.x; b < a.nbox;
       b += (long)gridDim.x * blockDim.x) {
    const int i = a.i0 + (int)(b % a.bx);
    const long row = b / a.bx;
    const double x = (a.x0 + ((double)i * a.dx));
    const double dist = x - a.xpos;
    double d;
    if (fabs(dist) < a.xwidth)
      d = dist;
    else if (a.xpos < a.xwidth && fabs(dist - a.l) < a.xwidth)
      d = dist - a.l;
    else if (a.xpos > a.l - a.xwidth && fabs(dist + a.l) < a.xwidth)
      d = dist + a.l;
    else
      continue;
    const double S = 1.0 / (1.0 + exp((-a.alpha) * d));
    psi[row * a.nx + i] = (T)((a.rho_low * S) + (a.rho_high * (1.0 - S)));
  }
}

} // namespace pfc_amd
