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
  const int tx = blockIdx.x % a.ntx;
  const long outer = blockIdx.x / a.ntx;
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
// Row helpers: fetch the value held at line position `pp(e)` for every owned
// element, through LDS (row layout with padding).
// ----------------------------------------------------------------------------
template <typename T, int E, int TL, typename PartnerPos>
__device__ __forceinline__ void row_fetch_partner(const cplx<T> (&v)[E],
                                                  cplx<T> (&p)[E], int t,
                                                  T *lds_re, T *lds_im,
                                                  PartnerPos pp) {
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int i = pad_pos<true>(t + TL * e);
    lds_re[i] = v[e].x;
    lds_im[i] = v[e].y;
  }
  __syncthreads();
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int i = pad_pos<true>(pp(t + TL * e));
    p[e] = cplx<T>{lds_re[i], lds_im[i]};
  }
  __syncthreads();
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
  LineFft<T, Plan, -1, false, true>::run(v, t, ll, a.tw);

  cplx<T> pz[E];
  row_fetch_partner<T, E, TL>(v, pz, t, re, im,
                              [](int k) { return k == 0 ? 0 : M - k; });
  if (!active) return;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int k = t + TL * e;
    const cplx<T> zk = v[e], zc = c_conj(pz[e]);
    const cplx<T> ev = c_scale(c_add(zk, zc), (T)0.5);
    const cplx<T> d = c_sub(zk, zc);
    const cplx<T> od = {(T)0.5 * d.y, (T)-0.5 * d.x}; // -i/2 * d
    out[k] = c_add(ev, c_mul(a.wn[k], od));
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
    cplx<T> xa = in[k];
    cplx<T> xb = c_conj(in[M - k]);
    if (k == 0) {
      xa.y = (T)0;
      xb.y = (T)0;
    }
    const cplx<T> ev = c_add(xa, xb);
    const cplx<T> od = c_mulc(c_sub(xa, xb), a.wn[k]); // * exp(+2 pi i k/nx)
    v[e] = cplx<T>{ev.x - od.y, ev.y + od.x};          // ev + i*od
  }
  const LdsLine<T> ll{lds, c * RP, 1, LPB * RP};
  LineFft<T, Plan, +1, false, true>::run(v, t, ll, a.tw);
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
  __device__ __forceinline__ T operator()(T u) const {
    return split ? u * (c2 + (c1 + u) * u) : u * u * u;
  }
};

template <typename T> struct PfcXArgs {
  Nonlin<T> nl;
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

  cplx<T> v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const T u = in[t + TL * e];
    v[e] = cplx<T>{u, a.nl(u)};
  }
  T *re = lds + c * RP;
  T *im = lds + LPB * RP + c * RP;
  const LdsLine<T> ll{lds, c * RP, 1, LPB * RP};
  LineFft<T, Plan, -1, false, true>::run(v, t, ll, a.tw);

  cplx<T> pz[E];
  row_fetch_partner<T, E, TL>(v, pz, t, re, im,
                              [](int k) { return k == 0 ? 0 : N - k; });
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
// Fused PFC step, z pass: for L adjacent kx columns at one ky,
//   A = Fz(Psi) ; B = Fz(Nl) ; A = opL*A + opN*B ; Psi = Fz^-1(A)
// with opL/opN evaluated per mode on the fly (no operator tables in HBM).
// ----------------------------------------------------------------------------
template <typename T> struct PfcZArgs {
  int zero_k0; // mean-split step: the k = 0 mode of the deviation stays 0
  cplx<T> *A;
  const cplx<T> *B;
  const cplx<T> *tw;
  PfcOp op;
  long s_pos; // P*ny
  int pitch;  // P
  int W;      // nx/2+1
  int ntx;
};

template <typename T, typename Plan, int L, bool SPLIT>
__global__ void __launch_bounds__(L *Plan::TL) k_pfc_z(PfcZArgs<T> a) {
  constexpr int TL = Plan::TL, E = Plan::E, N = Plan::N;
  constexpr int NP = row_padded(N);
  __shared__ T lds[(SPLIT ? 1 : 2) * L * NP];
  const int c = threadIdx.x % L, t = threadIdx.x / L;
  const int tx = blockIdx.x % a.ntx;
  const int y = blockIdx.x / a.ntx;
  const int x = tx * L + c;
  const bool active = x < a.W;
  const long base = (long)y * a.pitch + x;
  const LdsLine<T> ll{lds, c, L, L * NP};

  cplx<T> acc[E];
#pragma unroll
  for (int e = 0; e < E; ++e)
    acc[e] = active ? a.A[base + (long)(t + TL * e) * a.s_pos] : cplx<T>{0, 0};
  LineFft<T, Plan, -1, SPLIT, true>::run(acc, t, ll, a.tw);

  const double kxy = pfc_kxy(a.op, x, y);
  T opn[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    double l, n;
    pfc_op_mode(a.op, kxy, pfc_kc(t + TL * e, a.op.nz, a.op.fz), l, n);
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
  LineFft<T, Plan, +1, SPLIT, true>::run(acc, t, ll, a.tw);
  if (active) {
#pragma unroll
    for (int e = 0; e < E; ++e) a.A[base + (long)(t + TL * e) * a.s_pos] = acc[e];
  }
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

} // namespace pfc_amd
