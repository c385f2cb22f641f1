// adjoint_kernels.hpp -- pointwise kernels of the adjoint of the plain PFC
// ETD1 step.
//
// Forward step: psi' = A psi + B psi^3 with A = F^-1 diag(opL) F and
// B = F^-1 diag(opN) F.  opL and opN are real and even in k, so A and B are
// symmetric on real fields and the adjoint of the step at psi is
//   lambda <- A lambda + 3 psi^2 (B lambda).
// The derivative of the step with respect to the scalar r (= eps) is
//   dpsi'/dr = F^-1(dopL F(psi) + dopN F(psi^3)),
//   dopL = dt (-k^2) opL,  dopN = (-k^2)^2 (dt opL L - (opL - 1)) / L^2
// (dopN = 0 on the |L| <= 1e-14 branch, where opN = dt (-k^2) does not
// depend on r), and its pairing with an adjoint field lambda is, by
// Parseval over the half spectrum (weight 2 except on the kx = 0 and
// kx = nx/2 planes),
//   <lambda, dpsi'/dr> = (1/N) sum_k w_kx Re(conj(Lambda_k) X_k).
// Mean-split fields (the FP32-mixed mode) hold d = psi - m with m exact:
// the same step map written in (m, d), so its adjoint is the one above at
// psi = m + d.  The cube is then formed as psi^3 - m^3 (as the forward step
// does) and F(d) stands for F(psi): both differ from the plain terms only at
// k = 0, where dopL = dopN = 0.
#pragma once

#include "fft_kernels.hpp"

namespace pfc_amd {

// d opL / d eps and d opN / d eps for one mode (unscaled).
__device__ __forceinline__ void pfc_op_deriv(const PfcOp &o, double kxy, double kz,
                                             double &opL, double &opN, double &dL,
                                             double &dN) {
#pragma clang fp contract(off)
  const double k2 = (kxy + (kz * kz));
  const double k2_lap = -k2;
  const double one_m = 1.0 - k2;
  const double L = (k2_lap * (o.eps + (one_m * one_m)));
  const double oL = exp((L * o.dt));
  opL = oL;
  if (fabs(L) > 1e-14) {
    opN = ((oL - 1.0) / L * k2_lap);
    // dt*opL*L - (opL - 1) = x e^x - expm1(x), x = L dt
    const double x = L * o.dt;
    const double num = (x * oL) - expm1(x);
    dN = ((k2_lap * k2_lap) * num) / (L * L);
  } else {
    opN = (o.dt * k2_lap);
    dN = 0.0;
  }
  dL = (o.dt * k2_lap) * oL;
}

// n3 = psi^3 (the reference's psi*psi*psi), or psi^3 - m^3 from the
// deviation (Nonlin of the forward step)
template <typename T> __global__ void k_cube(const T *psi, T *out, long n, Nonlin<T> nl) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    out[i] = nl(psi[i]);
}

// Spectral part of the adjoint step over the outbox (pitch W, x fastest):
//   a = opL * Lam, b = opN * Lam  (written over P and Nh)
//   partial[block] = sum w_kx Re(conj(Lam) (dopL P + dopN Nh))   (if dr)
// One block of 256 threads per chunk of modes; the block sum is a fixed-order
// tree, so the result does not depend on scheduling.
template <typename T>
__global__ void __launch_bounds__(256)
    k_adj_spectral(PfcOp o, int W, const cplx<T> *lam, cplx<T> *P, cplx<T> *Nh,
                   double *partial, int want_dr) {
  __shared__ double red[256];
  const long total = (long)W * o.ny * o.nz;
  double acc = 0.0;
  for (long idx = blockIdx.x * 256L + threadIdx.x; idx < total; idx += (long)gridDim.x * 256L) {
    const int i = (int)(idx % W);
    const long jk = idx / W;
    const int j = (int)(jk % o.ny);
    const int k = (int)(jk / o.ny);
    double oL, oN, dL, dN;
    pfc_op_deriv(o, pfc_kxy(o, i, j), pfc_kc(k, o.nz, o.fz), oL, oN, dL, dN);
    const cplx<T> l = lam[idx];
    if (want_dr) {
      const cplx<T> p = P[idx], q = Nh[idx];
      const double xr = dL * (double)p.x + dN * (double)q.x;
      const double xi = dL * (double)p.y + dN * (double)q.y;
      const double w = (i == 0 || 2 * i == o.nx) ? 1.0 : 2.0;
      acc += w * ((double)l.x * xr + (double)l.y * xi);
    }
    P[idx] = cplx<T>{(T)(oL * (double)l.x), (T)(oL * (double)l.y)};
    Nh[idx] = cplx<T>{(T)(oN * (double)l.x), (T)(oN * (double)l.y)};
  }
  if (!want_dr) return;
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Fixed-order sum of n partials by one block of 256 threads into out[0].
__global__ void __launch_bounds__(256) k_sum_partials(const double *partial, int n,
                                                      double *out) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) acc += partial[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

// lambda = a + 3 psi^2 b with psi = m + field (m = 0 for a plain field)
template <typename T>
__global__ void k_adj_combine(const T *psi, const T *a, const T *b, T *lambda, long n, T m) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const T u = m + psi[i];
    lambda[i] = a[i] + (((T)3 * (u * u)) * b[i]);
  }
}

} // namespace pfc_amd
