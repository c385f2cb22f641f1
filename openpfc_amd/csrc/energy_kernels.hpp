// energy_kernels.hpp -- FP64 reductions of the plain PFC free energy
//
//   F = dV sum_x [ 1/2 psi (eps + (1 + lap)^2) psi + 1/4 psi^4 ]
//     = dV [ 1/(2N) sum_k |Psi_k|^2 (eps + (1 - k^2)^2) + 1/4 sum_x psi^4 ],
//
// the functional whose conserved gradient flow d psi/dt = lap dF/dpsi is the
// step's PDE (L = -k^2 (eps + (1-k^2)^2), N = psi^3).  The k sum runs over
// the half spectrum with weight 2 off the kx = 0 and kx = nx/2 planes.  For a
// mean-split field (d = psi - m) the k = 0 amplitude is N m + Re D_0, formed
// in double from the exact mean, and psi^4 = (m + d)^4 is summed in double,
// so the FP32 deviation never carries the O(1) part.  Every block sums its
// terms in a fixed-order tree; the partials are summed in a fixed order too.
#pragma once

#include "fft_kernels.hpp"

namespace pfc_amd {

// Block-tree sum of `acc` over 256 threads; thread 0 writes partial[block].
__device__ __forceinline__ void block_sum_256(double acc, double *partial) {
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// sum_k w |X_k|^2 (eps + (1 - k^2)^2) over the half spectrum (pitch W);
// k0 = N*m is added to Re X_0.
template <typename T>
__global__ void __launch_bounds__(256)
    k_fe_spectral(PfcOp o, int W, const cplx<T> *X, double k0, double *partial) {
  const long total = (long)W * o.ny * o.nz;
  double acc = 0.0;
  for (long idx = blockIdx.x * 256L + threadIdx.x; idx < total; idx += (long)gridDim.x * 256L) {
    const int i = (int)(idx % W);
    const long jk = idx / W;
    const int j = (int)(jk % o.ny);
    const int k = (int)(jk / o.ny);
    const double kz = pfc_kc(k, o.nz, o.fz);
    const double k2 = pfc_kxy(o, i, j) + kz * kz;
    const double one_m = 1.0 - k2;
    double re = (double)X[idx].x, im = (double)X[idx].y;
    if (idx == 0) re += k0;
    const double w = (i == 0 || 2 * i == o.nx) ? 1.0 : 2.0;
    acc += w * (re * re + im * im) * (o.eps + one_m * one_m);
  }
  block_sum_256(acc, partial);
}

// sum_x (m + psi)^4 in double
template <typename T>
__global__ void __launch_bounds__(256)
    k_fe_quartic(const T *psi, long n, double m, double *partial) {
  double acc = 0.0;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256L) {
    const double u = m + (double)psi[i];
    const double u2 = u * u;
    acc += u2 * u2;
  }
  block_sum_256(acc, partial);
}

} // namespace pfc_amd
