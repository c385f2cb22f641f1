"""The oracle's free energy of the plain PFC model,
F = dV [sum_k |Psi_k|^2 (eps + (1-k^2)^2) / (2N) + sum psi^4 / 4]
(no reference code computes it; it is the functional whose conserved
gradient flow is the step of docs/api/examples/10_complete_pfc_model.cpp:121-144).

Pinned by
  - the closed form for psi = m + A cos(q.x) with q on the grid:
    F/V = m^2 (eps+1)/2 + A^2 (eps + (1-q^2)^2)/4 + (m^4 + 3 m^2 A^2 + 3 A^4/8)/4
    (a dropped Parseval weight, a wrong sign of k^2 or a lost 1/2 fails it);
  - numpy's full complex fftn on random fields (no half-spectrum weights);
  - its decrease along an oracle trajectory (the step is a gradient flow).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

DX = math.pi / 4


@pytest.mark.parametrize("n,mode,m,A,eps", [((16, 16, 16), (2, 0, 0), -0.2, 0.1, -0.25),
                                            ((32, 16, 8), (8, 2, 1), 0.1, 0.3, -0.1),
                                            ((16, 8, 32), (8, 0, 16), -0.3, 0.05, 0.2)])
def test_single_mode_closed_form(n, mode, m, A, eps):
    nx, ny, nz = n
    L = [DX * v for v in n]
    q = [2 * math.pi * mode[i] / L[i] for i in range(3)]
    z, y, x = np.meshgrid(np.arange(nz) * DX, np.arange(ny) * DX, np.arange(nx) * DX,
                          indexing="ij")
    psi = m + A * np.cos(q[0] * x + q[1] * y + q[2] * z)
    q2 = sum(v * v for v in q)
    # the Nyquist plane of a real field carries cos only, so the mode at
    # index n/2 counts once with amplitude A (|Psi| = N A) instead of twice
    # with A/2; both give the same A^2/2 mean square except there
    nyq = any(2 * mode[i] == n[i] for i in range(3))
    ms = A * A if nyq else A * A / 2
    quart = m ** 4 + 6 * m * m * ms + (A ** 4 if nyq else 3 * A ** 4 / 8)
    V = nx * ny * nz * DX ** 3
    ref = V * (m * m * (eps + 1) / 2 + ms * (eps + (1 - q2) ** 2) / 2 + quart / 4)
    got = O.pfc_free_energy(psi, (DX,) * 3, eps)
    assert abs(got - ref) <= 1e-12 * abs(ref), (got, ref)


def test_matches_numpy_full_spectrum():
    n = (24, 12, 16)
    nx, ny, nz = n
    rng = np.random.default_rng(3)
    psi = -0.2 + 0.1 * rng.standard_normal((nz, ny, nx))
    eps = -0.25
    X = np.fft.fftn(psi)
    k = [2 * np.pi * np.fft.fftfreq(v, d=DX) for v in n]
    KZ, KY, KX = np.meshgrid(k[2], k[1], k[0], indexing="ij")
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    N = psi.size
    ref = DX ** 3 * ((np.abs(X) ** 2 * (eps + (1 - k2) ** 2)).sum() / (2 * N) + (psi ** 4).sum() / 4)
    got = O.pfc_free_energy(psi, (DX,) * 3, eps)
    assert abs(got - ref) <= 1e-12 * abs(ref), (got, ref)


def test_decreases_along_the_trajectory():
    n = (32, 32, 32)
    eps, dt = -0.25, 0.5
    psi = O.uniform_noise(n, 1234, -0.01, 0.01, base=-0.2)
    opL, opN = O.pfc_operators(n, (DX,) * 3, eps, dt)
    F = [O.pfc_free_energy(psi, (DX,) * 3, eps)]
    for _ in range(20):
        psi = O.pfc_step(psi, opL, opN, 5)
        F.append(O.pfc_free_energy(psi, (DX,) * 3, eps))
    assert all(b <= a for a, b in zip(F, F[1:])), F
    assert F[-1] < F[0]
