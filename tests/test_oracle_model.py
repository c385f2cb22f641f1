"""Pins the oracle's k-space helpers, ETD coefficients, PFC operators and step.

Known answers from the reference's tests:
  tests/unit/kernel/fft/test_kspace.cpp:73-104         (Nyquist folding)
  tests/unit/kernel/integrator/test_spectral_exp_coefficients.cpp:29-92
Mathematics that fixes the rest:
  - k = 0: opL = 1, opN = 0, so the mean is conserved exactly
    (tests/integration/scenarios/complete_simulation/test_mass_conservation.cpp:47);
  - a constant field is a fixed point;
  - a small single-mode perturbation of psi_bar grows at the rate of the
    linearised PDE d(psi)/dt = lap[eps*psi + psi^3 + (1+lap)^2 psi], i.e.
    sigma(k) = -k^2 (eps + 3 psi_bar^2 + (1 - k^2)^2); with dt -> 0 the ETD1
    amplitude after time t approaches exp(sigma t), which pins the sign of the
    Laplacian, the (1-k^2)^2 term and the 3 psi_bar^2 coupling through opN;
  - the committed numpy trajectories (tests/golden/pfc_*.npz) computed with an
    independent FFT and RNG.
"""
import math
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
DX, DT, EPS = math.pi / 4, 0.5, -0.25


def test_k_component_nyquist_folding():
    f = 0.1
    assert O.k_component(0, 128, f) == 0.0
    assert O.k_component(1, 128, f) == 1.0 * f
    assert O.k_component(32, 128, f) == 32.0 * f
    assert O.k_component(64, 128, f) == 64.0 * f
    assert O.k_component(65, 128, f) == -63.0 * f
    assert O.k_component(100, 128, f) == -28.0 * f
    assert O.k_component(127, 128, f) == -1.0 * f


@pytest.mark.parametrize("L,dt", [(-4.0, 0.25), (1.5, 0.25)])
def test_spectral_exp_coeffs_nonzero(L, dt):
    e, p = O.spectral_exp_coeffs(L, dt)
    assert abs(e - math.exp(L * dt)) <= 1e-14 * abs(math.exp(L * dt))
    assert abs(p - math.expm1(L * dt) / L) <= 1e-14 * abs(math.expm1(L * dt) / L)


def test_spectral_exp_coeffs_edge_cases():
    assert O.spectral_exp_coeffs(-2.0, 0.0) == (1.0, 0.0)
    e, p = O.spectral_exp_coeffs(0.0, 0.1)
    assert e == 1.0 and abs(p - 0.1) <= 1e-15
    e, p = O.spectral_exp_coeffs(1e-13, 0.1)          # below threshold: Taylor
    assert abs(p - (0.1 + 0.5 * 1e-13 * 0.01)) <= 1e-18
    e, p = O.spectral_exp_coeffs(1e-12, 0.1)          # at threshold: expm1 branch
    assert abs(p - math.expm1(1e-13) / 1e-12) <= 1e-12 * p


def test_pfc_operators_dc_mode_and_shape():
    n = (16, 8, 4)
    opL, opN = O.pfc_operators(n, (DX, 1.0, 0.5), EPS, DT)
    assert opL.shape == (4, 8, 9)
    assert opL[0, 0, 0] == 1.0 and opN[0, 0, 0] == 0.0
    # opL is even in each wave number: index j and ny - j give the same k^2
    assert np.array_equal(opL[:, 1, :], opL[:, 7, :])
    assert np.array_equal(opN[:, :, :], opN[[0, 3, 2, 1], :, :])
    assert np.all(opL >= 0) and np.all(np.isfinite(opN))


def test_constant_field_fixed_point_and_mean_conservation():
    n = (16, 16, 16)
    opL, opN = O.pfc_operators(n, (DX, DX, DX), EPS, DT)
    psi = np.full((16, 16, 16), -0.3)
    out = O.pfc_step(psi, opL, opN, 5)
    assert np.abs(out - psi).max() <= 1e-15
    psi = O.uniform_noise(n, 7, -0.01, 0.01, base=-0.2)
    out = O.pfc_step(psi, opL, opN, 50)
    assert abs(out.mean() - psi.mean()) <= 1e-12 * abs(psi.mean())


@pytest.mark.parametrize("mode", [(1, 0, 0), (4, 0, 0), (2, 2, 0), (4, 4, 4)])
def test_single_mode_growth_matches_linearised_pde(mode):
    """dx = pi/4 on 32 points: index 4 along an axis is |k| = 1.  ETD1 is
    first order in the explicit (nonlinear) part, so dt is small."""
    N = 32
    n = (N, N, N)
    dt, t_end, pbar, A = 2.5e-4, 0.5, -0.2, 1e-7
    steps = int(round(t_end / dt))
    opL, opN = O.pfc_operators(n, (DX, DX, DX), EPS, dt)
    z, y, x = np.meshgrid(*(np.arange(N),) * 3, indexing="ij")
    mx, my, mz = mode
    phase = 2 * np.pi * (mx * x + my * y + mz * z) / N
    psi = pbar + A * np.cos(phase)
    out = O.pfc_step(psi, opL, opN, steps)
    amp = 2 * np.mean((out - out.mean()) * np.cos(phase))
    f = 2 * np.pi / (DX * N)
    k2 = (f * mx) ** 2 + (f * my) ** 2 + (f * mz) ** 2
    sigma = -k2 * (EPS + 3 * pbar ** 2 + (1 - k2) ** 2)
    expected = A * math.exp(sigma * t_end)
    assert abs(amp - expected) <= 2e-3 * abs(expected), (amp, expected, sigma)


def test_trajectory_16_matches_numpy_fixture():
    d = np.load(os.path.join(GOLD, "pfc_16_20.npz"))
    n = (16, 16, 16)
    psi0 = O.uniform_noise(n, 1234, -0.01, 0.01, base=-0.2)
    assert np.array_equal(psi0, d["psi_0"])  # mt19937_64 stream and draw rule
    opL, opN = O.pfc_operators(n, (DX, DX, DX), EPS, DT)
    a = psi0
    for s in (10, 20):
        a = O.pfc_step(a, opL, opN, 10)
        ref = d[f"psi_{s}"]
        assert np.abs(a - ref).max() <= 1e-13 * np.abs(ref).max(), s


def test_trajectory_64_200_matches_numpy_fixture():
    """The 64^3 demo workload: 200 steps (t = 0 -> 100)."""
    d = np.load(os.path.join(GOLD, "pfc_64_200.npz"))
    n = (64, 64, 64)
    a = O.uniform_noise(n, 1234, -0.01, 0.01, base=-0.2)
    opL, opN = O.pfc_operators(n, (DX, DX, DX), EPS, DT)
    for s in (0, 50, 100, 150, 200):
        if s:
            a = O.pfc_step(a, opL, opN, 50)
        flat = a.reshape(-1)
        sub = d[f"sub_{s}"]
        scale = np.abs(sub).max()
        assert np.abs(flat[::64] - sub).max() <= 1e-11 * scale, s
        st = d[f"stats_{s}"]
        assert abs(flat.sum() - st[0]) <= 1e-9 * abs(st[0])
        assert abs(flat.min() - st[2]) <= 1e-11 and abs(flat.max() - st[3]) <= 1e-11


@pytest.mark.parametrize("n", [(12, 6, 8), (6, 10, 7), (8, 4, 2)])
def test_pfc_operators_same_bits_on_mirrored_kz_planes(n):
    """kz enters the operators only as kz*kz, and k_component(k) =
    -k_component(nz - k) exactly for 0 < k < nz, k != nz/2, so the planes k and
    nz - k of opL and opN hold the same bits (even and odd nz).  The z pass
    that tabulates nz/2 + 1 multiplier pairs per line rests on this."""
    spacing = (DX, 0.9, 1.3)
    opL, opN = O.pfc_operators(n, spacing, EPS, DT)
    nz = n[2]
    f = 2 * math.pi / (spacing[2] * nz)
    for k in range(1, nz):
        if 2 * k != nz:
            assert O.k_component(k, nz, f) == -O.k_component(nz - k, nz, f), k
    mirror = (-np.arange(nz)) % nz
    assert np.array_equal(opL, opL[mirror]) and np.array_equal(opN, opN[mirror])
    # the planes are not all alike: the check can fail
    assert not np.array_equal(opL[1], opL[0])
