"""Pins the oracle's 3-D R2C/C2R (the checker every GPU test relies on).

Checks, each chosen so a plausible slip (wrong axis order, sign, missing
conjugate, lost Nyquist term, wrong scale) fails:
  - brute-force O(N^2) DFT on tiny grids, even and odd nx;
  - committed numpy/pocketfft fixtures (tests/golden/fft_small.npz);
  - the reference's FFT known-answer tests
    (tests/unit/kernel/fft/test_fft.cpp:37-47 8-pt DC 21.991,
     test_fft.cpp:59-70 2-pt backward 1.5,
     tests/unit/runtime/gpu/test_fft_hip.cpp:105-117 DC 1000 -> 1000/64^3);
  - round trips and the half-complex rule (imaginary parts of the kx = 0 and
    kx = nx/2 planes are ignored by C2R).
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("shape", [(4, 2, 2), (6, 3, 2), (5, 3, 2), (8, 1, 1), (2, 4, 3),
                                   (3, 1, 5)])
def test_r2c_matches_brute_force_dft(shape):
    nx, ny, nz = shape
    rng = np.random.default_rng(sum(shape))
    a = rng.standard_normal((nz, ny, nx))
    got = O.r2c(a)
    ref = O.dft_r2c_brute(a)
    assert got.shape == (nz, ny, nx // 2 + 1)
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()


def test_brute_force_dft_definition():
    """The brute-force DFT itself: a single plane wave lands in one bin with
    amplitude N^3 (forward exp(-i k.x), unnormalised)."""
    nx, ny, nz = 6, 4, 5
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    kx, ky, kz = 2, 1, 3
    a = np.cos(2 * np.pi * (kx * x / nx + ky * y / ny + kz * z / nz))
    X = O.dft_r2c_brute(a)
    assert abs(X[kz, ky, kx] - nx * ny * nz / 2) < 1e-9
    assert abs(X[nz - kz, ny - ky, kx] - 0) < 1e-9  # conjugate partner is at -k (not stored)
    mask = np.ones_like(X, dtype=bool)
    mask[kz, ky, kx] = False
    assert np.abs(X[mask]).max() < 1e-9


def test_matches_numpy_fixtures():
    d = np.load(os.path.join(GOLD, "fft_small.npz"))
    names = sorted(k[3:] for k in d.files if k.startswith("in_"))
    assert len(names) == 4
    for name in names:
        a, ref = d["in_" + name], d["out_" + name]
        got = O.r2c(a)
        assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max()), name
        back = O.c2r(ref, a.shape[2])
        assert np.abs(back - a).max() <= 1e-13, name


def test_kat_forward_8pt_dc():
    x = np.array([0.000, 0.785, 1.571, 2.356, 3.142, 3.927, 4.712, 5.498]).reshape(1, 1, 8)
    X = O.r2c(x)
    assert X.size == 5
    assert abs(X.reshape(-1)[0].real - 21.991) <= 0.01


def test_kat_backward_2pt():
    X = np.array([1.0, 2.0], dtype=np.complex128).reshape(1, 1, 2)
    y = O.c2r(X, 2).reshape(-1)
    assert abs(y[0] - 1.5) <= 0.01
    assert abs(y[1] - (-0.5)) <= 1e-15


def test_kat_dc_backward_64():
    X = np.zeros((64, 64, 33), dtype=np.complex128)
    X[0, 0, 0] = 1000.0
    y = O.c2r(X, 64)
    assert np.allclose(y, 1000.0 / 64 ** 3, atol=1e-15, rtol=0)


@pytest.mark.parametrize("shape", [(16, 16, 16), (32, 8, 4), (12, 10, 6), (7, 5, 3)])
def test_roundtrip(shape):
    nx, ny, nz = shape
    a = np.random.default_rng(3).standard_normal((nz, ny, nx))
    assert np.abs(O.c2r(O.r2c(a), nx) - a).max() <= 1e-14 * 8


def test_c2r_ignores_imag_of_dc_and_nyquist_planes():
    nx, ny, nz = 8, 4, 4
    a = np.random.default_rng(4).standard_normal((nz, ny, nx))
    X = O.r2c(a)
    ref = O.c2r(X, nx)
    Y = X.copy()
    # imaginary part of X[kx=0] / X[kx=nx/2] at ky = kz = 0 is not representable
    # by a real field; a half-complex inverse must drop it
    Y[0, 0, 0] += 5j
    Y[0, 0, nx // 2] += 3j
    assert np.abs(O.c2r(Y, nx) - ref).max() <= 1e-14


def test_fft1d_against_direct_sum():
    for n in [1, 2, 3, 4, 5, 6, 7, 8, 12, 16, 30, 64, 100]:
        x = np.random.default_rng(n).standard_normal(n) + 1j * np.random.default_rng(n + 1).standard_normal(n)
        k = np.arange(n)
        W = np.exp(-2j * np.pi * np.outer(k, k) / n)
        assert np.abs(O.fft1d(x) - W @ x).max() <= 1e-12 * max(1, n)
        assert np.abs(O.fft1d(x, +1) - W.conj() @ x).max() <= 1e-12 * max(1, n)


@pytest.mark.parametrize("shape", [(48, 24, 12), (24, 12, 96), (192, 6, 3), (12, 48, 24)])
def test_mixed_radix_lengths_match_numpy(shape):
    """The lengths of the uneven-size ensembles (factors 3 and 2^k) against
    numpy's rfftn (pocketfft) and back."""
    nx, ny, nz = shape
    a = np.random.default_rng(nx * ny + nz).standard_normal((nz, ny, nx))
    ref = np.fft.rfftn(a)
    got = O.r2c(a)
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
    assert np.abs(O.c2r(got, nx) - a).max() <= 1e-13 * np.abs(a).max()
