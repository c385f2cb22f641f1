#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/ from an implementation that
shares no code with the C oracle: numpy's pocketfft for the transforms and a
pure-Python mt19937_64 for the random initial fields.

    python tests/golden/make_golden.py

Fixtures (numpy .npz, no pickles):
  fft_small.npz   real inputs of shapes (nz, ny, nx) in {(4,6,8), (3,5,6), (2,1,8),
                  (1,1,2)} and their unnormalised R2C spectra (x halved, x fastest),
                  i.e. numpy.fft.rfftn over the reference's (z, y, x) memory order.
  pfc_16_20.npz   plain PFC ETD1 (docs/api/examples/10_complete_pfc_model.cpp:121-144,
                  166-213), 16^3, dx = pi/4, dt = 0.5, eps = -0.25,
                  psi0 = -0.2 + U(-0.01, 0.01) from mt19937_64(1234); psi after
                  0, 10 and 20 steps (full fields).
  pfc_64_200.npz  the same model at 64^3 (the size the CPU finishes in seconds), 200 steps;
                  per snapshot at steps 0/50/100/150/200: sum, sum of squares,
                  min, max and every 64th value in linear order.
The uniform draw is lo + (hi-lo) * ((r >> 11) * 2^-53) for each 64-bit output r.
"""
from __future__ import annotations

import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


class MT19937_64:
    """std::mt19937_64 ([rand.predef]); 10000th output of seed 5489 is
    9981545732273789042."""

    def __init__(self, seed: int):
        self.mt = [0] * 312
        self.mt[0] = seed & 0xFFFFFFFFFFFFFFFF
        for i in range(1, 312):
            self.mt[i] = (6364136223846793005 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 62)) + i) \
                & 0xFFFFFFFFFFFFFFFF
        self.i = 312

    def _twist(self):
        mt = self.mt
        for i in range(312):
            x = (mt[i] & 0xFFFFFFFF80000000) | (mt[(i + 1) % 312] & 0x7FFFFFFF)
            xa = x >> 1
            if x & 1:
                xa ^= 0xB5026F5AA96619E9
            mt[i] = mt[(i + 156) % 312] ^ xa
        self.i = 0

    def next(self) -> int:
        if self.i >= 312:
            self._twist()
        x = self.mt[self.i]
        self.i += 1
        x ^= (x >> 29) & 0x5555555555555555
        x ^= (x << 17) & 0x71D67FFFEDA60000
        x ^= (x << 37) & 0xFFF7EEE000000000
        x ^= x >> 43
        return x & 0xFFFFFFFFFFFFFFFF


def uniform_field(n, seed, lo, hi, base):
    g = MT19937_64(seed)
    total = n[0] * n[1] * n[2]
    r = np.fromiter((g.next() >> 11 for _ in range(total)), dtype=np.uint64, count=total)
    u = r.astype(np.float64) * 2.0 ** -53
    return (base + (lo + (hi - lo) * u)).reshape(n[2], n[1], n[0])


def pfc_ops(n, d, eps, dt):
    """Operators on the (nz, ny, nx//2+1) grid, vectorised in numpy."""
    nx, ny, nz = n
    f = [2.0 * math.pi / (d[i] * n[i]) for i in range(3)]
    i = np.arange(nx // 2 + 1)
    j = np.arange(ny)
    k = np.arange(nz)
    kx = np.where(i <= nx // 2, i, i - nx) * f[0]
    ky = np.where(j <= ny // 2, j, j - ny) * f[1]
    kz = np.where(k <= nz // 2, k, k - nz) * f[2]
    k2 = (kz[:, None, None] ** 2 + ky[None, :, None] ** 2) + kx[None, None, :] ** 2
    L = -k2 * (eps + (1.0 - k2) ** 2)
    opL = np.exp(L * dt)
    with np.errstate(divide="ignore", invalid="ignore"):
        opN = np.where(np.abs(L) > 1e-14, np.expm1(L * dt) / L * (-k2), dt * (-k2))
    return opL, opN


def pfc_run(psi, opL, opN, steps):
    for _ in range(steps):
        F = np.fft.rfftn(psi)
        Nl = np.fft.rfftn(psi * psi * psi)
        psi = np.fft.irfftn(opL * F + opN * Nl, s=psi.shape, axes=(0, 1, 2))
    return psi


def main():
    rng = np.random.default_rng(20240611)
    fft = {}
    for s in [(4, 6, 8), (3, 5, 6), (2, 1, 8), (1, 1, 2)]:
        a = rng.standard_normal(s)
        fft[f"in_{s[2]}x{s[1]}x{s[0]}"] = a
        fft[f"out_{s[2]}x{s[1]}x{s[0]}"] = np.fft.rfftn(a)
    np.savez_compressed(os.path.join(HERE, "fft_small.npz"), **fft)

    dx, dt, eps = math.pi / 4, 0.5, -0.25
    n = (16, 16, 16)
    psi = uniform_field(n, 1234, -0.01, 0.01, -0.2)
    opL, opN = pfc_ops(n, (dx, dx, dx), eps, dt)
    out = {"psi_0": psi}
    a = psi
    for s in (10, 20):
        a = pfc_run(a, opL, opN, 10)
        out[f"psi_{s}"] = a
    np.savez_compressed(os.path.join(HERE, "pfc_16_20.npz"), **out)

    n = (64, 64, 64)
    psi = uniform_field(n, 1234, -0.01, 0.01, -0.2)
    opL, opN = pfc_ops(n, (dx, dx, dx), eps, dt)
    out = {}
    a = psi
    for s in (0, 50, 100, 150, 200):
        if s:
            a = pfc_run(a, opL, opN, 50)
        flat = a.reshape(-1)
        out[f"stats_{s}"] = np.array([flat.sum(), (flat * flat).sum(), flat.min(), flat.max()])
        out[f"sub_{s}"] = flat[::64].copy()
    np.savez_compressed(os.path.join(HERE, "pfc_64_200.npz"), **out)


if __name__ == "__main__":
    main()
