"""ctypes wrapper over the CPU oracle (``oracle/pfc_oracle.c``).

TEST INFRASTRUCTURE ONLY: the tests, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use this module as the checker.  The
product (``openpfc_amd``) never imports it.

Every function mirrors a C function whose header comment cites the
reference file:line it restates.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_pfc.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc only)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE],
                       check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_mt64_nth.restype = ctypes.c_uint64
        L.orc_mt64_nth.argtypes = [ctypes.c_uint64, ctypes.c_long]
        L.orc_add_uniform_noise.argtypes = [_dp, ctypes.c_size_t, ctypes.c_uint64,
                                            ctypes.c_double, ctypes.c_double]
        L.orc_fft1d.argtypes = [ctypes.c_int, _dp, ctypes.c_int]
        L.orc_r2c3d.argtypes = [ctypes.c_int] * 3 + [_dp, _dp]
        L.orc_c2r3d.argtypes = [ctypes.c_int] * 3 + [_dp, _dp, _dp]
        L.orc_dft_r2c3d_brute.argtypes = [ctypes.c_int] * 3 + [_dp, _dp]
        L.orc_k_component.restype = ctypes.c_double
        L.orc_k_component.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.orc_spectral_exp_coeffs.argtypes = [ctypes.c_double] * 3 + [_dp, _dp]
        L.orc_pfc_operators.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 5 + [_dp, _dp]
        L.orc_pfc_step.argtypes = [ctypes.c_int] * 3 + [_dp, _dp, _dp, ctypes.c_int]
        L.orc_tungsten_physics_for_mode.argtypes = [ctypes.c_double, _dp, _dp, _dp]
        L.orc_tungsten_operators.argtypes = ([ctypes.c_int] * 3 + [ctypes.c_double] * 3
                                             + [_dp, ctypes.c_double, _dp, _dp, _dp])
        L.orc_tungsten_step.argtypes = [ctypes.c_int] * 3 + [_dp] * 6 + [ctypes.c_int]
        L.orc_ic_constant.argtypes = [_dp, ctypes.c_size_t, ctypes.c_double]
        L.orc_bc_fixed.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 6 + [_dp]
        L.orc_pfc_operator_derivs.argtypes = ([ctypes.c_int] * 3 + [ctypes.c_double] * 5
                                              + [_dp, _dp])
        L.orc_pfc_adjoint_step.argtypes = [ctypes.c_int] * 3 + [_dp] * 6 + [_dp]
        L.orc_ic_single_seed.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 9 + [_dp]
        L.orc_add_counter_noise.argtypes = [_dp, ctypes.c_size_t, ctypes.c_uint64,
                                            ctypes.c_double, ctypes.c_double]
        L.orc_splitmix64.restype = ctypes.c_uint64
        L.orc_splitmix64.argtypes = [ctypes.c_uint64]
        L.orc_ic_single_seed_masked.argtypes = ([ctypes.c_int] * 3 + [ctypes.c_double] * 9
                                                + [ctypes.c_int, _dp])
        L.orc_pfc_free_energy.restype = ctypes.c_double
        L.orc_pfc_free_energy.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 4 + [_dp]
        L.orc_sum.restype = ctypes.c_double
        L.orc_sum.argtypes = [_dp, ctypes.c_size_t]
        L.orc_sum_sq.restype = ctypes.c_double
        L.orc_sum_sq.argtypes = [_dp, ctypes.c_size_t]
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_fft_seconds.restype = ctypes.c_double
        L.orc_fft_seconds.argtypes = [ctypes.c_int]
        L.orc_set_num_threads.argtypes = [ctypes.c_int]
        L.orc_uniform_real_seq.argtypes = [ctypes.c_uint64, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_int, _dp]
        L.orc_seed_make.argtypes = [_dp, _dp, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, _dp]
        L.orc_seed_grid.argtypes = ([ctypes.c_int] * 2 + [ctypes.c_double] * 2
                                    + [ctypes.c_int] * 2 + [ctypes.c_double] * 4 + [_dp])
        L.orc_random_seeds.argtypes = [ctypes.c_double, ctypes.c_double, _dp]
        L.orc_ic_seeds.argtypes = ([ctypes.c_int] * 3 + [ctypes.c_double] * 6
                                   + [_dp, ctypes.c_int, ctypes.c_int, _dp])
        L.orc_column_max.argtypes = [ctypes.c_int] * 3 + [_dp, _dp]
        L.orc_bc_moving.argtypes = ([ctypes.c_int] * 3 + [ctypes.c_double] * 8
                                    + [ctypes.POINTER(ctypes.c_int), _dp, _dp, _dp])
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.dtype in (np.float64, np.complex128) and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


# -- layout helpers ----------------------------------------------------------
# Arrays are C-ordered numpy arrays of shape (nz, ny, nx): x is the fastest
# index, exactly the reference's linear order (k outer, i inner).

def shape_real(n):
    nx, ny, nz = n
    return (nz, ny, nx)


def shape_complex(n):
    nx, ny, nz = n
    return (nz, ny, nx // 2 + 1)


def mt64_nth(seed: int, n: int) -> int:
    return int(lib().orc_mt64_nth(seed, n))


def uniform_noise(n, seed: int, lo: float, hi: float, base: float = 0.0) -> np.ndarray:
    a = np.full(shape_real(n), base, dtype=np.float64)
    lib().orc_add_uniform_noise(_p(a), a.size, seed, lo, hi)
    return a


def fft1d(x: np.ndarray, sign: int = -1) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().orc_fft1d(a.size, _p(a), sign)
    return a


def r2c(psi: np.ndarray) -> np.ndarray:
    nz, ny, nx = psi.shape
    a = np.ascontiguousarray(psi, dtype=np.float64)
    out = np.empty((nz, ny, nx // 2 + 1), dtype=np.complex128)
    lib().orc_r2c3d(nx, ny, nz, _p(a), _p(out))
    return out


def c2r(spec: np.ndarray, nx: int) -> np.ndarray:
    nz, ny, w = spec.shape
    assert w == nx // 2 + 1
    a = np.ascontiguousarray(spec, dtype=np.complex128)
    out = np.empty((nz, ny, nx), dtype=np.float64)
    lib().orc_c2r3d(nx, ny, nz, _p(a), _p(out), None)
    return out


def dft_r2c_brute(psi: np.ndarray) -> np.ndarray:
    nz, ny, nx = psi.shape
    a = np.ascontiguousarray(psi, dtype=np.float64)
    out = np.empty((nz, ny, nx // 2 + 1), dtype=np.complex128)
    lib().orc_dft_r2c3d_brute(nx, ny, nz, _p(a), _p(out))
    return out


def k_component(index: int, size: int, freq_scale: float) -> float:
    return lib().orc_k_component(index, size, freq_scale)


def spectral_exp_coeffs(L: float, dt: float, thr: float = 1e-12):
    e = ctypes.c_double()
    p = ctypes.c_double()
    lib().orc_spectral_exp_coeffs(L, dt, thr, ctypes.byref(e), ctypes.byref(p))
    return e.value, p.value


def pfc_operators(n, spacing, eps: float, dt: float):
    nx, ny, nz = n
    opL = np.empty(shape_complex(n), dtype=np.float64)
    opN = np.empty(shape_complex(n), dtype=np.float64)
    lib().orc_pfc_operators(nx, ny, nz, *spacing, eps, dt, _p(opL), _p(opN))
    return opL, opN


def pfc_step(psi: np.ndarray, opL: np.ndarray, opN: np.ndarray, nsteps: int) -> np.ndarray:
    nz, ny, nx = psi.shape
    a = np.ascontiguousarray(psi, dtype=np.float64).copy()
    lib().orc_pfc_step(nx, ny, nz, _p(opL), _p(opN), _p(a), nsteps)
    return a


def tungsten_physics_for_mode(k_lap: float, p):
    pa = np.ascontiguousarray(p, dtype=np.float64)
    f = ctypes.c_double()
    c = ctypes.c_double()
    lib().orc_tungsten_physics_for_mode(k_lap, _p(pa), ctypes.byref(f), ctypes.byref(c))
    return f.value, c.value


def tungsten_operators(n, spacing, p, dt: float):
    nx, ny, nz = n
    pa = np.ascontiguousarray(p, dtype=np.float64)
    f = np.empty(shape_complex(n))
    e = np.empty(shape_complex(n))
    w = np.empty(shape_complex(n))
    lib().orc_tungsten_operators(nx, ny, nz, *spacing, _p(pa), dt, _p(f), _p(e), _p(w))
    return f, e, w


def tungsten_step(psi, filterMF, expL, nW, q, nsteps: int):
    nz, ny, nx = psi.shape
    a = np.ascontiguousarray(psi, dtype=np.float64).copy()
    mf = np.empty_like(a)
    qa = np.ascontiguousarray(q, dtype=np.float64)
    lib().orc_tungsten_step(nx, ny, nz, _p(filterMF), _p(expL), _p(nW), _p(qa),
                            _p(a), _p(mf), nsteps)
    return a, mf


def ic_constant(n, value: float) -> np.ndarray:
    a = np.empty(shape_real(n))
    lib().orc_ic_constant(_p(a), a.size, value)
    return a


def pfc_operator_derivs(n, spacing, eps: float, dt: float):
    """d opL/d eps, d opN/d eps over the outbox (unscaled)."""
    nx, ny, nz = n
    a = np.empty(shape_complex(n))
    b = np.empty(shape_complex(n))
    lib().orc_pfc_operator_derivs(nx, ny, nz, *spacing, eps, dt, _p(a), _p(b))
    return a, b


def pfc_adjoint_step(psi, lam, opL, opN, dopL=None, dopN=None):
    """Adjoint of one step from psi; returns (new lambda, dr contribution or None)."""
    nz, ny, nx = psi.shape
    p = np.ascontiguousarray(psi, dtype=np.float64)
    l = np.ascontiguousarray(lam, dtype=np.float64).copy()
    dr = ctypes.c_double(0.0)
    want = dopL is not None
    dl = dopL if want else opL
    dn = dopN if want else opN
    lib().orc_pfc_adjoint_step(nx, ny, nz, _p(opL), _p(opN), _p(dl), _p(dn), _p(p), _p(l),
                               ctypes.cast(ctypes.pointer(dr), _dp) if want else None)
    return l, (dr.value if want else None)


def pfc_free_energy(psi: np.ndarray, spacing, eps: float) -> float:
    """F = dV [sum_k |Psi_k|^2 (eps + (1-k^2)^2) / (2N) + sum psi^4 / 4]."""
    nz, ny, nx = psi.shape
    p = np.ascontiguousarray(psi, dtype=np.float64)
    return float(lib().orc_pfc_free_energy(nx, ny, nz, *map(float, spacing), float(eps), _p(p)))


def bc_fixed(a: np.ndarray, x0: float, dx: float, rho_low: float, rho_high: float,
             xwidth: float = 20.0, alpha: float = 1.0) -> None:
    """FixedBC applied in place to a (nz, ny, nx) float64 array."""
    nz, ny, nx = a.shape
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    lib().orc_bc_fixed(nx, ny, nz, x0, dx, rho_low, rho_high, xwidth, alpha, _p(a))


def ic_single_seed(n, origin, spacing, rho: float, amp: float, radius: float = 64.0):
    nx, ny, nz = n
    a = np.empty(shape_real(n))
    lib().orc_ic_single_seed(nx, ny, nz, *origin, *spacing, rho, amp, radius, _p(a))
    return a


def sum_(a: np.ndarray) -> float:
    a = np.ascontiguousarray(a, dtype=np.float64)
    return lib().orc_sum(_p(a), a.size)


def sum_sq(a: np.ndarray) -> float:
    a = np.ascontiguousarray(a, dtype=np.float64)
    return lib().orc_sum_sq(_p(a), a.size)


def num_threads() -> int:
    return int(lib().orc_num_threads())


def set_num_threads(n: int) -> None:
    lib().orc_set_num_threads(int(n))


def fft_seconds(reset: bool = True) -> float:
    """Seconds pfc_step spent in transforms since the last reset."""
    return float(lib().orc_fft_seconds(int(reset)))


def add_counter_noise(a: np.ndarray, seed: int, lo: float, hi: float) -> np.ndarray:
    """psi += lo + (hi-lo)*u_i with the counter-based splitmix64 stream."""
    assert a.dtype == np.float64 and a.flags.c_contiguous
    lib().orc_add_counter_noise(_p(a), a.size, ctypes.c_uint64(seed), lo, hi)
    return a


def splitmix64(z: int) -> int:
    """The oracle's splitmix64 mixing function (one output of the generator
    whose state is z before its increment)."""
    return int(lib().orc_splitmix64(ctypes.c_uint64(z & ((1 << 64) - 1))))


def ic_single_seed_masked(a: np.ndarray, origin, spacing, rho, amp, radius, keep_outside):
    nz, ny, nx = a.shape
    lib().orc_ic_single_seed_masked(nx, ny, nz, *origin, *spacing, rho, amp, radius,
                                    int(keep_outside), _p(a))
    return a


# -- rotated seeds, SeedGrid, RandomSeeds, MovingBC ----------------------------
# A seed is a row of 24 doubles: location (3), six rotated wave vectors (18),
# rho, radius, amplitude.

def uniform_real_seq(seed: int, a: float, b: float, n: int) -> np.ndarray:
    """First n draws of std::uniform_real_distribution<double>(a, b) on
    std::mt19937_64(seed), as libstdc++ evaluates them."""
    out = np.empty(n)
    lib().orc_uniform_real_seq(ctypes.c_uint64(seed), a, b, n, _p(out))
    return out


def seed_make(location, orientation, radius: float, rho: float, amp: float) -> np.ndarray:
    loc = np.ascontiguousarray(location, dtype=np.float64)
    o = np.ascontiguousarray(orientation, dtype=np.float64)
    out = np.empty(24)
    lib().orc_seed_make(_p(loc), _p(o), radius, rho, amp, _p(out))
    return out


def seed_grid(size, spacing, Ny: int, Nz: int, X0: float, radius: float, rho: float,
              amp: float) -> np.ndarray:
    out = np.empty((Ny * Nz, 24))
    lib().orc_seed_grid(int(size[1]), int(size[2]), float(spacing[1]), float(spacing[2]),
                        Ny, Nz, X0, radius, rho, amp, _p(out))
    return out


def random_seeds(rho: float, amp: float) -> np.ndarray:
    out = np.empty((150, 24))
    lib().orc_random_seeds(rho, amp, _p(out))
    return out


def ic_seeds(a: np.ndarray, origin, spacing, seeds: np.ndarray, keep_outside=False):
    """The field of a seed list, in place on a (nz, ny, nx) array."""
    nz, ny, nx = a.shape
    sd = np.ascontiguousarray(seeds, dtype=np.float64).reshape(-1, 24)
    lib().orc_ic_seeds(nx, ny, nz, *map(float, origin), *map(float, spacing), _p(sd),
                       sd.shape[0], int(keep_outside), _p(a))
    return a


def column_max(a: np.ndarray) -> np.ndarray:
    nz, ny, nx = a.shape
    a = np.ascontiguousarray(a, dtype=np.float64)
    out = np.empty(nx)
    lib().orc_column_max(nx, ny, nz, _p(a), _p(out))
    return out


class MovingBC:
    """MovingBC (one rank) applied in place to (nz, ny, nx) arrays; keeps the
    reference's state (first call, column index, boundary position)."""

    def __init__(self, rho_low, rho_high, xwidth=15.0, alpha=1.0, disp=40.0, xpos=0.0,
                 threshold=0.1):
        self.p = (rho_low, rho_high, xwidth, alpha, disp, threshold)
        self.si = (ctypes.c_int * 2)(1, 0)
        self.sd = np.array([xpos], dtype=np.float64)

    @property
    def xpos(self) -> float:
        return float(self.sd[0])

    @property
    def idx(self) -> int:
        return int(self.si[1])

    def apply(self, a: np.ndarray, x0: float, dx: float) -> np.ndarray:
        nz, ny, nx = a.shape
        assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
        xline = np.empty(nx)
        rl, rh, w, al, disp, thr = self.p
        lib().orc_bc_moving(nx, ny, nz, x0, dx, rl, rh, w, al, disp, thr, self.si,
                            _p(self.sd), _p(xline), _p(a))
        return xline
