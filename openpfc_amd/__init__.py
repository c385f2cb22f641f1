"""openpfc_amd -- MI355X-native spectral PFC time step (HIP, gfx950).

Python access to the C ABI (``include/pfc_hip.h``), for tests and the
benchmark.  The reference-shaped host API (``pfc::fft::IFFT``, ``Model``,
``Simulator``) lives in C++ under ``openpfc_amd/host/``.

``HipFft`` mirrors the reference's ``FFT_Impl<rocfft>`` DataBuffer overloads
(include/openpfc/kernel/fft/detail/fft_heffte_backend.hpp:146-227): sizes are
checked and a mismatch raises ``ValueError`` (the reference throws
std::invalid_argument) before anything is enqueued.
"""
from __future__ import annotations

import ctypes
import os
import time

from . import _lib
from ._lib import (F32, F64, FLAG_F64_BUTTERFLIES, MODEL_PFC, MODEL_TUNGSTEN, PFC_HIP_OK,
                   TUNGSTEN_NPARAMS, PfcHipError, check, load)
from .ensemble import Ensemble, Run
from .tungsten import TungstenParams

__all__ = ["HipFft", "PfcStepper", "TungstenStepper", "TungstenParams", "Ensemble", "Run",
           "PfcHipError", "F32", "F64", "FLAG_F64_BUTTERFLIES", "load"]


def _stream_ptr(stream) -> ctypes.c_void_p:
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(int(getattr(stream, "cuda_stream", stream)))


def _ptr(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(int(t.data_ptr()))


class HipFft:
    """3-D R2C/C2R on one GPU with the reference's one-rank layout.

    Real boxes are (nz, ny, nx) tensors (x fastest), spectra are
    (nz, ny, nx//2+1) complex tensors.  ``precision`` is F64 or F32.
    """

    def __init__(self, nx: int, ny: int, nz: int, precision: int = F64, flags: int = 0):
        self._lib = load()
        h = ctypes.c_void_p()
        check(self._lib.pfc_hip_plan_create(nx, ny, nz, precision, int(flags), ctypes.byref(h)))
        self._h = h
        self.n = (nx, ny, nz)
        self.precision = precision
        self.flags = int(flags)
        self._fft_time = 0.0

    # -- lifetime -----------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.pfc_hip_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    # -- sizes (IFFT::size_inbox/size_outbox/size_workspace) -----------------
    def _sizes(self):
        a, b, c = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_size_t()
        check(self._lib.pfc_hip_sizes(self._h, ctypes.byref(a), ctypes.byref(b),
                                      ctypes.byref(c)))
        return a.value, b.value, c.value

    def size_inbox(self) -> int:
        return self._sizes()[0]

    def size_outbox(self) -> int:
        return self._sizes()[1]

    def size_workspace(self) -> int:
        return self._sizes()[2]

    def get_allocated_memory_bytes(self) -> int:
        b = ctypes.c_size_t()
        check(self._lib.pfc_hip_allocated_bytes(self._h, ctypes.byref(b)))
        return b.value

    def reset_fft_time(self) -> None:
        self._fft_time = 0.0

    def get_fft_time(self) -> float:
        return self._fft_time

    # -- launch variants ------------------------------------------------------
    ROLES = {"x_r2c": 0, "x_c2r": 1, "col_y": 2, "col_z": 3, "pfc_x": 4, "pfc_z": 5,
             "step_layout": 6, "t_x_fwd": 7, "t_y_fwd": 8, "t_z": 9, "t_y_inv": 10,
             "t_x_inv": 11, "x_tung": 12, "t_x_r2c": 13, "t_x_tung": 14, "t_x_fuse": 15,
             "step_fuse": 16}

    def variants(self, role) -> list:
        r = self.ROLES.get(role, role)
        c = ctypes.c_int()
        check(self._lib.pfc_hip_variant_count(self._h, r, ctypes.byref(c)))
        return [self._lib.pfc_hip_variant_name(self._h, r, i).decode() for i in range(c.value)]

    def get_variant(self, role) -> int:
        c = ctypes.c_int()
        check(self._lib.pfc_hip_variant_get(self._h, self.ROLES.get(role, role), ctypes.byref(c)))
        return c.value

    def set_variant(self, role, index: int) -> None:
        check(self._lib.pfc_hip_variant_set(self._h, self.ROLES.get(role, role), int(index)))

    # -- transforms -----------------------------------------------------------
    def _check(self, real, cplx, what: str):
        import torch
        rdt = torch.float64 if self.precision == F64 else torch.float32
        cdt = torch.complex128 if self.precision == F64 else torch.complex64
        if real.numel() != self.size_inbox():
            raise ValueError(f"HipFft::{what}: real buffer size {real.numel()} != "
                             f"size_inbox() {self.size_inbox()}")
        if cplx.numel() != self.size_outbox():
            raise ValueError(f"HipFft::{what}: complex buffer size {cplx.numel()} != "
                             f"size_outbox() {self.size_outbox()}")
        if real.dtype != rdt or cplx.dtype != cdt:
            raise ValueError(f"HipFft::{what}: dtypes {real.dtype}/{cplx.dtype} do not "
                             f"match the plan precision")
        if not (real.is_cuda and cplx.is_cuda):
            raise ValueError(f"HipFft::{what}: buffers must be device tensors")
        if not (real.is_contiguous() and cplx.is_contiguous()):
            raise ValueError(f"HipFft::{what}: buffers must be contiguous")

    def forward(self, real_in, cplx_out, stream=None):
        self._check(real_in, cplx_out, "forward")
        t0 = time.perf_counter()
        check(self._lib.pfc_hip_r2c(self._h, _ptr(real_in), _ptr(cplx_out),
                                    _stream_ptr(stream)))
        self._fft_time += time.perf_counter() - t0

    def backward(self, cplx_in, real_out, stream=None):
        self._check(real_out, cplx_in, "backward")
        t0 = time.perf_counter()
        check(self._lib.pfc_hip_c2r(self._h, _ptr(cplx_in), _ptr(real_out),
                                    _stream_ptr(stream)))
        self._fft_time += time.perf_counter() - t0


def _check_field(fft: HipFft, psi, what: str):
    import torch
    rdt = torch.float64 if fft.precision == F64 else torch.float32
    if psi.numel() != fft.size_inbox() or psi.dtype != rdt or not psi.is_cuda \
            or not psi.is_contiguous():
        raise ValueError(f"{what}: psi must be a contiguous device tensor of "
                         f"size_inbox() {fft.size_inbox()} elements of {rdt}")


def ic_constant(fft: HipFft, psi, value: float, stream=None) -> None:
    """Constant::apply (initial_conditions/constant.hpp:82-89) on the device."""
    _check_field(fft, psi, "ic_constant")
    check(fft._lib.pfc_hip_ic_constant(fft.handle, _ptr(psi), float(value),
                                       _stream_ptr(stream)))


def ic_single_seed(fft: HipFft, psi, origin, spacing, rho: float, amplitude: float,
                   radius: float = 64.0, keep_outside: bool = False, stream=None) -> None:
    """SingleSeed::apply (initial_conditions/single_seed.hpp:51-83) on the device.

    keep_outside=False writes 0.0 outside the sphere, as the reference does;
    True leaves the field untouched there."""
    _check_field(fft, psi, "ic_single_seed")
    o = (ctypes.c_double * 3)(*origin)
    s = (ctypes.c_double * 3)(*spacing)
    check(fft._lib.pfc_hip_ic_single_seed(fft.handle, _ptr(psi), o, s, float(rho),
                                          float(amplitude), float(radius),
                                          int(bool(keep_outside)), _stream_ptr(stream)))


def ic_seeds(fft: HipFft, psi, origin, spacing, seeds, keep_outside: bool = False,
             stream=None) -> None:
    """The field of a rotated-seed list on the device (the field::apply of
    SeedGrid / RandomSeeds, initial_conditions/seed_grid.hpp:108-118).  seeds:
    rows of 24 doubles (location, six wave vectors, rho, radius, amplitude);
    the first seed holding a point gives its value, 0.0 outside every seed
    (or the field kept with keep_outside)."""
    import numpy as np
    _check_field(fft, psi, "ic_seeds")
    rows = np.ascontiguousarray(seeds, dtype=np.float64).reshape(-1, 24)
    o = (ctypes.c_double * 3)(*origin)
    s = (ctypes.c_double * 3)(*spacing)
    check(fft._lib.pfc_hip_ic_seeds(fft.handle, _ptr(psi), o, s,
                                    rows.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                    rows.shape[0], int(bool(keep_outside)),
                                    _stream_ptr(stream)))


def column_max(fft: HipFft, psi, stream=None):
    """MovingBC's column profile: the maximum over (y, z) at every x index
    (numpy array of nx doubles; synchronises the stream)."""
    import numpy as np
    _check_field(fft, psi, "column_max")
    out = np.empty(fft.n[0])
    check(fft._lib.pfc_hip_column_max(fft.handle, _ptr(psi),
                                      out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                      _stream_ptr(stream)))
    return out


def bc_moving(fft: HipFft, psi, x0: float, dx: float, xpos: float, rho_low: float,
              rho_high: float, xwidth: float = 15.0, alpha: float = 1.0, stream=None) -> None:
    """MovingBC::fill_bc (boundary_conditions/moving_bc.hpp:187-219) on the
    device: the sigmoid band of half-width xwidth round fmod(xpos, nx*dx)."""
    _check_field(fft, psi, "bc_moving")
    check(fft._lib.pfc_hip_bc_moving(fft.handle, _ptr(psi), float(x0), float(dx), float(xpos),
                                     float(rho_low), float(rho_high), float(xwidth),
                                     float(alpha), _stream_ptr(stream)))


def ic_add_noise(fft: HipFft, psi, seed: int, lo: float, hi: float, stream=None) -> None:
    """psi += lo + (hi-lo)*u_i, u_i from the counter-based splitmix64 stream."""
    _check_field(fft, psi, "ic_add_noise")
    check(fft._lib.pfc_hip_ic_add_noise(fft.handle, _ptr(psi), ctypes.c_ulonglong(seed),
                                        float(lo), float(hi), _stream_ptr(stream)))


class _Stepper:
    """Shared parts of the device-resident steppers (field checks, pass
    timing, autotuning).  Subclasses set NPASSES, TUNE_ROLES (tuned roles in
    the library's slot order: natural-layout roles first, NATURAL_ROLES of
    them) and the operator setup."""

    NPASSES = 5
    TUNE_ROLES: list = []
    NATURAL_ROLES = 0

    def _check_psi(self, psi, what="step"):
        import torch
        rdt = torch.float64 if self.fft.precision == F64 else torch.float32
        if psi.numel() != self.fft.size_inbox() or psi.dtype != rdt or not psi.is_cuda \
                or not psi.is_contiguous():
            raise ValueError(f"{type(self).__name__}.{what}: psi must be a contiguous "
                             f"device tensor of size_inbox() {self.fft.size_inbox()} "
                             f"elements of {rdt}")

    def autotune(self, psi, reps: int = 2, stream=None) -> dict:
        """Time every variant of the step's passes on psi (restored afterwards),
        keep the fastest per pass and the faster layout; returns
        {role: [(name, ms), ...], "layout_ms": {...}, "chosen": {...}}."""
        self._check_psi(psi, "autotune")
        roles = list(self.TUNE_ROLES)
        layouts = self.fft.variants("step_layout")
        if "transposed" not in layouts:
            roles = roles[:self.NATURAL_ROLES]
        elif "natural" not in layouts:  # transposed only (FP64-butterfly plans)
            roles = roles[self.NATURAL_ROLES:]
        names = {r: self.fft.variants(r) for r in roles}
        total = sum(len(v) for v in names.values()) + 2
        out = (ctypes.c_double * total)()
        check(self.fft._lib.pfc_hip_pfc_autotune(self.fft.handle, _ptr(psi), int(reps),
                                                 _stream_ptr(stream), out, total))
        res, i = {}, 0
        for r in roles:
            res[r] = [(n, out[i + k]) for k, n in enumerate(names[r])]
            i += len(names[r])
        res["layout_ms"] = {"transposed": out[i], "natural": out[i + 1]}
        res["chosen"] = {r: names[r][self.fft.get_variant(r)] for r in roles if names[r]}
        res["chosen"]["step_layout"] = self.fft.variants("step_layout")[
            self.fft.get_variant("step_layout")]
        return res

    def step_timed(self, psi, nsteps: int, stream=None):
        """Run nsteps with HIP events around each pass; returns the mean
        duration of each of the NPASSES passes in ms (stream synchronised)."""
        self._check_psi(psi, "step_timed")
        out = (ctypes.c_double * self.NPASSES)()
        check(self.fft._lib.pfc_hip_pfc_step_timed(self.fft.handle, _ptr(psi), int(nsteps),
                                                   _stream_ptr(stream), out, self.NPASSES))
        return list(out)


class PfcStepper(_Stepper):
    """Device-resident plain conserved PFC ETD1 stepper over a HipFft plan.

    Mirrors PFCModel::initialize/step (docs/api/examples/10_complete_pfc_model.cpp:
    94-144,166-213): ``initialize(dt)`` builds opL/opN on the device,
    ``step(psi, n)`` advances ``psi`` in place by n steps.
    """

    NPASSES = 5
    TUNE_ROLES = ["pfc_x", "col_y", "pfc_z", "x_c2r", "t_x_fwd", "t_y_fwd", "t_z",
                  "t_y_inv", "t_x_inv", "t_x_fuse"]
    NATURAL_ROLES = 4

    def __init__(self, fft: HipFft, spacing, epsilon: float):
        self.fft = fft
        self.spacing = tuple(float(s) for s in spacing)
        self.epsilon = float(epsilon)
        self.dt = None

    def step_timed_fused(self, psi, nsteps: int, stream=None):
        """The launches of a fused multi-step run (nsteps >= 2): returns
        [forward x (first step), y forward, z, y inverse (means), inverse x
        (last step), fused boundary (mean)] in ms; the last entry is 0 when
        the plan does not fuse step boundaries."""
        self._check_psi(psi, "step_timed_fused")
        out = (ctypes.c_double * 6)()
        check(self.fft._lib.pfc_hip_pfc_step_timed(self.fft.handle, _ptr(psi), int(nsteps),
                                                   _stream_ptr(stream), out, 6))
        return list(out)

    def initialize(self, dt: float) -> None:
        params = (ctypes.c_double * 4)(*self.spacing, self.epsilon)
        check(self.fft._lib.pfc_hip_pfc_set_operators(self.fft.handle, MODEL_PFC, params,
                                                      4, float(dt)))
        self.dt = float(dt)

    def set_mean(self, mean, enable: bool = True) -> None:
        """Mean-split (FP32-mixed) mode: the field passed to step() then holds
        the deviation psi - mean; `mean` is kept in double and conserved."""
        check(self.fft._lib.pfc_hip_pfc_set_mean(self.fft.handle, int(bool(enable)),
                                                 float(mean)))
        self.mean = float(mean) if enable else None

    def step(self, psi, nsteps: int = 1, stream=None) -> None:
        self._check_psi(psi)
        check(self.fft._lib.pfc_hip_pfc_step(self.fft.handle, _ptr(psi), int(nsteps),
                                             _stream_ptr(stream)))

    def free_energy(self, psi, stream=None) -> float:
        """F = dV sum [psi (eps + (1+lap)^2) psi / 2 + psi^4 / 4], summed in
        double (psi the deviation in the mean-split mode)."""
        self._check_psi(psi, "free_energy")
        out = ctypes.c_double()
        check(self.fft._lib.pfc_hip_pfc_free_energy(self.fft.handle, _ptr(psi),
                                                    ctypes.byref(out), _stream_ptr(stream)))
        return out.value

    def operators(self, stream=None):
        """(opL, opN) as device tensors in the outbox layout, scaled by 1/(nx ny nz)."""
        import torch
        dt = torch.float64 if self.fft.precision == F64 else torch.float32
        nx, ny, nz = self.fft.n
        opL = torch.empty((nz, ny, nx // 2 + 1), dtype=dt, device="cuda")
        opN = torch.empty_like(opL)
        check(self.fft._lib.pfc_hip_pfc_operators(self.fft.handle, _ptr(opL), _ptr(opN),
                                                  _stream_ptr(stream)))
        return opL, opN


class TungstenStepper(_Stepper):
    """Device-resident Tungsten mean-field PFC stepper over a HipFft plan: the
    TungstenHIP drop-in (apps/tungsten/include/tungsten/hip/tungsten_model.hpp:
    57-330; the step of cpu/tungsten_model.hpp:306-366).

    ``initialize(dt)`` sets the operators from ``params`` (a TungstenParams:
    the reference's base parameters, derived ones computed as the reference
    does); ``step(psi, n, psi_mf)`` advances psi in place and, when psi_mf is
    given, leaves in it the mean field of the last step's input (what
    TungstenHIP::get_psiMF holds after step()).
    """

    NPASSES = 9
    TUNE_ROLES = ["x_r2c", "col_y", "pfc_z", "x_tung", "x_c2r", "t_x_r2c", "t_y_fwd",
                  "t_z", "t_y_inv", "t_x_tung", "t_x_inv"]
    NATURAL_ROLES = 5

    def __init__(self, fft: HipFft, spacing, params: TungstenParams | None = None):
        self.fft = fft
        self.spacing = tuple(float(s) for s in spacing)
        self.params = params if params is not None else TungstenParams()
        self.dt = None

    def initialize(self, dt: float) -> None:
        vals = self.params.abi_params(self.spacing)
        arr = (ctypes.c_double * TUNGSTEN_NPARAMS)(*vals)
        check(self.fft._lib.pfc_hip_pfc_set_operators(self.fft.handle, MODEL_TUNGSTEN, arr,
                                                      TUNGSTEN_NPARAMS, float(dt)))
        self.dt = float(dt)

    def step(self, psi, nsteps: int = 1, psi_mf=None, stream=None) -> None:
        self._check_psi(psi)
        if psi_mf is not None:
            self._check_psi(psi_mf, "step (psi_mf)")
        check(self.fft._lib.pfc_hip_tungsten_step(
            self.fft.handle, _ptr(psi), int(nsteps),
            _ptr(psi_mf) if psi_mf is not None else None, _stream_ptr(stream)))

    def operators(self, stream=None):
        """(filterMF, expL, nW) as device tensors in the outbox layout, each
        scaled by 1/(nx ny nz) as the step applies them."""
        import torch
        dt = torch.float64 if self.fft.precision == F64 else torch.float32
        nx, ny, nz = self.fft.n
        f = torch.empty((nz, ny, nx // 2 + 1), dtype=dt, device="cuda")
        e = torch.empty_like(f)
        n = torch.empty_like(f)
        check(self.fft._lib.pfc_hip_tungsten_operators(self.fft.handle, _ptr(f), _ptr(e),
                                                       _ptr(n), _stream_ptr(stream)))
        return f, e, n


# -- autotune cache (the file format of the C++ layer's autotune.hpp) --------
def device_id(device: int = 0) -> str:
    """'<arch>_<CUs>cu' of a device, e.g. 'gfx950_256cu': the cache file's prefix."""
    import torch
    p = torch.cuda.get_device_properties(device)
    return p.gcnArchName.split(":")[0] + f"_{p.multi_processor_count}cu"


def step_key(model: str, precision: int, n, mean_split: bool = False) -> str:
    """Cache key of a step: model/precision/extents[/mean_split]."""
    return (f"{model}/{'f64' if precision == F64 else 'f32'}/{n[0]}x{n[1]}x{n[2]}"
            + ("/mean_split" if mean_split else ""))


def autotune_cache_path(cache_dir: str | None = None, device: int = 0) -> str:
    """<cache_dir>/<device>_autotune_cache.json; cache_dir defaults to
    OPENPFC_GPU_AUTOTUNE_CACHE_DIR, else the working directory."""
    d = cache_dir or os.environ.get("OPENPFC_GPU_AUTOTUNE_CACHE_DIR", ".")
    return os.path.join(d, f"{device_id(device)}_autotune_cache.json")


def cache_entry(fft: HipFft) -> dict:
    """The plan's current variant name per step role (keys: role numbers)."""
    out = {}
    for role in range(64):
        c = ctypes.c_int()
        if fft._lib.pfc_hip_variant_count(fft._h, role, ctypes.byref(c)) != PFC_HIP_OK:
            break
        if c.value > 0:
            out[str(role)] = fft.variants(role)[fft.get_variant(role)]
    return out


def store_cache_entry(fft: HipFft, key: str, cache_dir: str | None = None) -> str:
    """Add the plan's variants under `key` to the cache file (atomically
    replaced); returns the file's path."""
    import json
    path = autotune_cache_path(cache_dir)
    try:
        with open(path) as fh:
            cache = json.load(fh)
        if not isinstance(cache, dict):
            cache = {}
    except (OSError, ValueError):
        cache = {}
    cache["device_id"] = device_id()
    cache.setdefault("entries", {})[key] = cache_entry(fft)
    tmp = f"{path}.tmp{os.getpid()}"
    with open(tmp, "w") as fh:
        json.dump(cache, fh, indent=1)
    os.replace(tmp, path)
    return path


def apply_cache_entry(fft: HipFft, key: str, cache_dir: str | None = None) -> bool:
    """Select the cached variants of `key`; False (nothing changed) when the
    file has no such entry or it names a variant the plan does not offer."""
    import json
    try:
        with open(autotune_cache_path(cache_dir)) as fh:
            entry = json.load(fh).get("entries", {}).get(key)
    except (OSError, ValueError, AttributeError):
        return False
    if not isinstance(entry, dict):
        return False
    picks = []
    for role in range(64):
        c = ctypes.c_int()
        if fft._lib.pfc_hip_variant_count(fft._h, role, ctypes.byref(c)) != PFC_HIP_OK:
            break
        if c.value <= 0:
            continue
        names = fft.variants(role)
        if entry.get(str(role)) not in names:
            return False
        picks.append((role, names.index(entry[str(role)])))
    for role, idx in picks:
        fft.set_variant(role, idx)
    return True
