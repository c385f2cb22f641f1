"""Ensembles of independent plain-PFC runs of uneven sizes on one GPU.

``Ensemble`` wraps ``pfc_hip_ensemble_*`` (``include/pfc_hip.h``): runs
with equal extents, spacing and dt form a group stepped by one batched plan
(transposed layout where the extents have it: the group's fields one after
another, each with its own eps and mean as launch arguments; else the
natural layout: one launch per pass for all of the group's fields, per-run
eps and mean in device memory), and one step of every group is captured
once into a hipGraph, so each step of the whole ensemble is a single graph
launch.
The fields of all runs live in one device tensor owned by this object;
``field(r)`` is a (nz, ny, nx) view of run r.

No reference counterpart: the reference steps one Model per Simulator
(include/openpfc/frontend/simulator.hpp:549-553); each run here is that
model's step (docs/api/examples/10_complete_pfc_model.cpp:121-144).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

from ._lib import F64, check, load

PARAMS_PER_RUN = 6  # dx, dy, dz, eps, dt, mean
_ROLES = {"x_c2r": 1, "col_y": 2, "pfc_x": 4, "pfc_z": 5}
_T_ROLES = {"t_x_fwd": 7, "t_y_fwd": 8, "t_z": 9, "t_y_inv": 10, "t_x_inv": 11,
            "t_x_fuse": 15, "step_fuse": 16}
_LAYOUT = 6


@dataclass(frozen=True)
class Run:
    """One plain-PFC run: extents (nx, ny, nz), spacing (dx, dy, dz), eps (r),
    dt and the mean of its field (used by the mean-split step)."""
    ext: tuple
    spacing: tuple
    eps: float
    dt: float
    mean: float = 0.0

    @staticmethod
    def cubic(n: int, dx: float, eps: float, dt: float, mean: float = 0.0) -> "Run":
        return Run((n, n, n), (dx, dx, dx), float(eps), float(dt), float(mean))

    def params(self) -> tuple:
        return (*map(float, self.spacing), float(self.eps), float(self.dt), float(self.mean))

    @property
    def size(self) -> int:
        return self.ext[0] * self.ext[1] * self.ext[2]


def _arrays(runs):
    n = len(runs)
    ext = (ctypes.c_int * (3 * n))(*[int(v) for r in runs for v in r.ext])
    par = (ctypes.c_double * (PARAMS_PER_RUN * n))(*[v for r in runs for v in r.params()])
    return ext, par


def layout(runs) -> tuple[list[int], int]:
    """(offsets, total): index of the first value of every run in the shared
    buffer and the buffer's length (a host computation of the library)."""
    lib = load()
    ext, par = _arrays(runs)
    offs = (ctypes.c_size_t * len(runs))()
    total = ctypes.c_size_t()
    check(lib.pfc_hip_ensemble_layout(len(runs), ext, par, offs, ctypes.byref(total)))
    return list(offs), total.value


def _stream_ptr(stream) -> ctypes.c_void_p:
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(int(getattr(stream, "cuda_stream", stream)))


class Ensemble:
    """Runs advanced together; ``mean_split`` selects the FP32-mixed
    mean-split step (each field then holds psi - mean of its run)."""

    def __init__(self, runs, precision: int = F64, mean_split: bool = False):
        import torch
        self._lib = load()
        self.runs = list(runs)
        self.precision = precision
        self.mean_split = bool(mean_split)
        self.offsets, self.total = layout(self.runs)
        dt = torch.float64 if precision == F64 else torch.float32
        self.buffer = torch.zeros(self.total, dtype=dt, device="cuda")
        ext, par = _arrays(self.runs)
        h = ctypes.c_void_p()
        check(self._lib.pfc_hip_ensemble_create(len(self.runs), ext, par, int(precision),
                                                int(self.mean_split),
                                                ctypes.c_void_p(self.buffer.data_ptr()),
                                                ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.pfc_hip_ensemble_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def dof(self) -> int:
        return sum(r.size for r in self.runs)

    def field(self, r: int):
        """(nz, ny, nx) view of run r's field in the shared buffer."""
        nx, ny, nz = self.runs[r].ext
        o = self.offsets[r]
        return self.buffer[o:o + nx * ny * nz].view(nz, ny, nx)

    def step(self, nsteps: int = 1, graph: bool = True, stream=None) -> None:
        check(self._lib.pfc_hip_ensemble_step(self._h, int(nsteps), int(bool(graph)),
                                              _stream_ptr(stream)))

    def autotune(self, reps: int = 2, stream=None) -> None:
        check(self._lib.pfc_hip_ensemble_autotune(self._h, int(reps), _stream_ptr(stream)))

    def groups(self) -> list[dict]:
        """Per group: extents, number of runs, first value, the step layout
        and the chosen variant of every pass of that layout's step."""
        n = ctypes.c_int()
        check(self._lib.pfc_hip_ensemble_groups(self._h, ctypes.byref(n)))
        out = []
        for g in range(n.value):
            plan, batch, off = ctypes.c_void_p(), ctypes.c_int(), ctypes.c_size_t()
            check(self._lib.pfc_hip_ensemble_group(self._h, g, ctypes.byref(plan),
                                                   ctypes.byref(batch), ctypes.byref(off)))
            def name_of(role):
                idx = ctypes.c_int()
                check(self._lib.pfc_hip_variant_get(plan, role, ctypes.byref(idx)))
                return self._lib.pfc_hip_variant_name(plan, role, idx.value).decode()
            layout = name_of(_LAYOUT)
            chosen = {"step_layout": layout}
            for name, role in (_T_ROLES if layout == "transposed" else _ROLES).items():
                cnt = ctypes.c_int()
                check(self._lib.pfc_hip_variant_count(plan, role, ctypes.byref(cnt)))
                if cnt.value > 0:
                    chosen[name] = name_of(role)
            first = next(r for r, o in enumerate(self.offsets) if o == off.value)
            out.append({"ext": tuple(self.runs[first].ext), "batch": batch.value,
                        "offset": off.value, "variants": chosen})
        return out
