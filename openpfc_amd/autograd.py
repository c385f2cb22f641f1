"""Differentiable PFC trajectories: the libtorch autograd::Function built from
csrc/torch_pfc.cpp (openpfc_amd/openpfc_amd_torch.so).

    psiT = pfc_trajectory(psi0, eps, nsteps, dt, spacing, every=20, mixed=False, snaps=0)
    loss = ((psiT - target) ** 2).sum(); loss.backward()
    # psi0.grad, eps.grad: adjoint gradients (pfc_hip_pfc_adjoint_step)

Checkpoints: snaps=0 keeps one every `every` steps and recomputes one
segment at a time in the backward; snaps >= 1 uses the binomial (revolve)
schedule with at most `snaps` fields held besides psi0, recomputing the
fewest steps that budget allows.  Both give the same gradients bit for bit;
schedule_cost(nsteps, every, snaps) gives (recomputed steps, fields held).
All work runs on torch's current stream.

psi0: contiguous float64 device tensor (nz, ny, nx); eps: float64 scalar
tensor.  mixed=True steps and differentiates in the FP32-mixed mode (FP32
deviation from the FP64 mean, conserved exactly); inputs and outputs stay
float64.  Loading fails loudly when the extension was not built.
"""
from __future__ import annotations

import importlib.machinery
import importlib.util
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
EXT_PATH = os.path.join(_HERE, "openpfc_amd_torch.so")
_ext = None


def load_extension():
    global _ext
    if _ext is not None:
        return _ext
    import torch  # noqa: F401  (the extension links against torch's libraries)
    if not os.path.exists(EXT_PATH):
        raise FileNotFoundError(f"{EXT_PATH} not found: build it with "
                                "`python -c 'import __graft_entry__ as g; g.build()'`")
    loader = importlib.machinery.ExtensionFileLoader("openpfc_amd_torch", EXT_PATH)
    spec = importlib.util.spec_from_file_location("openpfc_amd_torch", EXT_PATH, loader=loader)
    mod = importlib.util.module_from_spec(spec)
    loader.exec_module(mod)
    _ext = mod
    return mod


def pfc_trajectory(psi0, eps, nsteps: int, dt: float, spacing, every: int = 20,
                   mixed: bool = False, snaps: int = 0):
    return load_extension().pfc_trajectory(psi0, eps, int(nsteps), float(dt),
                                           [float(s) for s in spacing], int(every),
                                           bool(mixed), int(snaps))


def schedule_cost(nsteps: int, every: int = 20, snaps: int = 0):
    """(steps recomputed by the backward, fields held besides psi0) of the
    uniform schedule (snaps=0, a checkpoint every `every` steps) or of the
    binomial schedule with `snaps` slots."""
    return tuple(load_extension().schedule_cost(int(nsteps), int(every), int(snaps)))


def binomial_costs(nsteps: int, snaps: int) -> list:
    """Forward steps of the optimal reversal of m = 1..nsteps steps with the
    start state held and `snaps` free slots (the schedule's recurrence)."""
    return [int(c) for c in load_extension().binomial_costs(int(nsteps), int(snaps))]
