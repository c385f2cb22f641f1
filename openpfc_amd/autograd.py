"""Differentiable PFC trajectories: the libtorch autograd::Function built from
csrc/torch_pfc.cpp (openpfc_amd/openpfc_amd_torch.so).

    psiT = pfc_trajectory(psi0, eps, nsteps, dt, spacing, every=20, mixed=False)
    loss = ((psiT - target) ** 2).sum(); loss.backward()
    # psi0.grad, eps.grad: adjoint gradients (pfc_hip_pfc_adjoint_step),
    # checkpoints every `every` steps, segments recomputed in the backward

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
                   mixed: bool = False):
    return load_extension().pfc_trajectory(psi0, eps, int(nsteps), float(dt),
                                           [float(s) for s in spacing], int(every),
                                           bool(mixed))
