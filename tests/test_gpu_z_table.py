"""The z pass with tabulated multipliers (k_tz_mt, the "/mt" variants of role
t_z) against the per-element form (k_tz) of the same plan, lines per block
and exchange mode.

k_tz_mt evaluates each line's multipliers once per distinct kz*kz (slots
0..nz/2) and reads element k's pair from slot k or nz - k; its paired form
shares one table between the lines ky and ny - ky of a block.  The oracle's
operator tables hold the same bits on the planes k and nz - k
(test_oracle_model.py, test_oracle_tungsten.py), and the arithmetic that
applies the pairs is k_tz's, so the two forms must agree bit for bit:
FP64, FP32-mixed and FP32 with FP64 butterflies for the PFC step, and both
Tungsten z passes (mean-field filter, ETD combine).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

DX, DT, EPS = math.pi / 4, 0.5, -0.25
# a grid with the transposed layout (nx/2, ny, nz >= 128) whose 256-point z
# lines fill a wavefront with two lines per block (the paired-table forms)
N = (256, 128, 256)


def _pairs(names):
    """(plain, tabulated) variant indices of the same plan, G and exchange;
    "/mtp" forms hold lines ky and ny - ky in one block with a shared table."""
    out = []
    for i, m in enumerate(names):
        base = m.rsplit("/", 1)[0]
        if m.rsplit("/", 1)[-1] in ("mt", "mtp") and base in names:
            out.append((names.index(base), i))
    assert any(names[i].endswith("/mtp") for _, i in out), names
    return out


def _compare(torch, f, st, psi0, rdt, nsteps):
    layouts = f.variants("step_layout")
    f.set_variant("step_layout", layouts.index("transposed"))
    names = f.variants("t_z")
    for plain, mt in _pairs(names):
        outs = []
        for v in (plain, mt):
            print(f"t_z {names[v]}", flush=True)
            f.set_variant("t_z", v)
            psi = torch.from_numpy(psi0).to(rdt).cuda()
            st.step(psi, nsteps)
            outs.append(psi)
        assert torch.equal(outs[0], outs[1]), (names[plain], names[mt])


@pytest.mark.parametrize("prec", ["f64", "f32", "f32xp"])
def test_pfc_tabulated_z_same_bits(torch_gpu, prec):
    import openpfc_amd as P
    torch = torch_gpu
    flags = P.FLAG_F64_BUTTERFLIES if prec == "f32xp" else 0
    f = P.HipFft(*N, precision=P.F64 if prec == "f64" else P.F32, flags=flags)
    st = P.PfcStepper(f, (DX, DX, DX), EPS)
    st.initialize(DT)
    base = -0.2
    if prec != "f64":
        st.set_mean(base)
        base = 0.0
    psi0 = O.uniform_noise(N, 47, -0.01, 0.01, base=base)
    _compare(torch, f, st, psi0, torch.float64 if prec == "f64" else torch.float32, 2)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_tungsten_tabulated_z_same_bits(torch_gpu, prec):
    import openpfc_amd as P
    from openpfc_amd import TungstenParams
    from test_gpu_tungsten import DX_G, GOLDEN_PARAMS
    torch = torch_gpu
    f = P.HipFft(*N, precision=P.F64 if prec == "f64" else P.F32)
    st = P.TungstenStepper(f, (DX_G,) * 3, TungstenParams(**GOLDEN_PARAMS))
    st.initialize(1.0)
    psi0 = O.uniform_noise(N, 83, -0.05, 0.05, base=-0.1)
    _compare(torch, f, st, psi0, torch.float64 if prec == "f64" else torch.float32, 2)
