"""Free energy of the plain PFC model on the device (pfc_hip_pfc_free_energy)
against the oracle (tests/test_oracle_energy.py pins the oracle by a closed
form, numpy's full-spectrum sum and its decrease along trajectories):
FP64 within 1e-12; the FP32-mixed mode (deviation in FP32, mean and sums in
FP64) within 1e-7 of the FP64 value of the same field; decrease along the
device trajectory."""
import math

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

DX = math.pi / 4


@pytest.mark.parametrize("n", [(64, 64, 64), (48, 48, 48), (128, 64, 32), (256, 256, 256)])
def test_free_energy_matches_oracle(torch_gpu, n):
    import openpfc_amd as P
    torch = torch_gpu
    print(f"free energy n={n}", flush=True)
    psi = O.uniform_noise(n, 42, -0.1, 0.1, base=-0.2)
    f = P.HipFft(*n)
    st = P.PfcStepper(f, (DX,) * 3, -0.25)
    st.initialize(0.5)
    got = st.free_energy(torch.from_numpy(psi).cuda())
    ref = O.pfc_free_energy(psi, (DX,) * 3, -0.25)
    assert abs(got - ref) <= 1e-12 * abs(ref), (got, ref)


def test_free_energy_mixed_mode(torch_gpu):
    import openpfc_amd as P
    torch = torch_gpu
    n = (64, 64, 64)
    psi = O.uniform_noise(n, 43, -0.05, 0.05, base=-0.25)
    m = float(psi.mean())
    d32 = (psi - m).astype(np.float32)
    f = P.HipFft(*n, precision=P.F32)
    st = P.PfcStepper(f, (DX,) * 3, -0.25)
    st.initialize(0.5)
    st.set_mean(m)
    got = st.free_energy(torch.from_numpy(d32).cuda())
    ref = O.pfc_free_energy(m + d32.astype(np.float64), (DX,) * 3, -0.25)
    print(f"mixed free energy rel err {abs(got - ref) / abs(ref):.2e}")
    assert abs(got - ref) <= 1e-7 * abs(ref), (got, ref)


def test_free_energy_decreases_on_device(torch_gpu):
    import openpfc_amd as P
    torch = torch_gpu
    n = (64, 64, 64)
    f = P.HipFft(*n)
    st = P.PfcStepper(f, (DX,) * 3, -0.25)
    st.initialize(0.5)
    psi = torch.from_numpy(O.uniform_noise(n, 1234, -0.01, 0.01, base=-0.2)).cuda()
    F = [st.free_energy(psi)]
    for _ in range(20):
        st.step(psi, 5)
        F.append(st.free_energy(psi))
    assert all(b <= a for a, b in zip(F, F[1:])), F
