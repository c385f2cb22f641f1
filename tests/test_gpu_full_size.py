"""The 1024^3 FP64 drop-in workload at its own size.

Workload: plain conserved PFC on a 1024^3 periodic grid, dx = pi/4, dt = 0.5,
r = -0.25; psi0 = -0.2 + U(-0.01, 0.01) (counter-based splitmix64 stream,
seed 1234) with a one-mode bcc seed (rho = -0.2, amplitude 0.1, radius
64*dx) written inside the sphere and the noise kept outside, on a grid whose
origin is -N*dx/2 per axis (the benchmark's field).  The step is
docs/api/examples/10_complete_pfc_model.cpp:121-144.

* the device initial field equals the oracle's: the noise bit for bit, the
  seed points within 1e-15 (device against host cosines);
* 3 device steps against 3 oracle steps: max|d| / max|psi_ref| <= 1e-10;
* continued to 1000 steps: finite, the mean conserved to 1e-12 and the free
  energy non-increasing at every 100 steps.

The oracle's three 1024^3 steps take about half a minute on 16 host cores
and about 45 GB of host memory.
"""
import math
import time

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

N = 1024
DX = math.pi / 4
DT = 0.5
EPS = -0.25
SEED_AMP = 0.1


def _device_field(P, torch, f):
    psi = torch.empty(N ** 3, dtype=torch.float64, device="cuda")
    origin = (-0.5 * N * DX,) * 3
    P.ic_constant(f, psi, -0.2)
    P.ic_add_noise(f, psi, 1234, -0.01, 0.01)
    P.ic_single_seed(f, psi, origin, (DX,) * 3, -0.2, SEED_AMP, 64 * DX, keep_outside=True)
    torch.cuda.synchronize()
    return psi


def _oracle_field():
    a = np.full((N, N, N), -0.2)
    O.add_counter_noise(a, 1234, -0.01, 0.01)
    O.ic_single_seed_masked(a, (-0.5 * N * DX,) * 3, (DX,) * 3, -0.2, SEED_AMP, 64 * DX, 1)
    return a


def test_full_size_fp64_step_matches_oracle_and_keeps_invariants(torch_gpu):
    import openpfc_amd as P
    torch = torch_gpu
    print("1024^3 FP64: device initial field", flush=True)
    f = P.HipFft(N, N, N)
    st = P.PfcStepper(f, (DX,) * 3, EPS)
    st.initialize(DT)
    psi = _device_field(P, torch, f)
    t0 = time.perf_counter()
    ref = _oracle_field()
    got0 = psi.view(N, N, N).cpu().numpy()
    d = got0 - ref
    del got0
    # only seed points (device vs host cosines) may differ, by rounding only
    mx = float(np.abs(d).max())
    assert mx <= 1e-15, f"initial fields differ by {mx:.3e}"
    k, j, i = np.nonzero(d)
    del d
    x = lambda v: v * DX - 0.5 * N * DX  # noqa: E731
    rr = x(i) ** 2 + x(j) ** 2 + x(k) ** 2
    assert np.all(rr < (64 * DX) ** 2), "the noise outside the seed differs"
    print(f"{i.size} seed points differ by <= {mx:.1e}", flush=True)
    print(f"initial fields agree ({time.perf_counter() - t0:.1f} s); oracle 3 steps", flush=True)
    # the oracle steps from the same bits the device holds
    ref = psi.view(N, N, N).cpu().numpy().copy()
    mean0 = float(torch.mean(psi).item())
    F0 = st.free_energy(psi)
    opL, opN = O.pfc_operators((N, N, N), (DX,) * 3, EPS, DT)
    t0 = time.perf_counter()
    ref = O.pfc_step(ref, opL, opN, 3)
    del opL, opN
    print(f"oracle 3 steps {time.perf_counter() - t0:.1f} s on {O.num_threads()} threads",
          flush=True)
    st.step(psi, 3)
    torch.cuda.synchronize()
    got = psi.view(N, N, N).cpu().numpy()
    err = float(np.abs(got - ref).max() / np.abs(ref).max())
    print(f"3 steps: rel max err {err:.3e}", flush=True)
    assert err <= 1e-10, f"rel max err {err:.3e}"
    del got, ref
    energies = [F0]
    done = 3
    for target in range(100, 1001, 100):
        st.step(psi, target - done)
        done = target
        torch.cuda.synchronize()
        assert bool(torch.isfinite(psi).all().item()), f"non-finite field at step {done}"
        m = float(torch.mean(psi).item())
        F = st.free_energy(psi)
        print(f"step {done}: mean drift {m - mean0:.2e}, F {F:.10e}", flush=True)
        assert abs(m - mean0) <= 1e-12, f"mean drift {m - mean0:.3e} at step {done}"
        assert F <= energies[-1], f"free energy rose at step {done}: {energies[-1]} -> {F}"
        energies.append(F)
