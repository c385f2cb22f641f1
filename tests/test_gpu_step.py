"""GPU parity of the fused plain-PFC ETD1 step against the CPU oracle.

The oracle restates docs/api/examples/10_complete_pfc_model.cpp:121-144
(step) and :166-213 (operators) with the reference's FFT contract.  The bar:
max|psi_gpu - psi_ref| / max|psi_ref| <= 1e-10 in FP64 for the same seed and
step count.  Invariants (reference
tests/integration/scenarios/complete_simulation/test_mass_conservation.cpp:47):
the mean is conserved to 1e-12, a constant field is a fixed point.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

DX = math.pi / 4
DT = 0.5
EPS = -0.25


def _psi0(n, seed=1234):
    return O.uniform_noise(n, seed, -0.01, 0.01, base=-0.2)


def _run_gpu(torch, n, psi0, chunks, prec="f64", eps=EPS, dx=DX, dt=DT):
    import openpfc_amd as P
    f = P.HipFft(*n, precision=P.F64 if prec == "f64" else P.F32)
    st = P.PfcStepper(f, (dx, dx, dx), eps)
    st.initialize(dt)
    rdt = torch.float64 if prec == "f64" else torch.float32
    psi = torch.from_numpy(psi0).to(rdt).cuda()
    out = []
    for c in chunks:
        st.step(psi, c)
        out.append(psi.double().cpu().numpy().copy())
    return out


def _rel_max(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("n,steps", [((16, 16, 16), 20), ((32, 16, 8), 30),
                                     ((1024, 8, 4), 5), ((8, 1024, 4), 5),
                                     ((8, 4, 1024), 5), ((256, 128, 128), 10),
                                     ((1024, 256, 128), 3), ((256, 1024, 128), 3),
                                     ((256, 128, 1024), 3)])
def test_step_matches_oracle_small(torch_gpu, n, steps):
    print(f"step n={n} steps={steps}", flush=True)
    psi0 = _psi0(n)
    opL, opN = O.pfc_operators(n, (DX, DX, DX), EPS, DT)
    ref = O.pfc_step(psi0, opL, opN, steps)
    got = _run_gpu(torch_gpu, n, psi0, [steps])[0]
    err = _rel_max(got, ref)
    assert err <= 1e-10, f"rel max err {err:.3e}"


def test_step_64cubed_200_steps_config1(torch_gpu):
    """64^3, dx = pi/4, dt = 0.5, r = -0.25, psi0 = -0.2 + U(-0.01, 0.01)
    (mt19937_64 seed 1234), 200 steps; compared at steps 50/100/150/200."""
    n = (64, 64, 64)
    print("step 64^3 x 200", flush=True)
    psi0 = _psi0(n)
    opL, opN = O.pfc_operators(n, (DX, DX, DX), EPS, DT)
    got = _run_gpu(torch_gpu, n, psi0, [50, 50, 50, 50])
    a = psi0
    for i in range(4):
        a = O.pfc_step(a, opL, opN, 50)
        err = _rel_max(got[i], a)
        print(f"step {50 * (i + 1)}: rel max err {err:.3e}")
        assert err <= 1e-10, f"step {50 * (i + 1)}: rel max err {err:.3e}"
    mean0 = psi0.mean()
    assert abs(got[-1].mean() - mean0) <= 1e-12 * abs(mean0)


def test_constant_field_is_fixed_point(torch_gpu):
    n = (32, 32, 32)
    psi0 = np.full((32, 32, 32), -0.3)
    got = _run_gpu(torch_gpu, n, psi0, [10])[0]
    assert np.abs(got - psi0).max() <= 1e-14


def test_step_f32_tracks_f64(torch_gpu):
    """FP32 storage and arithmetic stays near the FP64 oracle at 64^3 x 200
    (relative L2 error reported; bar 1e-3)."""
    n = (64, 64, 64)
    psi0 = _psi0(n)
    opL, opN = O.pfc_operators(n, (DX, DX, DX), EPS, DT)
    ref = O.pfc_step(psi0, opL, opN, 200)
    got = _run_gpu(torch_gpu, n, psi0, [200], prec="f32")[0]
    l2 = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    print(f"f32 vs f64 oracle rel L2 {l2:.3e} rel max {_rel_max(got, ref):.3e}")
    assert l2 <= 1e-3


def test_step_before_operators_is_state_error(torch_gpu):
    torch = torch_gpu
    import openpfc_amd as P
    f = P.HipFft(8, 8, 8)
    st = P.PfcStepper(f, (1.0, 1.0, 1.0), -0.25)
    psi = torch.zeros(512, dtype=torch.float64, device="cuda")
    with pytest.raises(P.PfcHipError) as ei:
        st.step(psi, 1)
    assert ei.value.code == 5


@pytest.mark.parametrize("n,steps", [((64, 64, 64), 1000), ((256, 128, 128), 200)])
def test_fp32_mixed_within_1e4_of_fp64(torch_gpu, n, steps):
    """FP32 field and transforms, FP64 conserved mean (mean-split step):
    relative L2 error <= 1e-4 against the FP64 oracle after the run, and the
    mean exactly conserved."""
    torch = torch_gpu
    import openpfc_amd as P
    psi0 = _psi0(n)
    opL, opN = O.pfc_operators(n, (DX, DX, DX), EPS, DT)
    ref = O.pfc_step(psi0, opL, opN, steps)
    mean = float(psi0.mean())
    f = P.HipFft(*n, precision=P.F32)
    st = P.PfcStepper(f, (DX, DX, DX), EPS)
    st.initialize(DT)
    st.set_mean(mean)
    d = torch.from_numpy(psi0 - mean).float().cuda()
    st.step(d, steps)
    got = mean + d.double().cpu().numpy()
    l2 = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    print(f"fp32-mixed {n} x {steps}: rel L2 {l2:.3e}, rel max {_rel_max(got, ref):.3e}, "
          f"mean drift {abs(got.mean() - mean):.2e}")
    assert l2 <= 1e-4
    assert abs(got.mean() - mean) <= 1e-9


def test_fp64_mean_split_matches_plain(torch_gpu):
    """In FP64 the mean-split step is the same step: the mean is restored
    exactly and the field agrees with the oracle to the FP64 bar."""
    torch = torch_gpu
    import openpfc_amd as P
    n = (256, 128, 128)
    psi0 = _psi0(n)
    opL, opN = O.pfc_operators(n, (DX, DX, DX), EPS, DT)
    ref = O.pfc_step(psi0, opL, opN, 20)
    mean = float(psi0.mean())
    f = P.HipFft(*n)
    st = P.PfcStepper(f, (DX, DX, DX), EPS)
    st.initialize(DT)
    st.set_mean(mean)
    d = torch.from_numpy(psi0 - mean).cuda()
    st.step(d, 20)
    got = mean + d.cpu().numpy()
    assert _rel_max(got, ref) <= 1e-10


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 2e-6)])
def test_fused_multi_step_matches_single_steps(torch_gpu, prec, tol):
    """A K-step call fuses every step boundary (inverse x of one step with the
    forward x of the next, psi kept in registers); K single-step calls do
    not.  The two agree to rounding, and the fused run's launch timing
    reports the boundary pass."""
    import openpfc_amd as P
    torch = torch_gpu
    n = (256, 128, 128)
    psi0 = _psi0(n)
    f = P.HipFft(*n, precision=P.F64 if prec == "f64" else P.F32)
    assert f.variants("step_layout")[f.get_variant("step_layout")] == "transposed"
    assert len(f.variants("t_x_fuse")) > 0
    f.set_variant("step_fuse", 1)
    st = P.PfcStepper(f, (DX, DX, DX), EPS)
    st.initialize(DT)
    rdt = torch.float64 if prec == "f64" else torch.float32
    a = torch.from_numpy(psi0).to(rdt).cuda()
    b = a.clone()
    st.step(a, 12)
    for _ in range(12):
        st.step(b, 1)
    err = float(((a - b).abs().max() / b.abs().max()).item())
    print(f"fused vs single steps ({prec}): rel max {err:.2e}")
    assert err <= tol
    opL, opN = O.pfc_operators(n, (DX, DX, DX), EPS, DT)
    ref = O.pfc_step(psi0, opL, opN, 12)
    assert _rel_max(a.double().cpu().numpy(), ref) <= (1e-10 if prec == "f64" else 1e-5)
    t = st.step_timed_fused(a, 4)
    assert len(t) == 6 and all(v > 0 for v in t), t
