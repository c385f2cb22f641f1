"""Adjoint of the plain PFC step on the device (pfc_hip_pfc_adjoint_step) and
the libtorch autograd::Function over trajectories (openpfc_amd_torch):

* one adjoint step (lambda and the eps term) against the oracle's adjoint,
  which the CPU tests pin by central differences, within 1e-10;
* gradients of J = ||psi_T - target||^2 with respect to psi0 (random
  directions) and eps against FP64 central finite differences of the
  device forward, within 1e-6 relative;
* the gradients do not depend on the checkpoint interval (the segments are
  recomputed bit for bit);
* the FP32-mixed mode (FP32 deviation from the exact FP64 mean): one adjoint
  step against the oracle's FP64 adjoint at psi = mean + d within 1e-5, and
  trajectory gradients within 1e-4 of FP64 central differences.
"""
import ctypes
import math

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

DX = math.pi / 4
DT = 0.5


def _fields(torch, n, seed):
    rng = np.random.default_rng(seed)
    shape = (n[2], n[1], n[0])
    psi0 = -0.2 + 0.1 * rng.uniform(-1, 1, shape)
    target = -0.2 + 0.1 * rng.uniform(-1, 1, shape)
    return psi0, target


def test_adjoint_step_matches_oracle(torch_gpu):
    import openpfc_amd as P
    torch = torch_gpu
    n = (32, 16, 16)
    eps = -0.3
    psi, lam = _fields(torch, n, 5)
    f = P.HipFft(*n)
    st = P.PfcStepper(f, (DX,) * 3, eps)
    st.initialize(DT)
    dpsi = torch.from_numpy(psi).cuda()
    dlam = torch.from_numpy(lam).cuda()
    dr = ctypes.c_double(0.0)
    P.check(f._lib.pfc_hip_pfc_adjoint_step(f.handle, ctypes.c_void_p(dpsi.data_ptr()),
                                            ctypes.c_void_p(dlam.data_ptr()),
                                            ctypes.byref(dr), None))
    opL, opN = O.pfc_operators(n, (DX,) * 3, eps, DT)
    dL, dN = O.pfc_operator_derivs(n, (DX,) * 3, eps, DT)
    ref, ref_dr = O.pfc_adjoint_step(psi, lam, opL, opN, dL, dN)
    got = dlam.cpu().numpy()
    assert np.abs(got - ref).max() <= 1e-10 * np.abs(ref).max()
    assert abs(dr.value - ref_dr) <= 1e-10 * abs(ref_dr), (dr.value, ref_dr)
    # without the eps term the field update is the same
    dlam2 = torch.from_numpy(lam).cuda()
    P.check(f._lib.pfc_hip_pfc_adjoint_step(f.handle, ctypes.c_void_p(dpsi.data_ptr()),
                                            ctypes.c_void_p(dlam2.data_ptr()), None, None))
    assert torch.equal(dlam2, dlam)


def _loss(torch, traj, psi0, eps, target, nsteps, every=4):
    e = torch.tensor(eps, dtype=torch.float64)
    psiT = traj(psi0, e, nsteps, DT, (DX,) * 3, every)
    return float(((psiT - target) ** 2).sum().item())


@pytest.mark.parametrize("n,nsteps", [((32, 32, 32), 10), ((64, 32, 16), 7)])
def test_trajectory_gradients_match_central_differences(torch_gpu, n, nsteps):
    from openpfc_amd.autograd import pfc_trajectory
    torch = torch_gpu
    psi0_np, target_np = _fields(torch, n, 7)
    target = torch.from_numpy(target_np).cuda()
    psi0 = torch.from_numpy(psi0_np).cuda().requires_grad_(True)
    eps = torch.tensor(-0.3, dtype=torch.float64, requires_grad=True)
    psiT = pfc_trajectory(psi0, eps, nsteps, DT, (DX,) * 3, every=4)
    J = ((psiT - target) ** 2).sum()
    J.backward()
    g, ge = psi0.grad.detach(), float(eps.grad.item())
    rng = np.random.default_rng(8)
    base = psi0.detach()
    for _ in range(2):
        v = torch.from_numpy(rng.standard_normal(psi0_np.shape)).cuda()
        h = 1e-6
        fd = (_loss(torch, pfc_trajectory, base + h * v, -0.3, target, nsteps) -
              _loss(torch, pfc_trajectory, base - h * v, -0.3, target, nsteps)) / (2 * h)
        ad = float((g * v).sum().item())
        assert abs(fd - ad) <= 1e-6 * abs(ad), (fd, ad)
    h = 1e-6
    fd = (_loss(torch, pfc_trajectory, base, -0.3 + h, target, nsteps) -
          _loss(torch, pfc_trajectory, base, -0.3 - h, target, nsteps)) / (2 * h)
    assert abs(fd - ge) <= 1e-6 * abs(ge), (fd, ge)


def test_gradients_independent_of_checkpoint_interval(torch_gpu):
    from openpfc_amd.autograd import pfc_trajectory
    torch = torch_gpu
    n = (32, 16, 16)
    psi0_np, target_np = _fields(torch, n, 9)
    target = torch.from_numpy(target_np).cuda()
    out = []
    for every in (1, 3, 10):
        psi0 = torch.from_numpy(psi0_np).cuda().requires_grad_(True)
        eps = torch.tensor(-0.25, dtype=torch.float64, requires_grad=True)
        J = ((pfc_trajectory(psi0, eps, 10, DT, (DX,) * 3, every) - target) ** 2).sum()
        J.backward()
        out.append((psi0.grad.clone(), float(eps.grad.item())))
    for g, ge in out[1:]:
        assert torch.equal(g, out[0][0]) and ge == out[0][1]


def test_mixed_adjoint_step_matches_fp64_oracle(torch_gpu):
    import openpfc_amd as P
    torch = torch_gpu
    n = (32, 16, 16)
    eps = -0.3
    psi, lam = _fields(torch, n, 15)
    m = float(psi.mean())
    f = P.HipFft(*n, precision=P.F32)
    st = P.PfcStepper(f, (DX,) * 3, eps)
    st.initialize(DT)
    st.set_mean(m)
    dpsi = torch.from_numpy(psi - m).float().cuda()
    dlam = torch.from_numpy(lam).float().cuda()
    dr = ctypes.c_double(0.0)
    P.check(f._lib.pfc_hip_pfc_adjoint_step(f.handle, ctypes.c_void_p(dpsi.data_ptr()),
                                            ctypes.c_void_p(dlam.data_ptr()),
                                            ctypes.byref(dr), None))
    opL, opN = O.pfc_operators(n, (DX,) * 3, eps, DT)
    dL, dN = O.pfc_operator_derivs(n, (DX,) * 3, eps, DT)
    ref, ref_dr = O.pfc_adjoint_step(psi, lam, opL, opN, dL, dN)
    got = dlam.double().cpu().numpy()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"mixed adjoint step: rel max {err:.2e}, dr rel {abs(dr.value - ref_dr) / abs(ref_dr):.2e}")
    assert err <= 1e-5
    assert abs(dr.value - ref_dr) <= 1e-5 * abs(ref_dr), (dr.value, ref_dr)


def test_mixed_trajectory_gradients_close_to_fp64_central_differences(torch_gpu):
    from openpfc_amd.autograd import pfc_trajectory
    torch = torch_gpu
    n, nsteps = (32, 32, 32), 10
    psi0_np, target_np = _fields(torch, n, 17)
    target = torch.from_numpy(target_np).cuda()
    psi0 = torch.from_numpy(psi0_np).cuda().requires_grad_(True)
    eps = torch.tensor(-0.3, dtype=torch.float64, requires_grad=True)
    psiT = pfc_trajectory(psi0, eps, nsteps, DT, (DX,) * 3, every=4, mixed=True)
    ref_T = pfc_trajectory(psi0.detach(), eps.detach(), nsteps, DT, (DX,) * 3, every=4)
    assert psiT.dtype == torch.float64
    fwd = float(((psiT.detach() - ref_T).norm() / ref_T.norm()).item())
    assert fwd <= 1e-5, fwd
    J = ((psiT - target) ** 2).sum()
    J.backward()
    g, ge = psi0.grad.detach(), float(eps.grad.item())
    rng = np.random.default_rng(18)
    base = psi0.detach()
    v = torch.from_numpy(rng.standard_normal(psi0_np.shape)).cuda()
    h = 1e-6
    fd = (_loss(torch, pfc_trajectory, base + h * v, -0.3, target, nsteps) -
          _loss(torch, pfc_trajectory, base - h * v, -0.3, target, nsteps)) / (2 * h)
    ad = float((g * v).sum().item())
    fdr = (_loss(torch, pfc_trajectory, base, -0.3 + h, target, nsteps) -
           _loss(torch, pfc_trajectory, base, -0.3 - h, target, nsteps)) / (2 * h)
    print(f"mixed gradients: psi0 direction rel {abs(ad - fd) / abs(fd):.2e}, "
          f"r rel {abs(ge - fdr) / abs(fdr):.2e}, forward rel L2 {fwd:.2e}")
    assert abs(fd - ad) <= 1e-4 * abs(fd), (fd, ad)
    assert abs(fdr - ge) <= 1e-4 * abs(fdr), (fdr, ge)


def test_binomial_schedule_gradients_equal_uniform(torch_gpu):
    """The binomial (revolve) schedule applies the same adjoint steps to the
    same recomputed states as the uniform one: gradients bit for bit."""
    from openpfc_amd.autograd import pfc_trajectory
    torch = torch_gpu
    n, nsteps = (32, 16, 16), 23
    psi0_np, target_np = _fields(torch, n, 19)
    target = torch.from_numpy(target_np).cuda()
    out = []
    for every, snaps in ((5, 0), (1, 0), (5, 1), (5, 3), (5, 6), (5, 40)):
        print(f"schedule every={every} snaps={snaps}", flush=True)
        psi0 = torch.from_numpy(psi0_np).cuda().requires_grad_(True)
        eps = torch.tensor(-0.25, dtype=torch.float64, requires_grad=True)
        psiT = pfc_trajectory(psi0, eps, nsteps, DT, (DX,) * 3, every=every, snaps=snaps)
        J = ((psiT - target) ** 2).sum()
        J.backward()
        out.append((psiT.detach().clone(), psi0.grad.clone(), float(eps.grad.item())))
    for T, g, ge in out[1:]:
        assert torch.equal(T, out[0][0])
        assert torch.equal(g, out[0][1]) and ge == out[0][2]


def test_trajectory_on_a_side_stream_matches_default_stream(torch_gpu):
    """Every library call of the autograd Function is enqueued on torch's
    current stream, so work under torch.cuda.stream(s) (a non-blocking pool
    stream) gives the default stream's results."""
    from openpfc_amd.autograd import pfc_trajectory
    torch = torch_gpu
    n, nsteps = (64, 32, 32), 12
    psi0_np, target_np = _fields(torch, n, 21)
    res = []
    for side in (False, True):
        print(f"trajectory side stream={side}", flush=True)
        s = torch.cuda.Stream() if side else torch.cuda.current_stream()
        with torch.cuda.stream(s):
            # inputs produced on this stream right before the call
            psi0 = torch.from_numpy(psi0_np).cuda().mul_(1.0).requires_grad_(True)
            target = torch.from_numpy(target_np).cuda().add_(0.0)
            eps = torch.tensor(-0.3, dtype=torch.float64, requires_grad=True)
            psiT = pfc_trajectory(psi0, eps, nsteps, DT, (DX,) * 3, every=5, snaps=2)
            J = ((psiT - target) ** 2).sum()
            J.backward()
            g = psi0.grad.clone()
        s.synchronize()
        res.append((g, float(eps.grad.item()), float(J.item())))
    assert torch.equal(res[0][0], res[1][0])
    assert res[0][1] == res[1][1] and res[0][2] == res[1][2]
