"""GPU parity of the fused Tungsten mean-field step against the CPU oracle.

The oracle restates apps/tungsten/include/tungsten/cpu/tungsten_model.hpp:
223-366 and is pinned by the reference's golden norms (test_oracle_tungsten).
Here the HIP step (both spectral layouts) is checked against:
  * the golden norms themselves (apps/tungsten/tests/test_tungsten.cpp:
    288-299,347-354): 32^3, dt = 1, sum(psi^2) within 0.1 after each step;
  * the reference's CPU-vs-HIP recipe (test_tungsten_cpu_vs_hip.cpp:30-90):
    32^3, spacing 1, T = 0.5, dt = 0.01, psi = -0.4 + 0.1 sin(2 pi i/n),
    10 steps, every point within 1e-10 of the CPU result;
  * the oracle on larger and anisotropic grids, max|d| / max|psi| <= 1e-10;
  * the mean field psi_MF of the last step and the operator tables.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

DX_G = 1.1107207345395915
GOLDEN = [11965.0889218507, 11259.9705028338, 11050.6245088282, 10903.2783913748,
          10782.9029639299, 10678.0834708269, 10583.6114287826, 10496.5733807390,
          10415.1258360383, 10338.0182669812, 10264.3672697202]
GOLDEN_PARAMS = dict(n0=-0.10, n_sol=-0.047, n_vap=-0.464, T0=156000.0, T=3300.0,
                     Bx=0.8582, alpha_farTol=0.001, alpha_highOrd=4, alpha=0.50,
                     lambda_=0.22, stabP=0.2, shift_s=0.1898, shift_u=0.3341, p2=1.0,
                     p3=-0.5, p4=0.333333333, q20=-0.0037, q21=1.0, q30=-12.4567,
                     q31=20.0, q40=45.0)


def _params(**kw):
    from openpfc_amd import TungstenParams
    return TungstenParams(**{**GOLDEN_PARAMS, **kw})


def _oracle_ops(n, spacing, prm, dt):
    return O.tungsten_operators(n, spacing, prm.operator_params(), dt)


def _seed_field(n, dx):
    a = np.full((n[2], n[1], n[0]), -0.4)
    x0 = -0.5 * dx * n[0]
    O.ic_single_seed_masked(a, (x0, -0.5 * dx * n[1], -0.5 * dx * n[2]), (dx,) * 3,
                            -0.047, 0.215936, 18.0, 1)
    return a


def _stepper(torch, n, spacing, prm, dt, prec="f64", layout=None):
    import openpfc_amd as P
    f = P.HipFft(*n, precision=P.F64 if prec == "f64" else P.F32)
    if layout is not None:
        names = f.variants("step_layout")
        if layout not in names:
            pytest.skip(f"{layout} layout not available for {n}")
        f.set_variant("step_layout", names.index(layout))
    st = P.TungstenStepper(f, spacing, prm)
    st.initialize(dt)
    return f, st


def _rel_max(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_tungsten_golden_norms_gpu(torch_gpu):
    torch = torch_gpu
    n = (32, 32, 32)
    prm = _params()
    _, st = _stepper(torch, n, (DX_G,) * 3, prm, 1.0)
    psi0 = _seed_field(n, DX_G)
    fMF, expL, nW = _oracle_ops(n, (DX_G,) * 3, prm, 1.0)
    psi = torch.from_numpy(psi0).cuda()
    ref = psi0.copy()
    norms = [float((psi0 * psi0).sum())]
    for i in range(10):
        st.step(psi, 1)
        ref, _ = O.tungsten_step(ref, fMF, expL, nW, prm.nonlinear_params(), 1)
        got = psi.cpu().numpy()
        norms.append(float((got * got).sum()))
        err = _rel_max(got, ref)
        assert err <= 1e-10, f"step {i + 1}: rel max err {err:.3e}"
    for i, (g, r) in enumerate(zip(norms, GOLDEN)):
        assert abs(g - r) <= 0.1, (i, g, r)


def test_tungsten_cpu_vs_hip_recipe(torch_gpu):
    """test_tungsten_cpu_vs_hip.cpp:30-90 with the oracle as the CPU model."""
    torch = torch_gpu
    n = (32, 32, 32)
    prm = _params(n0=-0.4, T=0.5)
    _, st = _stepper(torch, n, (1.0, 1.0, 1.0), prm, 0.01)
    N = n[0] * n[1] * n[2]
    psi0 = (-0.4 + 0.1 * np.sin(2.0 * math.pi * np.arange(N) / N)).reshape(32, 32, 32)
    fMF, expL, nW = _oracle_ops(n, (1.0,) * 3, prm, 0.01)
    ref, _ = O.tungsten_step(psi0, fMF, expL, nW, prm.nonlinear_params(), 10)
    psi = torch.from_numpy(psi0).cuda()
    st.step(psi, 10)
    d = np.abs(psi.cpu().numpy() - ref)
    assert int((d > 1e-10).sum()) == 0 and float(d.max()) < 1e-10, float(d.max())


@pytest.mark.parametrize("n,steps,layout", [
    ((64, 64, 64), 20, "natural"),
    ((32, 16, 8), 10, "natural"),
    ((256, 128, 128), 5, "natural"),
    ((256, 128, 128), 5, "transposed"),
    ((512, 256, 128), 3, "transposed"),
    ((256, 512, 128), 3, "transposed"),
    ((256, 128, 512), 3, "transposed"),
    ((1024, 128, 128), 2, "transposed"),
])
def test_tungsten_matches_oracle(torch_gpu, n, steps, layout):
    print(f"tungsten n={n} steps={steps} layout={layout}", flush=True)
    torch = torch_gpu
    prm = _params()
    _, st = _stepper(torch, n, (DX_G,) * 3, prm, 1.0, layout=layout)
    psi0 = O.uniform_noise(n, 77, -0.05, 0.05, base=-0.1)
    fMF, expL, nW = _oracle_ops(n, (DX_G,) * 3, prm, 1.0)
    ref, ref_mf = O.tungsten_step(psi0, fMF, expL, nW, prm.nonlinear_params(), steps)
    psi = torch.from_numpy(psi0).cuda()
    mf = torch.empty_like(psi)
    st.step(psi, steps, psi_mf=mf)
    err = _rel_max(psi.cpu().numpy(), ref)
    assert err <= 1e-10, f"psi rel max err {err:.3e}"
    err_mf = _rel_max(mf.cpu().numpy(), ref_mf)
    assert err_mf <= 1e-10, f"psi_MF rel max err {err_mf:.3e}"
    # mass conservation (exp(0) = 1, n_weight(0) = 0)
    assert abs(float(psi.mean()) - float(psi0.mean())) <= 1e-12


@pytest.mark.parametrize("layout", ["natural", "transposed"])
def test_tungsten_fp32(torch_gpu, layout):
    """TungstenHIP<float>: FP32 throughout, coefficients rounded from FP64."""
    torch = torch_gpu
    n = (256, 128, 128)
    prm = _params()
    _, st = _stepper(torch, n, (DX_G,) * 3, prm, 1.0, prec="f32", layout=layout)
    psi0 = O.uniform_noise(n, 78, -0.05, 0.05, base=-0.1)
    fMF, expL, nW = _oracle_ops(n, (DX_G,) * 3, prm, 1.0)
    ref, _ = O.tungsten_step(psi0, fMF, expL, nW, prm.nonlinear_params(), 5)
    psi = torch.from_numpy(psi0).float().cuda()
    st.step(psi, 5)
    got = psi.double().cpu().numpy()
    rel_l2 = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    assert rel_l2 <= 1e-5, rel_l2


def test_tungsten_operator_tables(torch_gpu):
    """Device operators vs the oracle tables: filterMF and exp(L dt) within a
    few ulps, n_weight within the cancellation bound of expm1(x)/L."""
    torch = torch_gpu
    n = (32, 16, 8)
    prm = _params()
    _, st = _stepper(torch, n, (DX_G, 0.9, 1.3), prm, 0.7)
    f, e, w = (t.cpu().numpy() * (n[0] * n[1] * n[2]) for t in st.operators())
    fMF, expL, nW = _oracle_ops(n, (DX_G, 0.9, 1.3), prm, 0.7)
    eps = np.finfo(np.float64).eps
    # fMF: exp of identically rounded arguments (1 ulp each side).  expL and
    # n_weight: a 1-ulp difference of fMF moves L*dt by |kLap*dt*q2*fMF|*eps
    # (< 2 eps here) and exp/expm1 add 1 ulp each side.
    assert np.all(np.abs(f - fMF) <= 4 * eps * np.abs(fMF))
    assert np.all(np.abs(e - expL) <= 16 * eps * np.abs(expL) + 1e-300)
    assert np.all(np.abs(w - nW) <= 16 * eps * np.abs(nW) + 1e-300)
    assert e[0, 0, 0] == 1.0 and w[0, 0, 0] == 0.0


def test_tungsten_variants(torch_gpu):
    """Every variant of the Tungsten-specific passes gives the same step."""
    torch = torch_gpu
    n = (256, 128, 128)
    prm = _params()
    f, st = _stepper(torch, n, (DX_G,) * 3, prm, 1.0)
    psi0 = O.uniform_noise(n, 79, -0.05, 0.05, base=-0.1)
    fMF, expL, nW = _oracle_ops(n, (DX_G,) * 3, prm, 1.0)
    ref, _ = O.tungsten_step(psi0, fMF, expL, nW, prm.nonlinear_params(), 2)
    layouts = f.variants("step_layout")
    for layout, roles in (("natural", ["x_tung", "pfc_z"]),
                          ("transposed", ["t_x_r2c", "t_x_tung", "t_z"])):
        f.set_variant("step_layout", layouts.index(layout))
        for role in roles:
            names = f.variants(role)
            for i, name in enumerate(names):
                print(f"variant {layout} {role} {name}", flush=True)
                f.set_variant(role, i)
                psi = torch.from_numpy(psi0).cuda()
                st.step(psi, 2)
                err = _rel_max(psi.cpu().numpy(), ref)
                assert err <= 1e-10, f"{role} {name}: {err:.3e}"
            f.set_variant(role, 0)


def test_tungsten_step_timed_and_autotune(torch_gpu):
    torch = torch_gpu
    n = (256, 128, 128)
    _, st = _stepper(torch, n, (DX_G,) * 3, _params(), 1.0)
    psi = torch.from_numpy(O.uniform_noise(n, 80, -0.05, 0.05, base=-0.1)).cuda()
    ms = st.step_timed(psi, 2)
    assert len(ms) == 9 and all(m > 0 for m in ms)
    before = psi.clone()
    res = st.autotune(psi, reps=1)
    assert torch.equal(psi, before)
    assert set(res["chosen"]) >= {"x_r2c", "x_tung", "t_x_r2c", "t_x_tung", "step_layout"}
    assert res["layout_ms"]["natural"] > 0 and res["layout_ms"]["transposed"] > 0


def test_tungsten_errors(torch_gpu):
    import ctypes
    import openpfc_amd as P
    torch = torch_gpu
    f = P.HipFft(16, 16, 16)
    lib = f._lib
    bad = (ctypes.c_double * 16)(*([1.0] * 16))
    assert lib.pfc_hip_pfc_set_operators(f.handle, P.MODEL_TUNGSTEN, bad, 16, 1.0) == 1
    st = P.TungstenStepper(f, (1.0, 1.0, 1.0), _params())
    psi = torch.zeros(16 ** 3, dtype=torch.float64, device="cuda")
    with pytest.raises(P.PfcHipError):
        st.step(psi, 1)  # before initialize
    st.initialize(0.5)
    with pytest.raises(P.PfcHipError):
        P.PfcStepper(f, (1.0, 1.0, 1.0), -0.25).set_mean(-0.1)  # mean split: PFC only
    with pytest.raises(P.PfcHipError):
        st.step(psi, 1, psi_mf=psi)  # aliasing
    with pytest.raises(ValueError):
        st.step(psi[:10], 1)
    # pfc operators are not available on a Tungsten plan
    assert lib.pfc_hip_pfc_operators(f.handle, ctypes.c_void_p(psi.data_ptr()),
                                     ctypes.c_void_p(psi.data_ptr()), None) == 5
