"""Every launch variant (radix plan, tile width, rows per block, exchange
mode) of every pass gives the oracle's answer: the autotuner may pick any."""
import math

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

DX, DT, EPS = math.pi / 4, 0.5, -0.25
GRIDS = [(256, 128, 512), (1024, 4, 8), (8, 1024, 4), (8, 4, 1024), (512, 256, 4)]
# grids where the transposed step layout is available (nx/2, ny, nz >= 128)
T_GRIDS = [(256, 128, 128), (1024, 128, 256), (256, 512, 128), (256, 128, 1024)]


@pytest.mark.parametrize("n", GRIDS)
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_all_transform_variants(torch_gpu, n, prec):
    torch = torch_gpu
    import openpfc_amd as P
    nx, ny, nz = n
    f = P.HipFft(nx, ny, nz, P.F64 if prec == "f64" else P.F32)
    rdt = torch.float64 if prec == "f64" else torch.float32
    cdt = torch.complex128 if prec == "f64" else torch.complex64
    tol = 1e-13 if prec == "f64" else 3e-6
    rng = np.random.default_rng(7)
    a = rng.standard_normal((nz, ny, nx))
    ref = O.r2c(a)
    back = O.c2r(ref, nx)
    x = torch.from_numpy(a).to(rdt).cuda()
    Xin = torch.from_numpy(ref).to(cdt).cuda()
    for role in ["x_r2c", "x_c2r", "col_y", "col_z"]:
        names = f.variants(role)
        for i, name in enumerate(names):
            print(f"{n} {prec} {role}[{i}] {name}", flush=True)
            f.set_variant(role, i)
            X = torch.empty((nz, ny, nx // 2 + 1), dtype=cdt, device="cuda")
            f.forward(x, X)
            e1 = np.abs(X.cpu().numpy() - ref).max() / np.abs(ref).max()
            y = torch.empty_like(x)
            f.backward(Xin, y)
            e2 = np.abs(y.double().cpu().numpy() - back).max() / np.abs(back).max()
            assert e1 <= tol and e2 <= tol, f"{role} {name}: r2c {e1:.2e} c2r {e2:.2e}"
        f.set_variant(role, 0)


@pytest.mark.parametrize("n", GRIDS)
def test_all_step_variants(torch_gpu, n):
    torch = torch_gpu
    import openpfc_amd as P
    nx, ny, nz = n
    psi0 = O.uniform_noise(n, 99, -0.01, 0.01, base=-0.2)
    opL, opN = O.pfc_operators(n, (DX, DX, DX), EPS, DT)
    ref = O.pfc_step(psi0, opL, opN, 3)
    f = P.HipFft(nx, ny, nz)
    st = P.PfcStepper(f, (DX, DX, DX), EPS)
    st.initialize(DT)
    for role in ["pfc_x", "col_y", "pfc_z", "x_c2r"]:
        for i, name in enumerate(f.variants(role)):
            print(f"{n} step {role}[{i}] {name}", flush=True)
            f.set_variant(role, i)
            psi = torch.from_numpy(psi0).cuda()
            st.step(psi, 3)
            err = np.abs(psi.cpu().numpy() - ref).max() / np.abs(ref).max()
            assert err <= 1e-12, f"{role} {name}: {err:.2e}"
        f.set_variant(role, 0)


@pytest.mark.parametrize("n", T_GRIDS)
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_all_transposed_step_variants(torch_gpu, n, prec):
    """Both step layouts and every variant of the transposed passes."""
    torch = torch_gpu
    import openpfc_amd as P
    psi0 = O.uniform_noise(n, 41, -0.01, 0.01, base=-0.2)
    opL, opN = O.pfc_operators(n, (DX, DX, DX), EPS, DT)
    ref = O.pfc_step(psi0, opL, opN, 2)
    f = P.HipFft(*n, precision=P.F64 if prec == "f64" else P.F32)
    assert f.variants("step_layout") == ["transposed", "natural"]
    st = P.PfcStepper(f, (DX, DX, DX), EPS)
    st.initialize(DT)
    rdt = torch.float64 if prec == "f64" else torch.float32
    tol = 1e-12 if prec == "f64" else 2e-5
    # per-step runs, then fused step boundaries (st.step(psi, 2) runs the
    # fused pass, role t_x_fuse, between the two steps) with every variant
    assert f.variants("step_fuse") == ["per_step", "fused"]
    cases = [("step_layout", 1, 0)] + [(r, i, fu) for fu in (0, 1)
                                       for r in ["t_x_fwd", "t_y_fwd", "t_z", "t_y_inv",
                                                 "t_x_inv"] + (["t_x_fuse"] if fu else [])
                                       for i in range(len(f.variants(r)))]
    for role, i, fu in cases:
        name = f.variants(role)[i]
        print(f"{n} {prec} fuse={fu} {role}[{i}] {name}", flush=True)
        f.set_variant("step_fuse", fu)
        f.set_variant(role, i)
        psi = torch.from_numpy(psi0).to(rdt).cuda()
        st.step(psi, 2)
        err = np.abs(psi.double().cpu().numpy() - ref).max() / np.abs(ref).max()
        assert err <= tol, f"{role} {name} fuse={fu}: {err:.2e}"
        f.set_variant(role, 0)
    f.set_variant("step_fuse", 0)


def test_operators_match_oracle(torch_gpu):
    """On-the-fly opL/opN equal the reference formula (oracle, glibc exp) to
    a few ulp, scaled by 1/N^3."""
    import openpfc_amd as P
    n = (64, 32, 16)
    f = P.HipFft(*n)
    st = P.PfcStepper(f, (DX, 1.1, 0.9), EPS)
    st.initialize(DT)
    gL, gN = (t.cpu().numpy() for t in st.operators())
    rL, rN = O.pfc_operators(n, (DX, 1.1, 0.9), EPS, DT)
    s = 1.0 / (64 * 32 * 16)  # a power of two: scaling is exact
    # ulps of the reference value, floored at 1e-290: below that opL is deep in
    # exp's underflow tail (L*dt < -650), where device and glibc exp may round
    # denormal results differently; such values cannot reach a double result.
    # opL: exp may differ from glibc by 1 ulp.  opN = (opL-1)/L*(-k2) inherits
    # that ulp amplified by the cancellation in opL-1, i.e. relative error
    # eps*opL/|1-opL| on top of a few ulp of its own arithmetic.
    eps = np.finfo(np.float64).eps
    tolN = 4 * np.spacing(np.abs(rN * s)) + 2 * eps * np.abs(rN * s) * rL / np.maximum(
        np.abs(1 - rL), 1e-300)
    for g, r, tol in ((gL, rL * s, 2 * np.spacing(np.maximum(np.abs(rL * s), 1e-290))),
                      (gN, rN * s, tolN)):
        ratio = np.abs(g - r) / np.maximum(tol, 1e-300)
        i = np.unravel_index(np.argmax(ratio), ratio.shape)
        print(f"worst |diff|/tol {ratio.max():.2f} at {i} (ref {r[i]:.6e}), "
              f"exact fraction {(g == r).mean():.3f}")
        assert ratio.max() <= 1.0


def test_autotune_restores_field_and_picks_valid(torch_gpu):
    torch = torch_gpu
    import openpfc_amd as P
    n = (256, 128, 128)
    f = P.HipFft(*n)
    st = P.PfcStepper(f, (DX, DX, DX), EPS)
    st.initialize(DT)
    psi0 = O.uniform_noise(n, 5, -0.01, 0.01, base=-0.2)
    psi = torch.from_numpy(psi0).cuda()
    res = st.autotune(psi, reps=1)
    assert np.array_equal(psi.cpu().numpy(), psi0)
    for role in ["pfc_x", "col_y", "pfc_z", "x_c2r", "t_x_fwd", "t_y_fwd", "t_z",
                 "t_y_inv", "t_x_inv"]:
        assert res["chosen"][role] in f.variants(role)
        assert all(ms > 0 for _, ms in res[role])
    assert res["layout_ms"]["transposed"] > 0 and res["layout_ms"]["natural"] > 0
    assert res["chosen"]["step_layout"] in ("transposed", "natural")
    opL, opN = O.pfc_operators(n, (DX, DX, DX), EPS, DT)
    st.step(psi, 2)
    ref = O.pfc_step(psi0, opL, opN, 2)
    assert np.abs(psi.cpu().numpy() - ref).max() / np.abs(ref).max() <= 1e-12
