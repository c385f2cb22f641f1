"""FP32 fields with FP64 butterflies (PFC_HIP_FLAG_F64_BUTTERFLIES).

* every launch variant of the five transposed passes against the FP64
  oracle after two steps, within 4e-6 relative (single precision storage,
  double arithmetic in the transforms);
* the mean-split step over 200 steps at 256x128x128 against the oracle:
  the FP64-butterfly error is at most a quarter of the FP32-butterfly error
  of the same step, and the mean is conserved;
* the flag is refused without PFC_HIP_F32, on grids without the transposed
  step, and for the Tungsten model; the plan offers the transposed layout
  only and no fused step boundary.
"""
import ctypes
import math

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

DX = math.pi / 4
DT = 0.5
EPS = -0.25
GRIDS = [(256, 128, 128), (1024, 128, 256), (256, 512, 128), (256, 128, 1024)]


@pytest.mark.parametrize("n", GRIDS)
def test_all_xp_variants_match_oracle(torch_gpu, n):
    import openpfc_amd as P
    torch = torch_gpu
    psi0 = O.uniform_noise(n, 43, -0.01, 0.01, base=-0.2)
    opL, opN = O.pfc_operators(n, (DX, DX, DX), EPS, DT)
    ref = O.pfc_step(psi0, opL, opN, 2)
    f = P.HipFft(*n, precision=P.F32, flags=P.FLAG_F64_BUTTERFLIES)
    assert f.variants("step_layout") == ["transposed"]
    assert f.variants("step_fuse") == ["per_step"]
    st = P.PfcStepper(f, (DX, DX, DX), EPS)
    st.initialize(DT)
    for role in ["t_x_fwd", "t_y_fwd", "t_z", "t_y_inv", "t_x_inv"]:
        names = f.variants(role)
        assert names and all("/xp" in v for v in names), (role, names)
        for i, name in enumerate(names):
            print(f"{n} xp {role}[{i}] {name}", flush=True)
            f.set_variant(role, i)
            psi = torch.from_numpy(psi0).float().cuda()
            st.step(psi, 2)
            err = np.abs(psi.double().cpu().numpy() - ref).max() / np.abs(ref).max()
            assert err <= 4e-6, f"{role} {name}: {err:.2e}"
        f.set_variant(role, 0)


def test_xp_mean_split_error_below_fp32_butterflies(torch_gpu):
    import openpfc_amd as P
    torch = torch_gpu
    n, steps = (256, 128, 128), 200
    psi0 = O.uniform_noise(n, 47, -0.01, 0.01, base=-0.2)
    opL, opN = O.pfc_operators(n, (DX, DX, DX), EPS, DT)
    ref = O.pfc_step(psi0, opL, opN, steps)
    mean = float(np.mean(psi0))
    err = {}
    for flags in (0, P.FLAG_F64_BUTTERFLIES):
        f = P.HipFft(*n, precision=P.F32, flags=flags)
        st = P.PfcStepper(f, (DX, DX, DX), EPS)
        st.initialize(DT)
        st.set_mean(mean)
        d = torch.from_numpy(psi0 - mean).float().cuda()
        st.step(d, steps)
        got = d.double().cpu().numpy() + mean
        err[flags] = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
        drift = abs(float(d.double().mean().item()))
        print(f"flags {flags}: rel L2 {err[flags]:.3e}, mean drift {drift:.1e}", flush=True)
        assert drift <= 1e-9
    assert err[P.FLAG_F64_BUTTERFLIES] <= 0.25 * err[0], err


def test_xp_flag_refusals(torch_gpu):
    import openpfc_amd as P
    lib = P.load()
    h = ctypes.c_void_p()
    assert lib.pfc_hip_plan_create(256, 128, 128, P.F64, P.FLAG_F64_BUTTERFLIES,
                                   ctypes.byref(h)) == 1
    assert lib.pfc_hip_plan_create(64, 64, 64, P.F32, P.FLAG_F64_BUTTERFLIES,
                                   ctypes.byref(h)) == 2
    assert lib.pfc_hip_plan_create(256, 128, 128, P.F32, 4, ctypes.byref(h)) == 1
    f = P.HipFft(256, 128, 128, precision=P.F32, flags=P.FLAG_F64_BUTTERFLIES)
    params = (ctypes.c_double * 17)(1.0, 1.0, 1.0, *([1.0] * 14))
    assert lib.pfc_hip_pfc_set_operators(f.handle, P.MODEL_TUNGSTEN, params, 17, 0.01) == 2
