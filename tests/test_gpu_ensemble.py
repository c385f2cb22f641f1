"""Mixed-radix lengths, batched plans and ensembles on the device.

* The 3-D R2C at the mixed-radix lengths 24, 48, 96 and 192 against numpy's
  rfftn (pocketfft), FP64 and FP32.
* The fused PFC step at those lengths against the oracle
  (docs/api/examples/10_complete_pfc_model.cpp:121-144), max rel <= 1e-10.
* A batched plan of B fields gives, field by field, the bits of B single
  plans with the same eps / mean (plain FP64 and mean-split FP32; on
  extents with the transposed step in both layouts, with fused step
  boundaries and with FP64 butterflies).
* Ensembles of uneven sizes (one graph launch per step): every run matches
  the oracle with its own eps; graph replays equal the eagerly enqueued
  step bit for bit, with eager work between replays; per-run eps changed
  between replays is picked up; autotuning keeps the fields; FP32-mixed runs
  stay within FP32 error of the FP64 oracle with the mean conserved.
"""
import ctypes
import math

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

DX = math.pi / 4
DT = 0.5


def _rel_max(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("n", [(48, 48, 48), (24, 96, 192), (192, 24, 48), (96, 192, 24)])
def test_mixed_radix_r2c_c2r_match_numpy(torch_gpu, n, prec):
    import openpfc_amd as P
    torch = torch_gpu
    print(f"mixed-radix fft n={n} {prec}", flush=True)
    nx, ny, nz = n
    f = P.HipFft(nx, ny, nz, precision=P.F64 if prec == "f64" else P.F32)
    a = np.random.default_rng(nx + 3 * ny + 7 * nz).standard_normal((nz, ny, nx))
    rdt, cdt = (torch.float64, torch.complex128) if prec == "f64" else (torch.float32,
                                                                        torch.complex64)
    x = torch.from_numpy(a).to(rdt).cuda()
    X = torch.empty((nz, ny, nx // 2 + 1), dtype=cdt, device="cuda")
    f.forward(x, X)
    ref = np.fft.rfftn(a)
    tol = 1e-13 if prec == "f64" else 2e-6
    got = X.cpu().numpy()
    assert np.abs(got - ref).max() <= tol * np.abs(ref).max() * math.log2(nx * ny * nz)
    back = torch.empty_like(x)
    f.backward(X, back)
    assert np.abs(back.double().cpu().numpy() - a).max() <= 10 * tol * np.abs(a).max() * \
        math.log2(nx * ny * nz)


@pytest.mark.parametrize("n,steps", [((48, 48, 48), 20), ((24, 96, 192), 5),
                                     ((192, 48, 96), 5), ((96, 96, 96), 10)])
def test_mixed_radix_step_matches_oracle(torch_gpu, n, steps):
    import openpfc_amd as P
    torch = torch_gpu
    print(f"mixed-radix step n={n} steps={steps}", flush=True)
    psi0 = O.uniform_noise(n, 99, -0.01, 0.01, base=-0.2)
    opL, opN = O.pfc_operators(n, (DX, DX, DX), -0.25, DT)
    ref = O.pfc_step(psi0, opL, opN, steps)
    f = P.HipFft(*n)
    st = P.PfcStepper(f, (DX, DX, DX), -0.25)
    st.initialize(DT)
    psi = torch.from_numpy(psi0).cuda()
    st.step(psi, steps)
    err = _rel_max(psi.cpu().numpy(), ref)
    assert err <= 1e-10, f"rel max err {err:.3e}"


def _batched(lib, n, batch, prec, eps, mean, split):
    import openpfc_amd as P
    h = ctypes.c_void_p()
    P.check(lib.pfc_hip_plan_create_batched(*n, batch, prec, 0, ctypes.byref(h)))
    op = (ctypes.c_double * 4)(DX, DX, DX, eps[0])
    P.check(lib.pfc_hip_pfc_set_operators(h, 0, op, 4, DT))
    e = (ctypes.c_double * batch)(*eps)
    m = (ctypes.c_double * batch)(*mean)
    P.check(lib.pfc_hip_pfc_set_members(h, e, m, int(split)))
    return h


@pytest.mark.parametrize("prec,split", [("f64", False), ("f32", True)])
def test_batched_plan_equals_single_plans(torch_gpu, prec, split):
    import openpfc_amd as P
    torch = torch_gpu
    lib = P.load()
    n = (32, 48, 24)
    print(f"batched plan n={n} {prec} split={split}", flush=True)
    eps = [-0.20, -0.31, -0.25]
    mean = [-0.18, -0.22, -0.27]
    pr = P.F64 if prec == "f64" else P.F32
    rdt = torch.float64 if prec == "f64" else torch.float32
    fields = [O.uniform_noise(n, 7 + b, -0.01, 0.01, base=0.0 if split else mean[b])
              for b in range(3)]
    h = _batched(lib, n, 3, pr, eps, mean, split)
    try:
        stacked = torch.from_numpy(np.stack(fields)).to(rdt).cuda().contiguous()
        s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        P.check(lib.pfc_hip_pfc_step(h, ctypes.c_void_p(stacked.data_ptr()), 6, s))
        for b in range(3):
            f = P.HipFft(*n, precision=pr)
            st = P.PfcStepper(f, (DX, DX, DX), eps[b])
            st.initialize(DT)
            if split:
                st.set_mean(mean[b])
            one = torch.from_numpy(fields[b]).to(rdt).cuda()
            st.step(one, 6)
            assert torch.equal(one, stacked[b]), f"field {b} differs"
    finally:
        lib.pfc_hip_plan_destroy(h)


@pytest.mark.parametrize("prec,split,flags", [("f64", False, 0), ("f32", True, 0),
                                               ("f32", True, 1)])
@pytest.mark.parametrize("layout,fuse", [(0, 0), (0, 1), (1, 0)])
def test_batched_transposed_equals_single_plans(torch_gpu, prec, split, flags, layout, fuse):
    """Extents with the transposed step: a batched plan steps its fields one
    after another there (layout 0, with or without fused step boundaries,
    FP64, FP32 and FP32 with FP64 butterflies) and all at once in the natural
    layout (1); either way each field gets the bits of a single plan with the
    same eps, mean and variants."""
    import openpfc_amd as P
    torch = torch_gpu
    lib = P.load()
    if flags and (layout == 1 or fuse):
        pytest.skip("plans with FP64 butterflies have the transposed layout and single steps "
                    "only")
    n = (256, 128, 128)
    print(f"batched transposed n={n} {prec} split={split} flags={flags} layout={layout} "
          f"fuse={fuse}", flush=True)
    eps = [-0.20, -0.31, -0.25]
    mean = [-0.18, -0.22, -0.27]
    pr = P.F64 if prec == "f64" else P.F32
    rdt = torch.float64 if prec == "f64" else torch.float32
    fields = [O.uniform_noise(n, 17 + b, -0.01, 0.01, base=0.0 if split else mean[b])
              for b in range(3)]
    h = ctypes.c_void_p()
    P.check(lib.pfc_hip_plan_create_batched(*n, 3, pr, flags, ctypes.byref(h)))
    try:
        op = (ctypes.c_double * 4)(DX, DX, DX, eps[0])
        P.check(lib.pfc_hip_pfc_set_operators(h, 0, op, 4, DT))
        P.check(lib.pfc_hip_pfc_set_members(h, (ctypes.c_double * 3)(*eps),
                                            (ctypes.c_double * 3)(*mean), int(split)))
        cnt = ctypes.c_int()
        P.check(lib.pfc_hip_variant_count(h, 6, ctypes.byref(cnt)))
        assert cnt.value == (1 if flags else 2)
        if not flags:
            P.check(lib.pfc_hip_variant_set(h, 6, layout))
        if fuse:
            P.check(lib.pfc_hip_variant_set(h, 16, 1))
        stacked = torch.from_numpy(np.stack(fields)).to(rdt).cuda().contiguous()
        s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        P.check(lib.pfc_hip_pfc_step(h, ctypes.c_void_p(stacked.data_ptr()), 4, s))
        for b in range(3):
            f = P.HipFft(*n, precision=pr, flags=flags)
            st = P.PfcStepper(f, (DX, DX, DX), eps[b])
            st.initialize(DT)
            if split:
                st.set_mean(mean[b])
            if not flags:
                f.set_variant("step_layout", layout)
            if fuse:
                f.set_variant("step_fuse", 1)
            one = torch.from_numpy(fields[b]).to(rdt).cuda()
            st.step(one, 4)
            assert torch.equal(one, stacked[b]), f"field {b} differs"
    finally:
        lib.pfc_hip_plan_destroy(h)


def test_batched_plan_rejects_single_field_calls(torch_gpu):
    import openpfc_amd as P
    lib = P.load()
    h = _batched(lib, (16, 16, 16), 2, P.F64, [-0.25, -0.2], [0.0, 0.0], False)
    try:
        d = torch_gpu.zeros(2 * 16 ** 3, dtype=torch_gpu.float64, device="cuda")
        p = ctypes.c_void_p(d.data_ptr())
        assert lib.pfc_hip_r2c(h, p, p, None) == P._lib.PFC_HIP_EUNSUPPORTED
        assert lib.pfc_hip_ic_constant(h, p, 1.0, None) == P._lib.PFC_HIP_EUNSUPPORTED
        assert lib.pfc_hip_pfc_set_mean(h, 1, 0.1) == P._lib.PFC_HIP_EUNSUPPORTED
        assert b"batched" in lib.pfc_hip_last_error()
        lam = ctypes.c_double()
        assert lib.pfc_hip_pfc_adjoint_step(h, p, ctypes.c_void_p(d.data_ptr() + 8), None,
                                            None) == P._lib.PFC_HIP_EUNSUPPORTED
    finally:
        lib.pfc_hip_plan_destroy(h)


SIZES = [16, 24, 32, 48, 24, 16, 96, 32]


def _runs(seed, sizes=SIZES, dx=DX, dt=DT):
    from openpfc_amd import Run
    rng = np.random.default_rng(seed)
    return [Run.cubic(n, dx, rng.uniform(-0.35, -0.15), dt, rng.uniform(-0.30, -0.15))
            for n in sizes]


def _initial(runs, seed):
    return [O.uniform_noise(r.ext, seed + i, -0.01, 0.01, base=r.mean) for i, r in
            enumerate(runs)]


def _load(ens, psi0, split):
    import torch
    for i, a in enumerate(psi0):
        v = a - ens.runs[i].mean if split else a
        ens.field(i).copy_(torch.from_numpy(v))


def _oracle(run, psi0, steps):
    opL, opN = O.pfc_operators(run.ext, run.spacing, run.eps, run.dt)
    return O.pfc_step(psi0, opL, opN, steps)


def test_ensemble_fp64_matches_oracle(torch_gpu):
    from openpfc_amd import Ensemble
    runs = _runs(11)
    psi0 = _initial(runs, 100)
    ens = Ensemble(runs)
    print(f"ensemble fp64 sizes={SIZES} groups={len(ens.groups())}", flush=True)
    assert len(ens.groups()) == 5
    _load(ens, psi0, False)
    ens.step(10)
    torch_gpu.cuda.synchronize()
    for i, r in enumerate(runs):
        ref = _oracle(r, psi0[i], 10)
        err = _rel_max(ens.field(i).cpu().numpy(), ref)
        assert err <= 1e-10, f"run {i} (n={r.ext[0]}): rel max err {err:.3e}"
    ens.close()


def test_ensemble_graph_replays_equal_eager(torch_gpu):
    from openpfc_amd import Ensemble
    torch = torch_gpu
    runs = _runs(12)
    psi0 = _initial(runs, 200)
    g, e = Ensemble(runs), Ensemble(runs)
    _load(g, psi0, False)
    _load(e, psi0, False)
    side = torch.randn(1 << 20, dtype=torch.float64, device="cuda")
    for k in range(4):
        print(f"graph replay {k}", flush=True)
        g.step(1, graph=True)
        side.mul_(1.0001).add_(0.5)  # the program's own eager work between replays
        torch.cuda.synchronize()
        e.step(1, graph=False)
        torch.cuda.synchronize()
        assert torch.equal(g.buffer, e.buffer), f"replay {k} differs from the eager step"
    g.close()
    e.close()


def test_ensemble_members_changed_between_replays(torch_gpu):
    import openpfc_amd as P
    from openpfc_amd import Ensemble
    runs = _runs(13, sizes=[32, 32, 48])
    psi0 = _initial(runs, 300)
    ens = Ensemble(runs)
    _load(ens, psi0, False)
    ens.step(3)  # captured with the original eps
    new_eps = [-0.33, -0.17]
    plan, batch, off = ctypes.c_void_p(), ctypes.c_int(), ctypes.c_size_t()
    P.check(ens._lib.pfc_hip_ensemble_group(ens._h, 0, ctypes.byref(plan), ctypes.byref(batch),
                                            ctypes.byref(off)))
    assert batch.value == 2 and off.value == 0
    torch_gpu.cuda.synchronize()
    P.check(ens._lib.pfc_hip_pfc_set_members(plan, (ctypes.c_double * 2)(*new_eps), None, 0))
    ens.step(3)
    torch_gpu.cuda.synchronize()
    for i, r in enumerate(runs):
        mid = _oracle(r, psi0[i], 3)
        r2 = P.Run(r.ext, r.spacing, new_eps[i], r.dt) if i < 2 else r
        ref = _oracle(r2, mid, 3)
        err = _rel_max(ens.field(i).cpu().numpy(), ref)
        assert err <= 1e-10, f"run {i}: rel max err {err:.3e}"
    ens.close()


def test_ensemble_single_run_group_and_split_switch_between_replays(torch_gpu):
    """A single-run group holds eps and the mean as launch arguments of its
    kernels: changing them (and switching the mean-split mode) between graph
    replays must reach the replays, not only the eager path."""
    import openpfc_amd as P
    from openpfc_amd import Ensemble
    torch = torch_gpu
    runs = _runs(16, sizes=[32, 32, 48])
    psi0 = _initial(runs, 600)
    ens = Ensemble(runs)
    _load(ens, psi0, False)
    ens.step(3, graph=True)  # captured with the original eps
    plan, batch, off = ctypes.c_void_p(), ctypes.c_int(), ctypes.c_size_t()
    P.check(ens._lib.pfc_hip_ensemble_group(ens._h, 1, ctypes.byref(plan), ctypes.byref(batch),
                                            ctypes.byref(off)))
    assert batch.value == 1
    torch.cuda.synchronize()
    new_eps = -0.31
    print("single-run group: eps changed between replays", flush=True)
    P.check(ens._lib.pfc_hip_pfc_set_members(plan, (ctypes.c_double * 1)(new_eps), None, 0))
    ens.step(3, graph=True)
    torch.cuda.synchronize()
    r2 = P.Run(runs[2].ext, runs[2].spacing, new_eps, runs[2].dt)
    mid = _oracle(r2, _oracle(runs[2], psi0[2], 3), 3)
    err = _rel_max(ens.field(2).cpu().numpy(), mid)
    assert err <= 1e-10, f"eps change not replayed: rel max err {err:.3e}"
    # switch the single-run group to the mean-split step: the field then
    # holds the deviation from its (conserved) mean
    f2 = ens.field(2)
    m = float(f2.mean().item())
    f2.sub_(m)
    torch.cuda.synchronize()
    print("single-run group: mean-split switched on between replays", flush=True)
    P.check(ens._lib.pfc_hip_pfc_set_members(plan, (ctypes.c_double * 1)(new_eps),
                                             (ctypes.c_double * 1)(m), 1))
    ens.step(3, graph=True)
    torch.cuda.synchronize()
    ref = _oracle(r2, mid, 3)
    got = ens.field(2).cpu().numpy() + m
    err = _rel_max(got, ref)
    assert err <= 1e-10, f"mean-split switch not replayed: rel max err {err:.3e}"
    assert abs(float(ens.field(2).mean().item())) <= 1e-12
    # the two-run group was untouched throughout
    for i in range(2):
        err = _rel_max(ens.field(i).cpu().numpy(), _oracle(runs[i], psi0[i], 9))
        assert err <= 1e-10, f"run {i}: rel max err {err:.3e}"
    ens.close()


def test_ensemble_autotune_keeps_fields_and_parity(torch_gpu):
    from openpfc_amd import Ensemble
    torch = torch_gpu
    runs = _runs(14, sizes=[48, 64, 48, 96])
    psi0 = _initial(runs, 400)
    ens = Ensemble(runs)
    _load(ens, psi0, False)
    before = ens.buffer.clone()
    print("ensemble autotune", flush=True)
    ens.autotune(reps=1)
    torch.cuda.synchronize()
    assert torch.equal(before, ens.buffer)
    for g in ens.groups():  # every size here is below the transposed step's 128
        assert g["variants"]["step_layout"] == "natural"
        assert set(g["variants"]) == {"step_layout", "x_c2r", "col_y", "pfc_x", "pfc_z"}
    ens.step(5)
    torch.cuda.synchronize()
    for i, r in enumerate(runs):
        err = _rel_max(ens.field(i).cpu().numpy(), _oracle(r, psi0[i], 5))
        assert err <= 1e-10, f"run {i}: rel max err {err:.3e}"
    ens.close()


def test_ensemble_fp32_mixed_close_to_fp64_oracle(torch_gpu):
    import openpfc_amd as P
    from openpfc_amd import Ensemble, Run
    torch = torch_gpu
    base = _runs(15)
    psi0 = _initial(base, 500)
    # the conserved mean of each run is the exact average of its field
    runs = [Run(r.ext, r.spacing, r.eps, r.dt, float(np.mean(psi0[i])))
            for i, r in enumerate(base)]
    ens = Ensemble(runs, precision=P.F32, mean_split=True)
    _load(ens, psi0, True)
    steps = 20
    ens.step(steps)
    torch.cuda.synchronize()
    for i, r in enumerate(runs):
        ref = _oracle(r, psi0[i], steps)
        d = ens.field(i).double().cpu().numpy()
        got = r.mean + d
        rel_l2 = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
        print(f"run {i} n={r.ext[0]} rel_l2 {rel_l2:.2e} mean(d) {d.mean():.1e}")
        assert rel_l2 <= 1e-5, f"run {i}: rel L2 {rel_l2:.3e}"
        assert abs(d.mean()) <= 1e-8, f"run {i}: mean drift {d.mean():.3e}"
    ens.close()
