"""Rotated seeds, SeedGrid, RandomSeeds and MovingBC on the device, and the
reference's seed_grid + fixed / moving Tungsten configurations through the
App, against the oracle.

* pfc_hip_ic_seeds against orc_ic_seeds (FP64 and FP32 plans, 0.0 outside
  and keep_outside);
* pfc_hip_column_max against numpy's max over (y, z);
* pfc_hip_bc_moving against the oracle's band, including wrapped bands;
* apps/tungsten/inputs_json/tungsten_fixed_bc.json and
  tungsten_moving_bc.json (same model, domain, initial and boundary
  conditions; 20 steps instead of 20000) run by openpfc_amd_app, every saved
  field within 1e-10 of the oracle replaying the Simulator's order;
  tungsten_moving_bc_options.json (no "xpos") is rejected with the
  reference's message.
"""
import json
import math
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(ROOT, "openpfc_amd", "host", "openpfc_amd_app")
DX = 1.1107207345395915
PARAMS = {"n0": -0.10, "alpha": 0.50, "n_sol": -0.047, "n_vap": -0.464, "T": 3300.0,
          "T0": 156000.0, "Bx": 0.8582, "alpha_farTol": 0.001, "alpha_highOrd": 4,
          "lambda": 0.22, "stabP": 0.2, "shift_u": 0.3341, "shift_s": 0.1898, "p2": 1.0,
          "p3": -0.5, "p4": 0.333333333, "q20": -0.0037, "q21": 1.0, "q30": -12.4567,
          "q31": 20.0, "q40": 45.0}
GRID = {"type": "seed_grid", "X0": 30.0, "Ny": 2, "Nz": 2, "radius": 25.0,
        "amplitude": 0.215936, "rho": -0.047}


def _dev(torch, a, f64=True):
    return torch.from_numpy(np.ascontiguousarray(a).ravel()).to(
        "cuda", torch.float64 if f64 else torch.float32)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_device_seed_fields_match_oracle(torch_gpu, prec):
    import openpfc_amd as P
    torch = torch_gpu
    f64 = prec == "f64"
    tol = 4e-15 if f64 else 4e-7
    print(f"seed fields {prec}", flush=True)
    grid = O.seed_grid((128, 64, 64), (DX,) * 3, 2, 2, 30.0, 25.0, -0.047, 0.215936)
    f = P.HipFft(128, 64, 64, precision=P.F64 if f64 else P.F32)
    psi = _dev(torch, np.full((64, 64, 128), 3.0), f64)
    P.ic_seeds(f, psi, (0, 0, 0), (DX,) * 3, grid)
    ref = O.ic_seeds(np.full((64, 64, 128), 3.0), (0, 0, 0), (DX,) * 3, grid)
    got = psi.cpu().double().numpy().reshape(ref.shape)
    assert np.abs(got - ref).max() <= tol * 3
    assert np.count_nonzero(ref) > 10000 and np.count_nonzero(ref == 0.0) > 10000
    psi = _dev(torch, np.full((64, 64, 128), 3.0), f64)
    P.ic_seeds(f, psi, (0, 0, 0), (DX,) * 3, grid, keep_outside=True)
    ref = O.ic_seeds(np.full((64, 64, 128), 3.0), (0, 0, 0), (DX,) * 3, grid, keep_outside=True)
    assert np.abs(psi.cpu().double().numpy().reshape(ref.shape) - ref).max() <= tol * 3
    # RandomSeeds: 150 overlapping seeds on the reference's [-128, 128] box
    rnd = O.random_seeds(-0.047, 0.215936)
    g = P.HipFft(64, 64, 64, precision=P.F64 if f64 else P.F32)
    psi = _dev(torch, np.zeros((64, 64, 64)), f64)
    P.ic_seeds(g, psi, (-128,) * 3, (4.0,) * 3, rnd)
    ref = O.ic_seeds(np.zeros((64, 64, 64)), (-128,) * 3, (4.0,) * 3, rnd)
    assert np.abs(psi.cpu().double().numpy().reshape(ref.shape) - ref).max() <= tol * 3


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_device_column_max(torch_gpu, prec):
    import openpfc_amd as P
    torch = torch_gpu
    f64 = prec == "f64"
    for n in ((128, 64, 64), (2048, 4, 96), (64, 1, 1)):
        print(f"column max {prec} {n}", flush=True)
        a = np.random.default_rng(n[0]).standard_normal((n[2], n[1], n[0]))
        if not f64:
            a = a.astype(np.float32).astype(np.float64)
        f = P.HipFft(*n, precision=P.F64 if f64 else P.F32)
        got = P.column_max(f, _dev(torch, a, f64))
        assert np.array_equal(got, a.max(axis=(0, 1)))


def test_device_moving_band_matches_oracle(torch_gpu):
    import openpfc_amd as P
    torch = torch_gpu
    n = (128, 8, 4)
    f = P.HipFft(*n)
    l = n[0] * DX
    for xpos in (60.3, 127.17225402106772, 2.0, l + 5.0, 300.0):
        print(f"moving band xpos {xpos}", flush=True)
        a = np.random.default_rng(7).standard_normal((n[2], n[1], n[0]))
        d = _dev(torch, a)
        P.bc_moving(f, d, 0.0, DX, xpos, -0.464, -0.1, 15.0, 1.0)
        # the oracle's band for this position: threshold above every value,
        # disp 0, so the boundary stays at xpos
        bc = O.MovingBC(-0.464, -0.1, xwidth=15.0, alpha=1.0, disp=0.0, xpos=xpos,
                        threshold=1e300)
        ref = a.copy()
        bc.apply(ref, 0.0, DX)
        assert bc.xpos == xpos
        got = d.cpu().numpy().reshape(ref.shape)
        assert np.abs(got - ref).max() <= 4e-16
        assert np.count_nonzero(got != a) > 0


def _settings(out_pattern, bc, t1=20.0):
    return {
        "model": {"name": "tungsten", "params": PARAMS},
        "domain": {"Lx": 128, "Ly": 64, "Lz": 64, "dx": DX, "dy": DX, "dz": DX,
                   "origin": "corner"},
        "timestepping": {"t0": 0.0, "t1": t1, "dt": 1.0, "saveat": 5.0},
        "fields": [{"name": "psi", "data": out_pattern}],
        "initial_conditions": [{"target": "psi", "type": "constant", "n0": -0.10},
                               dict(GRID, target="psi")],
        "boundary_conditions": [dict(bc, target="psi")],
    }


def _oracle_run(bc_apply, steps):
    from openpfc_amd import TungstenParams
    prm = TungstenParams.from_dict(PARAMS)
    n = (128, 64, 64)
    grid = O.seed_grid(n, (DX,) * 3, 2, 2, 30.0, 25.0, -0.047, 0.215936)
    psi = O.ic_seeds(np.full((64, 64, 128), -0.10), (0, 0, 0), (DX,) * 3, grid)
    fMF, expL, nW = O.tungsten_operators(n, (DX,) * 3, prm.operator_params(), 1.0)
    bc_apply(psi)
    out = {0: psi.copy()}
    for k in range(1, steps + 1):
        bc_apply(psi)
        psi, _ = O.tungsten_step(psi, fMF, expL, nW, prm.nonlinear_params(), 1)
        if k % 5 == 0:
            out[k // 5] = psi.copy()
    return out


def _run_app(cfg):
    r = subprocess.run([APP, str(cfg)], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    return r


@pytest.mark.parametrize("kind", ["fixed", "moving"])
def test_app_seed_grid_configs_match_oracle(torch_gpu, tmp_path, kind):
    if kind == "fixed":
        bc = {"type": "fixed", "rho_low": -0.464, "rho_high": -0.100}

        def apply(a):
            O.bc_fixed(a, 0.0, DX, -0.464, -0.100)
    else:
        bc = {"type": "moving", "rho_low": -0.464, "rho_high": -0.10, "width": 15.0,
              "alpha": 1.0, "disp": 40.0, "xpos": 127.17225402106772}
        mb = O.MovingBC(-0.464, -0.10, xwidth=15.0, alpha=1.0, disp=40.0,
                        xpos=127.17225402106772)

        def apply(a):
            mb.apply(a, 0.0, DX)
    cfg = tmp_path / "run.json"
    cfg.write_text(json.dumps(_settings(str(tmp_path / "out" / "u_%04d.bin"), bc)))
    print(f"app tungsten seed_grid + {kind}", flush=True)
    r = _run_app(cfg)
    assert r.returncode == 0, r.stderr
    assert "20 steps" in r.stdout
    ref = _oracle_run(apply, 20)
    for k, want in ref.items():
        got = np.fromfile(tmp_path / "out" / f"u_{k:04d}.bin", dtype="<f8").reshape(want.shape)
        err = float(np.abs(got - want).max() / np.abs(want).max())
        print(f"save {k}: rel max err {err:.3e}", flush=True)
        assert err <= 1e-10, (k, err)


def test_app_rejects_moving_without_xpos(torch_gpu, tmp_path):
    # apps/tungsten/inputs_json/tungsten_moving_bc_options.json has no "xpos"
    bc = {"type": "moving", "initial_position": "end", "rho_low": -0.464, "rho_high": -0.10,
          "width": 3.0, "alpha": 10.0, "disp": 80}
    cfg = tmp_path / "run.json"
    cfg.write_text(json.dumps(_settings(str(tmp_path / "u_%04d.bin"), bc)))
    r = _run_app(cfg)
    assert r.returncode != 0
    assert "missing or invalid 'xpos' field" in (r.stdout + r.stderr)
