"""The settings-file App (openpfc_amd_app, C++ over the C ABI) on the device:

* a Tungsten run from a JSON file (constant + single_seed initial
  conditions, a fixed boundary condition, psi written at every save) against
  the oracle replaying the Simulator's order (ICs, BCs, save 0; then per
  step: BCs, step, save), every saved file within 1e-10;
* the same settings as TOML give the same files bit for bit;
* a restart (from_file + simulator {result_counter, increment}) continues
  bit for bit where the uninterrupted run is;
* the device FixedBC against the oracle.
"""
import json
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


def settings(out_pattern, t1=10.0, n=32):
    return {
        "model": {"name": "tungsten", "params": PARAMS},
        "domain": {"Lx": n, "Ly": n, "Lz": n, "dx": DX, "dy": DX, "dz": DX, "origin": "center"},
        "timestepping": {"t0": 0.0, "t1": t1, "dt": 1.0, "saveat": 1.0},
        "plan_options": {"backend": "hip"},
        "fields": [{"name": "psi", "data": out_pattern}],
        "initial_conditions": [{"target": "psi", "type": "constant", "n0": -0.4},
                               {"target": "psi", "type": "single_seed", "amp_eq": 0.215936,
                                "rho_seed": -0.047}],
        "boundary_conditions": [{"target": "psi", "type": "fixed", "rho_low": -0.464,
                                 "rho_high": -0.100}],
    }


def run_app(path):
    r = subprocess.run([APP, str(path)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    return r.stdout


def oracle_trajectory(n, steps):
    from openpfc_amd import TungstenParams
    prm = TungstenParams.from_dict(PARAMS)
    x0 = -0.5 * DX * n
    psi = O.ic_constant((n, n, n), -0.4)
    seed = O.ic_single_seed((n, n, n), (x0,) * 3, (DX,) * 3, -0.047, 0.215936, 64.0)
    psi = seed  # SingleSeed overwrites every point (0.0 outside the radius)
    O.bc_fixed(psi, x0, DX, -0.464, -0.100)
    fMF, expL, nW = O.tungsten_operators((n, n, n), (DX,) * 3, prm.operator_params(), 1.0)
    out = [psi.copy()]
    for _ in range(steps):
        O.bc_fixed(psi, x0, DX, -0.464, -0.100)
        psi, _ = O.tungsten_step(psi, fMF, expL, nW, prm.nonlinear_params(), 1)
        out.append(psi.copy())
    return out


def test_app_tungsten_json_against_oracle(torch_gpu, tmp_path):
    n = 32
    cfg = tmp_path / "run.json"
    cfg.write_text(json.dumps(settings(str(tmp_path / "out" / "psi_%d.bin"))))
    out = run_app(cfg)
    assert "10 steps" in out
    ref = oracle_trajectory(n, 10)
    for k in range(11):
        got = np.fromfile(tmp_path / "out" / f"psi_{k}.bin", dtype="<f8").reshape(n, n, n)
        err = float(np.abs(got - ref[k]).max() / np.abs(ref[k]).max())
        assert err <= 1e-10, (k, err)


def test_app_toml_equals_json_and_restart(torch_gpu, tmp_path):
    import tomli  # noqa: F401  (the TOML text below is written by hand)
    n = 32
    full = tmp_path / "full"
    cfg = tmp_path / "run.json"
    cfg.write_text(json.dumps(settings(str(full / "psi_%d.bin"))))
    run_app(cfg)
    # TOML form of the same settings
    toml_dir = tmp_path / "toml"
    lines = ['[model]', 'name = "tungsten"', '[model.params]']
    lines += [f"{k} = {json.dumps(v)}" for k, v in PARAMS.items()]
    lines += ['[domain]', f'Lx = {n}', f'Ly = {n}', f'Lz = {n}', f'dx = {DX!r}', f'dy = {DX!r}',
              f'dz = {DX!r}', 'origin = "center"', '[timestepping]', 't0 = 0.0', 't1 = 10.0',
              'dt = 1.0', 'saveat = 1.0', '[plan_options]', 'backend = "rocm"',
              '[[fields]]', 'name = "psi"', f'data = "{toml_dir}/psi_%d.bin"',
              '[[initial_conditions]]', 'target = "psi"', 'type = "constant"', 'n0 = -0.4',
              '[[initial_conditions]]', 'target = "psi"', 'type = "single_seed"',
              'amp_eq = 0.215936', 'rho_seed = -0.047',
              '[[boundary_conditions]]', 'target = "psi"', 'type = "fixed"',
              'rho_low = -0.464', 'rho_high = -0.100']
    tcfg = tmp_path / "run.toml"
    tcfg.write_text("\n".join(lines) + "\n")
    run_app(tcfg)
    for k in range(11):
        a = (full / f"psi_{k}.bin").read_bytes()
        assert a == (toml_dir / f"psi_{k}.bin").read_bytes(), k
    # restart from the file of step 5: files 6..10 are the same bits
    rdir = tmp_path / "restart"
    s = settings(str(rdir / "psi_%d.bin"))
    s["initial_conditions"] = [{"target": "psi", "type": "from_file",
                                "filename": str(full / "psi_5.bin")}]
    s["simulator"] = {"result_counter": 5, "increment": 5}
    rcfg = tmp_path / "restart.json"
    rcfg.write_text(json.dumps(s))
    run_app(rcfg)
    assert not (rdir / "psi_5.bin").exists()
    for k in range(6, 11):
        assert (rdir / f"psi_{k}.bin").read_bytes() == (full / f"psi_{k}.bin").read_bytes(), k


def test_device_fixed_bc_matches_oracle(torch_gpu):
    import ctypes
    import openpfc_amd as P
    torch = torch_gpu
    n = (128, 8, 4)
    f = P.HipFft(*n)
    x0, dx = 0.0, 0.75  # xpos = 76: the band covers x in (56, 96)
    a = np.random.default_rng(3).standard_normal(n[0] * n[1] * n[2])
    d = torch.from_numpy(a).cuda()
    P.check(f._lib.pfc_hip_bc_fixed(f.handle, ctypes.c_void_p(d.data_ptr()), x0, dx, -0.464,
                                    -0.1, 20.0, 1.0, None))
    ref = a.reshape(n[2], n[1], n[0]).copy()
    O.bc_fixed(ref, x0, dx, -0.464, -0.1)
    got = d.cpu().numpy().reshape(n[2], n[1], n[0])
    assert (ref != a.reshape(ref.shape)).sum() == 53 * n[1] * n[2]  # 53 band columns
    assert np.abs(got - ref).max() <= 4e-16
