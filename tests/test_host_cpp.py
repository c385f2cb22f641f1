"""Drives the C++ host-layer checks (openpfc_amd/host/tests/test_host.cpp).

CPU: Time/Simulator step counts and save points (time.hpp:732-754,972-978;
simulator_integrator.hpp:43-61), World coordinates, k-space helpers,
SingleSeed/Constant values, error behaviour.  GPU: HipFft overloads against a
direct DFT, the reference's size-mismatch messages, and a PFCModelHIP run
through Simulator compared with the CPU oracle from the same initial field.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "openpfc_amd", "host", "test_host")


def _run(*args, timeout=120):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} missing: run __graft_entry__.build()")
    r = subprocess.run([BIN, *args], capture_output=True, text=True, timeout=timeout)
    print(r.stdout, r.stderr)
    return r


def test_host_layer_cpu():
    r = _run("cpu")
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_host_layer_hipfft(torch_gpu):
    r = _run("gpu-fft")
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("n,steps", [(32, 30), (64, 50)])
def test_host_pfc_model_simulator_matches_oracle(torch_gpu, tmp_path, n, steps):
    out = tmp_path / "psi.bin"
    r = _run("gpu-sim", str(n), str(steps), str(out))
    assert r.returncode == 0, r.stderr
    got = np.fromfile(out, dtype=np.float64).reshape(n, n, n)
    psi0 = np.full((n, n, n), -0.2)
    O.add_counter_noise(psi0, 1234, -0.01, 0.01)
    dx = math.pi / 4
    opL, opN = O.pfc_operators((n, n, n), (dx, dx, dx), -0.25, 0.5)
    ref = O.pfc_step(psi0, opL, opN, steps)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err <= 1e-10, err
    F = float(next(ln for ln in r.stdout.splitlines() if ln.startswith("free_energy")).split()[1])
    Fref = O.pfc_free_energy(ref, (dx, dx, dx), -0.25)
    assert abs(F - Fref) <= 1e-10 * abs(Fref), (F, Fref)


@pytest.mark.gpu
def test_host_tungsten_hip_cpu_vs_hip_recipe(torch_gpu, tmp_path):
    """TungstenHIP<double> (C++ drop-in) on the reference's CPU-vs-HIP recipe
    (apps/tungsten/tests/test_tungsten_cpu_vs_hip.cpp:30-90): every point
    within 1e-10 of the oracle after 10 steps; psi_MF of the last step too."""
    from openpfc_amd.tungsten import TungstenParams
    out, mfout = tmp_path / "psi.bin", tmp_path / "mf.bin"
    r = _run("gpu-tung", "10", str(out), str(mfout))
    assert r.returncode == 0, r.stderr
    n = 32
    got = np.fromfile(out, dtype=np.float64).reshape(n, n, n)
    mf = np.fromfile(mfout, dtype=np.float64).reshape(n, n, n)
    prm = TungstenParams(n0=-0.4, T=0.5)
    fMF, expL, nW = O.tungsten_operators((n, n, n), (1.0, 1.0, 1.0), prm.operator_params(),
                                         0.01)
    N = n ** 3
    psi0 = (-0.4 + 0.1 * np.sin(2.0 * math.pi * np.arange(N) / N)).reshape(n, n, n)
    ref, ref_mf = O.tungsten_step(psi0, fMF, expL, nW, prm.nonlinear_params(), 10)
    assert float(np.abs(got - ref).max()) < 1e-10
    assert float(np.abs(mf - ref_mf).max()) < 1e-10


@pytest.mark.gpu
def test_host_reference_call_sites(torch_gpu):
    """create_hip(decomposition, rank, MPI_COMM_WORLD), create_with_backend
    (layout, rank, options, Backend::HIP, MPI_COMM_WORLD), the model and the
    Simulator with their communicator arguments: the same field, bit for bit,
    as the World-based construction."""
    r = _run("gpu-callsites", "32", "12")
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_host_autotune_cache(torch_gpu, tmp_path):
    """PFCModelHIP::initialize tunes and stores its step variants in
    <dir>/<device>_autotune_cache.json, the next model of that size loads
    them; OPENPFC_GPU_AUTOTUNE_MODE=cache_only refuses a size without an
    entry, OPENPFC_GPU_AUTOTUNE_DISABLE=1 keeps the defaults, and an entry
    naming an unknown variant is tuned again."""
    r = _run("gpu-autotune", str(tmp_path))
    assert r.returncode == 0, r.stderr
    files = list(tmp_path.glob("gfx*_autotune_cache.json"))
    assert len(files) == 1, files
