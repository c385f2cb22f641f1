#!/usr/bin/env python3
"""Benchmark of the fused spectral PFC ETD1 step on one MI355X.

Workload (default): plain conserved PFC, 1024^3 periodic grid, FP64,
dx = pi/4, dt = 0.5, r = -0.25; psi0 = -0.2 + U(-0.01, 0.01) (counter-based
splitmix64 stream, seed 1234) with a one-mode bcc seed (rho = -0.2,
amplitude 0.1, radius 64*dx, untouched outside) at the centre of a grid whose
origin is -N*dx/2 per axis.  A step is one ETD1 step:
psi_hat = opL*F(psi) + opN*F(psi^3), psi = F^-1(psi_hat)
(docs/api/examples/10_complete_pfc_model.cpp:121-144).

    python bench.py [--gpus 1] [--steps K] [--warmup W]
    python bench.py --model tungsten     # Tungsten mean-field step (see below)

Prints one JSON line.  `value` is FP64 steps/s with the field resident in
HBM, timed over one K-step call (which fuses step boundaries when the
autotuner found that faster); `step_hbm_frac` is the per-step model bytes
2R + 12C (R a real field, C a half spectrum) over the step time, and
`run_hbm_frac` the bytes the timed run itself needs per step (12C + 2R/K when
step boundaries are fused).  The same workload in FP32-mixed mode (FP32 deviation field and
transforms, FP64 conserved mean) is reported under "fp32_mixed" with its
error against the FP64 run after the same steps.  `roofline` is the
dominant kernel's algorithmic bytes per launch over its mean launch duration
(HIP events on the stream the kernels run on); `stream_copy` is the
library's plain streaming copy on arrays of one half spectrum, the bandwidth a
pass could reach on the same GPU, and `roofline.frac_of_stream_copy` /
`step_frac_of_stream_copy` compare against it.  `cpu_baseline` times the CPU
oracle (the repository's C restatement of the reference path, OpenMP) on one
step of the same 1024^3 workload.

--model tungsten times the Tungsten mean-field step (TungstenHIP drop-in) on
the reference's scalability workload (apps/tungsten/src/tungsten_scalability.cpp:
95-135): spacing 1, n0 = -0.4, T = 0.5 (other parameters at their defaults),
dt = 0.01, psi = -0.4 + 0.1 sin(2 pi i/N^3), 3 warm-up steps, FP64 with the
FP32 mode (TungstenHIP<float>) beside it.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "steps/s at 1024³: FP32-mixed vs FP64, ns/DOF/step, % of HBM peak"
HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
# One-mode amplitude of the seed.  The reference demo's 0.3 makes the explicit
# psi^3 term of ETD1 unstable at dt = 0.5 (the CPU oracle diverges within 8
# steps as well); 0.1 is stable and near the equilibrium amplitude.
SEED_AMP = 0.1
PASS_NAMES = ["x_fwd_pfc", "y_fwd", "z_fused", "y_inv", "x_inv_c2r"]
# (kernel symbol, variant role) of each pass in the two step layouts
PASS_KERNELS = {
    "transposed": [("k_tx_fwd", "t_x_fwd"), ("k_ty_fwd", "t_y_fwd"), ("k_tz", "t_z"),
                   ("k_ty_inv", "t_y_inv"), ("k_tx_inv", "t_x_inv")],
    "natural": [("k_pfc_x", "pfc_x"), ("k_col", "col_y"), ("k_pfc_z", "pfc_z"),
                ("k_col", "col_y"), ("k_x_c2r", "x_c2r")]}
STEP_ROLES = ["step_layout", "pfc_x", "col_y", "pfc_z", "x_c2r", "t_x_fwd", "t_y_fwd",
              "t_z", "t_y_inv", "t_x_inv", "t_x_fuse", "step_fuse"]
# launches of a fused multi-step run (pfc_hip_pfc_step_timed with 6 entries)
FUSED_NAMES = ["x_fwd_first", "y_fwd", "z_fused", "y_inv", "x_inv_last", "x_boundary_fused"]
TUNG_METRIC = "Tungsten mean-field step, steps/s at 1024^3 FP64 (% of HBM peak)"
TUNG_PASS_NAMES = ["x_fwd_r2c", "y_fwd", "z_meanfield", "y_inv_mf", "x_fused_nl",
                   "y_fwd_n", "z_combine", "y_inv", "x_inv_c2r"]
TUNG_PASS_KERNELS = {
    "transposed": [("k_tx_r2c", "t_x_r2c"), ("k_ty_fwd", "t_y_fwd"), ("k_tz", "t_z"),
                   ("k_ty_inv", "t_y_inv"), ("k_tx_tung", "t_x_tung"), ("k_ty_fwd", "t_y_fwd"),
                   ("k_tz", "t_z"), ("k_ty_inv", "t_y_inv"), ("k_tx_inv", "t_x_inv")],
    "natural": [("k_x_r2c", "x_r2c"), ("k_col", "col_y"), ("k_pfc_z", "pfc_z"),
                ("k_col", "col_y"), ("k_x_tung", "x_tung"), ("k_col", "col_y"),
                ("k_pfc_z", "pfc_z"), ("k_col", "col_y"), ("k_x_c2r", "x_c2r")]}
TUNG_STEP_ROLES = ["step_layout", "x_r2c", "col_y", "pfc_z", "x_tung", "x_c2r", "t_x_r2c",
                   "t_y_fwd", "t_z", "t_y_inv", "t_x_tung", "t_x_inv"]


# kernels whose last template argument selects non-temporal line accesses
NT_KERNELS = ("k_tz", "k_tx_inv", "k_tx_fwd_w", "k_ty_fwd_w", "k_ty_inv_w")


def variant_key(kernel, prec, n, variant, direction=None, mode=0):
    """rocprof short name (openpfc_amd.profile_summary.short_name) of a variant;
    `mode` is the z-pass kind (0 PFC, 1 Tungsten mean field, 2 Tungsten combine)."""
    import re
    t = "double" if prec == "f64" else "float"
    if kernel == "k_tx_fuse":  # "rA/EA+rB/EB/G8/split[/wN]": nx/2- and nx-point plans
        half, full = variant.split("+")
        ra, ea = half.split("/")
        pb = full.split("/")
        wpe = pb[4][1:] if len(pb) > 4 and pb[4][:1] == "w" else "1"  # waves per SIMD cap
        return (f"k_tx_fuse<{t},N{n // 2}/E{ea[1:]},r{ra[1:]}+N{n}/E{pb[1][1:]},r{pb[0][1:]},"
                f"{pb[2][1:]},{'true' if pb[3] == 'split' else 'false'},{wpe}>")
    parts = variant.split("/")
    nt = "nt" in parts  # non-temporal lines: the kernel's last template argument
    xp = "xp" in parts  # a plan with double butterflies
    parts = [p for p in parts if p not in ("nt", "xp")]
    if kernel == "k_tx_fwd" and "r2c2" in parts:
        kernel, n = "k_tx_fwd2", n // 2
    if kernel == "k_tx_fwd" and "wide" in parts:
        kernel = "k_tx_fwd_w"
    wide_nr = [p for p in parts if p.startswith("wide") and p != "wide"]
    if kernel in ("k_ty_fwd", "k_ty_inv") and wide_nr:
        kernel = kernel + "_w"
    radices = parts[0][1:]
    e = parts[1][1:]
    nn = n
    if kernel in ("k_tx_inv", "k_x_c2r", "k_x_r2c", "k_tx_r2c", "k_x_tung", "k_tx_tung"):
        nn = n // 2
    rest = []
    for p in parts[2:]:
        if p.startswith(("G", "L", "rows")):
            rest.append(re.sub(r"^(G|L|rows)", "", p))
        elif p in ("split", "full"):
            rest.append("true" if p == "split" else "false")
    if kernel == "k_col":
        rest = [str(direction)] + rest
    pair = "mtp" in parts  # lines ky and ny - ky sharing one table
    if kernel == "k_tz" and ("mt" in parts or pair):  # multipliers tabulated per line
        kernel = "k_tz_mt"
    if kernel in ("k_tz", "k_tz_mt", "k_pfc_z"):
        rest.append(str(mode))
        if kernel == "k_tz_mt":
            rest.append("true" if pair else "false")
        if kernel == "k_tz":
            rest.append("true" if "pre" in parts else "false")
    elif kernel in ("k_ty_fwd_w", "k_ty_inv_w"):
        rest.append(wide_nr[0][4:])
    elif kernel.startswith("k_t") and kernel != "k_tx_fwd_w":
        rest.append("false" if "direct" in parts else "true")
        if kernel in ("k_tx_r2c", "k_tx_tung"):
            rest.append("true" if "wide" in parts else "false")
        wpe = [p[1:] for p in parts if p[:1] == "w" and p[1:].isdigit()]
        if kernel == "k_tx_tung" and wpe:  # registers capped for wpe waves per SIMD
            kernel = "k_tx_tung_o"
            rest.append(wpe[0])
    if kernel in NT_KERNELS:
        rest.append("true" if nt else "false")
    t = "double" if prec == "f64" else "float"
    return (f"{kernel}<{t},N{nn}/E{e}{'/xp' if xp else ''},r{radices}"
            + ("," + ",".join(rest) if rest else "") + ">")


def pass_bytes(n: int, esize: int):
    """Algorithmic HBM bytes per launch of each pass: one read and one write of
    every array the pass streams (opL/opN are computed per mode, not read)."""
    R = esize * n ** 3                     # one real field
    C = 2 * esize * (n // 2 + 1) * n * n   # one half spectrum
    return {"x_fwd_pfc": R + 2 * C, "y_fwd": 4 * C, "z_fused": 3 * C,
            "y_inv": 2 * C, "x_inv_c2r": C + R}


def fused_pass_bytes(n: int, esize: int):
    """Algorithmic bytes per launch of the steady-state passes of a fused
    multi-step run: y forward 4C, z 3C, y inverse 2C, fused boundary C + 2C."""
    C = 2 * esize * (n // 2 + 1) * n * n
    return {"y_fwd": 4 * C, "z_fused": 3 * C, "y_inv": 2 * C, "x_boundary_fused": 3 * C}


def tung_pass_bytes(n: int, esize: int):
    """Algorithmic bytes per launch of the nine Tungsten passes (3R + 18C in
    all): psi is read by the forward x pass and again by the fused x pass, the
    first z pass writes the full spectrum of psi and the mean field, the
    second reads both spectra; the multipliers are computed per mode."""
    R = esize * n ** 3
    C = 2 * esize * (n // 2 + 1) * n * n
    return {"x_fwd_r2c": R + C, "y_fwd": 2 * C, "z_meanfield": 3 * C, "y_inv_mf": 2 * C,
            "x_fused_nl": R + 2 * C, "y_fwd_n": 2 * C, "z_combine": 3 * C, "y_inv": 2 * C,
            "x_inv_c2r": C + R}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def setup(P, torch, n, prec, dx, dt, r):
    fft = P.HipFft(n, n, n, precision=prec)
    st = P.PfcStepper(fft, (dx, dx, dx), r)
    st.initialize(dt)
    dtype = torch.float64 if prec == P.F64 else torch.float32
    psi = torch.empty(n ** 3, dtype=dtype, device="cuda")
    origin = (-0.5 * n * dx,) * 3
    P.ic_constant(fft, psi, -0.2)
    P.ic_add_noise(fft, psi, 1234, -0.01, 0.01)
    P.ic_single_seed(fft, psi, origin, (dx, dx, dx), -0.2, SEED_AMP, 64 * dx, keep_outside=True)
    torch.cuda.synchronize()
    return fft, st, psi


def choose_variants(P, fft, st, psi, args, log_prefix, roles=STEP_ROLES):
    """Autotune (untimed) or apply forced variants; returns {role: name}."""
    import json as _json
    if args.variants:
        with open(args.variants) as fh:
            want = _json.load(fh)
        for role in roles:
            if role in want:
                names = fft.variants(role)
                if want[role] in names:
                    fft.set_variant(role, names.index(want[role]))
    elif not args.no_autotune:
        t0 = time.perf_counter()
        st.autotune(psi, reps=2)
        log(f"{log_prefix} autotune {time.perf_counter() - t0:.1f} s")
    return {r: fft.variants(r)[fft.get_variant(r)] for r in roles if fft.variants(r)}


def time_steps(torch, st, psi, steps, warmup):
    st.step(psi, warmup)
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(s)
    st.step(psi, steps)
    e1.record(s)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return e0.elapsed_time(e1) / steps, wall


def cpu_info() -> dict:
    """Host the CPU baseline ran on: CPU model, cores, OpenMP setting."""
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(),
            "cores_available": len(os.sched_getaffinity(0)),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(n, dx, dt, r, nsteps, warmup=3, n_single=256):
    """The oracle (the C restatement of the reference path) on this host, with
    the method of apps/tungsten/src/tungsten_scalability.cpp:95-135: set up,
    `warmup` untimed steps, then `nsteps` timed steps, with the share of the
    time spent in transforms.  The first row uses every OpenMP thread on the
    n^3 workload; the second is the reference's literal one-rank,
    single-threaded run, sampled at n_single^3 so that it stays within
    seconds (its ns per DOF-step carries over to n^3 up to the log n of the
    FFT)."""
    import numpy as np
    from oracle import oracle as O

    def run(N, steps):
        a = np.full((N, N, N), -0.2)
        O.add_counter_noise(a, 1234, -0.01, 0.01)
        O.ic_single_seed_masked(a, (-0.5 * N * dx,) * 3, (dx, dx, dx), -0.2, SEED_AMP,
                                64 * dx, 1)
        opL, opN = O.pfc_operators((N, N, N), (dx, dx, dx), r, dt)
        O.pfc_step(a, opL, opN, warmup)
        O.fft_seconds(reset=True)
        t0 = time.perf_counter()
        O.pfc_step(a, opL, opN, steps)
        el = time.perf_counter() - t0
        return steps / el, O.fft_seconds(reset=True) / el

    threads = O.num_threads()
    sps, share = run(n, nsteps)
    O.set_num_threads(1)
    try:
        sps1, share1 = run(n_single, 1)
    finally:
        O.set_num_threads(threads)
    return {"value": sps, "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": f"{nsteps} timed step(s) after {warmup} warm-up steps of the same "
                      f"{n}^3 FP64 workload, oracle C restatement (own Stockham FFT, "
                      f"OpenMP {threads} threads)",
            "fft_share": share, "warmup_steps": warmup, **cpu_info(),
            "single_thread": {"grid": n_single, "steps_per_s": sps1,
                              "ns_per_dof_step": 1e9 / sps1 / n_single ** 3,
                              "fft_share": share1,
                              "sample": f"1 timed step after {warmup} warm-up steps at "
                                        f"{n_single}^3, 1 thread (1 rank)"}}


def cpp_app_leg(P, torch, n, dx, dt, r, steps, entry):
    """The C++ drop-in path at n^3: the settings-file App (openpfc_amd_app,
    PFCModelHIP under the Simulator) with the variants of `entry` handed over
    through its autotune cache (OPENPFC_GPU_AUTOTUNE_MODE=cache_only, so a
    cache the App cannot apply fails the leg).  Two runs of 3 and 3 + steps
    steps; their difference over `steps` is the App's time per step without
    its first step (which allocates the step buffers)."""
    import re
    import shutil
    import subprocess
    import tempfile
    app = os.path.join(ROOT, "openpfc_amd", "host", "openpfc_amd_app")
    d = tempfile.mkdtemp(prefix="pfc_app_")
    try:
        key = P.step_key("pfc", P.F64, (n, n, n))
        with open(os.path.join(d, f"{P.device_id()}_autotune_cache.json"), "w") as fh:
            json.dump({"device_id": P.device_id(), "entries": {key: entry}}, fh)
        env = dict(os.environ, OPENPFC_GPU_AUTOTUNE_CACHE_DIR=d,
                   OPENPFC_GPU_AUTOTUNE_MODE="cache_only")
        totals = {}
        for k in (3, 3 + steps):
            cfg = {"model": {"name": "pfc", "params": {"epsilon": r}},
                   "domain": {"Lx": n, "Ly": n, "Lz": n, "dx": dx, "dy": dx, "dz": dx,
                              "origin": "center"},
                   "timestepping": {"t0": 0.0, "t1": k * dt, "dt": dt, "saveat": k * dt},
                   "initial_conditions": [
                       {"target": "density", "type": "constant", "n0": -0.2},
                       {"target": "density", "type": "uniform_noise", "seed": 1234,
                        "low": -0.01, "high": 0.01}]}
            path = os.path.join(d, f"run{k}.json")
            with open(path, "w") as fh:
                json.dump(cfg, fh)
            log(f"c++ app leg: {k} steps")
            res = subprocess.run([app, path], capture_output=True, text=True, timeout=600,
                                 env=env)
            if res.returncode != 0:
                raise RuntimeError(f"openpfc_amd_app failed: {res.stdout}{res.stderr}")
            m = re.search(r"\[app\] (\d+) steps in ([0-9.]+) s", res.stdout)
            if not m or int(m.group(1)) != k:
                raise RuntimeError(f"openpfc_amd_app output not understood: {res.stdout}")
            totals[k] = float(m.group(2))
        ms = (totals[3 + steps] - totals[3]) * 1e3 / steps
        return {"ms_per_step": ms, "steps_per_s": 1e3 / ms, "steps": steps,
                "method": f"openpfc_amd_app wall time of {3 + steps} minus 3 steps, "
                          "variants from the Python leg via the autotune cache"}
    finally:
        shutil.rmtree(d, ignore_errors=True)


def stream_copy(P, torch, nbytes, reps=5):
    """GB/s of the library's plain streaming copy (pfc_hip_stream_copy: one
    16-byte value per thread, nbytes read + nbytes written per launch) on
    the current stream: the bandwidth a pass could reach on this GPU."""
    import ctypes
    from openpfc_amd._lib import check, load
    lib = load()
    nbytes = nbytes // 16 * 16
    a = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    s = torch.cuda.current_stream()
    sp = ctypes.c_void_p(s.cuda_stream)

    def launch():
        check(lib.pfc_hip_stream_copy(ctypes.c_void_p(a.data_ptr()),
                                      ctypes.c_void_p(b.data_ptr()), nbytes, sp))
    launch()
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(s)
        launch()
        e1.record(s)
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1))
    ok = bool(torch.equal(a, b))
    del a, b
    torch.cuda.empty_cache()
    return {"GBps": 2 * nbytes / (best * 1e-3) / 1e9, "bytes_per_launch": 2 * nbytes,
            "ms": best, "copy_correct": ok}


def copy_calibration(torch, nbytes=4 << 30, reps=5):
    """Device-to-device copy of nbytes (read nbytes + write nbytes per launch)."""
    a = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(reps):
        b.copy_(a)
    e1.record(s)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    torch.cuda.empty_cache()
    return {"bytes_per_launch": 2 * nbytes, "ms": ms, "GBps": 2 * nbytes / (ms * 1e-3) / 1e9}


def accuracy_run(P, torch, n, dx, dt, r, steps, every=100):
    """Accuracy run: FP64 and the two FP32-mixed modes (FP32 butterflies;
    FP64 butterflies on FP32 fields) advanced in lockstep from the same
    field; relative L2 / max error of each against FP64 every `every` steps
    (the FP64 GPU step is held to 1e-10 of the CPU oracle)."""
    fft64, st64, psi = setup(P, torch, n, P.F64, dx, dt, r)
    mean0 = float(psi.mean().item())
    mixed = {}
    for key, flags in (("fp32_mixed", 0), ("fp32_mixed_f64_butterflies",
                                            P.FLAG_F64_BUTTERFLIES)):
        f = P.HipFft(n, n, n, precision=P.F32, flags=flags)
        st = P.PfcStepper(f, (dx, dx, dx), r)
        st.initialize(dt)
        st.set_mean(mean0)
        mixed[key] = (f, st, (psi - mean0).float())
    # control: FP64 from the initial field rounded to FP32 (one FP32-sized
    # perturbation of the state, no FP32 arithmetic afterwards)
    fftc = P.HipFft(n, n, n, precision=P.F64)
    stc = P.PfcStepper(fftc, (dx, dx, dx), r)
    stc.initialize(dt)
    pc = psi.float().double()
    rows = []
    done = 0
    while done < steps:
        k = min(every, steps - done)
        st64.step(psi, k)
        for f, st, d in mixed.values():
            st.step(d, k)
        stc.step(pc, k)
        done += k
        nrm = torch.linalg.vector_norm(psi)
        mx0 = psi.abs().max()
        row = {"step": done}
        for key, (f, st, d) in mixed.items():
            diff = d.double() + mean0 - psi
            row[key] = {"rel_l2": (torch.linalg.vector_norm(diff) / nrm).item(),
                        "rel_max": (diff.abs().max() / mx0).item(),
                        "mean_drift": abs((d.double().mean() + mean0).item() - mean0)}
            del diff
        row["fp64_from_fp32_rounded_start_rel_l2"] = \
            (torch.linalg.vector_norm(pc - psi) / nrm).item()
        rows.append(row)
        log(f"accuracy step {done}: " + ", ".join(
            f"{key} rel L2 {row[key]['rel_l2']:.3e} rel max {row[key]['rel_max']:.3e}"
            for key in mixed) + f" (fp64 control {row['fp64_from_fp32_rounded_start_rel_l2']:.3e})")
    return {"metric": "fp32-mixed error vs fp64", "grid": [n, n, n], "steps": steps,
            "control": "fp64 run started from the fp32-rounded initial field",
            "bar": {"rel_l2": 1e-4, "after_steps": 1000},
            "rows": rows}


def tung_setup(P, torch, n, prec):
    """tungsten_scalability.cpp:95-135: spacing 1, n0 = -0.4, T = 0.5, dt = 0.01,
    psi = -0.4 + 0.1 sin(2 pi i / N^3)."""
    fft = P.HipFft(n, n, n, precision=prec)
    st = P.TungstenStepper(fft, (1.0, 1.0, 1.0), P.TungstenParams(n0=-0.4, T=0.5))
    st.initialize(0.01)
    i = torch.arange(n ** 3, dtype=torch.float64, device="cuda")
    psi = -0.4 + 0.1 * torch.sin(2.0 * math.pi * i / float(n ** 3))
    del i
    if prec == P.F32:
        psi = psi.float()
    torch.cuda.synchronize()
    return fft, st, psi


def tung_cpu_baseline(n, nsteps):
    import numpy as np
    from oracle import oracle as O

    import openpfc_amd.tungsten as T
    prm = T.TungstenParams(n0=-0.4, T=0.5)
    N = n ** 3
    a = (-0.4 + 0.1 * np.sin(2.0 * math.pi * np.arange(N, dtype=np.float64) / N))
    a = a.reshape(n, n, n)
    fMF, expL, nW = O.tungsten_operators((n, n, n), (1.0, 1.0, 1.0), prm.operator_params(),
                                         0.01)
    t0 = time.perf_counter()
    O.tungsten_step(a, fMF, expL, nW, prm.nonlinear_params(), nsteps)
    return nsteps / (time.perf_counter() - t0), O.num_threads()


def tungsten_main(args, P, torch):
    n = args.n
    log(f"tungsten fp64 {n}^3: setup")
    fft, st, psi = tung_setup(P, torch, n, P.F64)
    psi0 = psi.clone() if not args.no_fp32 else None
    chosen = choose_variants(P, fft, st, psi, args, "tungsten fp64", TUNG_STEP_ROLES)
    if args.save_variants:
        with open(args.save_variants, "w") as fh:
            json.dump(chosen, fh, indent=1)
    ms, wall = time_steps(torch, st, psi, args.steps, args.warmup)
    kms = st.step_timed(psi, 3)
    finite = bool(torch.isfinite(psi).all().item())
    del st, fft
    torch.cuda.empty_cache()
    pb = tung_pass_bytes(n, 8)
    dom = max(range(9), key=lambda i: kms[i])
    dname = TUNG_PASS_NAMES[dom]
    achieved = pb[dname] / (kms[dom] * 1e-3)
    step_bytes = sum(pb.values())
    layout = chosen.get("step_layout", "natural")
    kern, role = TUNG_PASS_KERNELS[layout][dom]
    kkey = variant_key(kern, "f64", n, chosen[role],
                       direction=(-1 if dom in (1, 5) else 1),
                       mode={2: 1, 6: 2}.get(dom, 0))
    traffic = None
    tpath = args.traffic or os.path.join(ROOT, "profiles", f"traffic_tungsten_step_{n}_f64.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            tk = json.load(fh).get("kernels", {})
        if kkey in tk:
            traffic = (tk[kkey][0] + tk[kkey][1]) * 1e9
    out = {
        "metric": TUNG_METRIC, "value": 1000.0 / ms, "unit": "steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: -0.4 + 0.1 sin(2 pi i/N^3) (tungsten_scalability workload)",
        "config": {"workload": f"Tungsten mean-field ETD1 step {n}^3 FP64", "grid": [n, n, n],
                   "spacing": 1.0, "dt": 0.01, "n0": -0.4, "T": 0.5},
        "ns_per_dof_step": ms * 1e6 / n ** 3,
        "step_hbm_frac": step_bytes / (ms * 1e-3) / HBM_PEAK,
        "step_alg_bytes": step_bytes, "wall_s_timed": wall,
        "kernels_ms": dict(zip(TUNG_PASS_NAMES, kms)),
        "roofline": {"bound": "hbm", "kernel": kkey, "pass": dname,
                     "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "alg_bytes_per_launch": pb[dname]},
        "variants": chosen, "checks": {"finite": finite}}
    log(f"tungsten fp64 done: {ms:.3f} ms/step")
    if not args.no_fp32:
        fft = P.HipFft(n, n, n, precision=P.F32)
        st = P.TungstenStepper(fft, (1.0, 1.0, 1.0), P.TungstenParams(n0=-0.4, T=0.5))
        st.initialize(0.01)
        d = psi0.float()
        del psi0
        chosen32 = choose_variants(P, fft, st, d, args, "tungsten fp32", TUNG_STEP_ROLES)
        ms32, _ = time_steps(torch, st, d, args.steps, args.warmup)
        k32 = st.step_timed(d, 3)
        diff = d.double() - psi
        l2 = (torch.linalg.vector_norm(diff) / torch.linalg.vector_norm(psi)).item()
        del d, diff, st, fft
        torch.cuda.empty_cache()
        out["fp32"] = {"mode": "TungstenHIP<float>: fp32 field, transforms and multipliers",
                       "steps_per_s": 1000.0 / ms32, "ms_per_step": ms32,
                       "speedup_vs_fp64": ms / ms32,
                       "step_hbm_frac": sum(tung_pass_bytes(n, 4).values()) / (ms32 * 1e-3)
                       / HBM_PEAK,
                       "kernels_ms": dict(zip(TUNG_PASS_NAMES, k32)), "variants": chosen32,
                       "rel_l2_vs_fp64_after": {"steps": args.warmup + args.steps + 3,
                                                "rel_l2": l2}}
    del psi
    torch.cuda.empty_cache()
    if not args.no_cpu_baseline:
        log("tungsten cpu baseline (oracle) start")
        sps, threads = tung_cpu_baseline(n, args.cpu_steps)
        out["cpu_baseline"] = {
            "value": sps, "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": f"{args.cpu_steps} step(s) of the same {n}^3 FP64 workload, oracle C "
                      f"restatement (own Stockham FFT, OpenMP {threads} threads)"}
    if args.calibrate_copy:
        out["copy_calibration"] = copy_calibration(torch)
    print(json.dumps(out), flush=True)
    return 0


def adjoint_run(P, torch, n=256, nsteps=200, every=20):
    """256^3 x 200-step adjoint run: gradients of J = ||psi_T - psi_target||^2 with respect
    to psi0 and r through the libtorch autograd::Function (adjoint steps,
    checkpoints every `every` steps), psi_target from the seeded run at
    r = -0.30; checked against FP64 central differences (one random
    direction for psi0, h = 1e-6 * |psi0|_max; h = 1e-6 for r)."""
    import numpy as np

    from openpfc_amd.autograd import pfc_trajectory
    dx, dt = math.pi / 4, 0.5
    fft = P.HipFft(n, n, n)
    psi = torch.empty(n ** 3, dtype=torch.float64, device="cuda")
    P.ic_constant(fft, psi, -0.2)
    P.ic_add_noise(fft, psi, 1234, -0.01, 0.01)
    psi0 = psi.view(n, n, n).clone()
    del fft
    sp = (dx, dx, dx)
    with torch.no_grad():
        target = pfc_trajectory(psi0, torch.tensor(-0.30, dtype=torch.float64), nsteps, dt, sp,
                                every)
    x0 = psi0.clone().requires_grad_(True)
    r = torch.tensor(-0.25, dtype=torch.float64, requires_grad=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    psiT = pfc_trajectory(x0, r, nsteps, dt, sp, every)
    J = ((psiT - target) ** 2).sum()
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    J.backward()
    torch.cuda.synchronize()
    t_bwd = time.perf_counter() - t0
    g, gr = x0.grad, float(r.grad.item())

    def loss(p0, rv):
        with torch.no_grad():
            pt = pfc_trajectory(p0, torch.tensor(rv, dtype=torch.float64), nsteps, dt, sp, every)
            return float(((pt - target) ** 2).sum().item())

    rng = np.random.default_rng(11)
    v = torch.from_numpy(rng.standard_normal((n, n, n))).cuda()
    h = 1e-6 * float(psi0.abs().max().item())
    fd = (loss(psi0 + h * v, -0.25) - loss(psi0 - h * v, -0.25)) / (2 * h)
    ad = float((g * v).sum().item())
    hr = 1e-6
    fdr = (loss(psi0, -0.25 + hr) - loss(psi0, -0.25 - hr)) / (2 * hr)
    # FP32-mixed: FP32 deviation field and FP32 adjoint, FP64 mean and r sum
    xm = psi0.clone().requires_grad_(True)
    rm = torch.tensor(-0.25, dtype=torch.float64, requires_grad=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    psiTm = pfc_trajectory(xm, rm, nsteps, dt, sp, every, mixed=True)
    Jm = ((psiTm - target) ** 2).sum()
    torch.cuda.synchronize()
    tm_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    Jm.backward()
    torch.cuda.synchronize()
    tm_bwd = time.perf_counter() - t0
    adm = float((xm.grad * v).sum().item())
    grm = float(rm.grad.item())
    mixed = {"psi0_direction_rel_err": abs(adm - fd) / abs(fd),
             "r_rel_err": abs(grm - fdr) / abs(fdr),
             "psiT_rel_l2_vs_fp64": float(((psiTm.detach() - psiT.detach()).norm() /
                                           psiT.detach().norm()).item()),
             "seconds": {"forward_with_checkpoints": tm_fwd, "backward": tm_bwd}}
    return {"metric": "adjoint gradients vs FP64 central differences", "grid": [n, n, n],
            "steps": nsteps, "checkpoint_every": every, "precision": "f64",
            "loss": float(J.item()),
            "psi0_direction": {"adjoint": ad, "central_difference": fd,
                               "rel_err": abs(ad - fd) / abs(fd), "h": h},
            "r": {"adjoint": gr, "central_difference": fdr, "rel_err": abs(gr - fdr) / abs(fdr),
                  "h": hr},
            "seconds": {"forward_with_checkpoints": t_fwd, "backward": t_bwd},
            "fp32_mixed": mixed}


ENS_SIZES = [32, 48, 64, 96, 128, 192, 256]
ENS_METRIC = "ensemble DOF-steps/s (64 uneven PFC runs, FP32-mixed, one graph launch per step)"


def ensemble_draw(nruns=64, seed=4):
    """Ensemble draw: sizes from ENS_SIZES, r in U(-0.35, -0.15) and the
    mean in U(-0.30, -0.15), all from one seeded numpy generator."""
    import numpy as np
    rng = np.random.default_rng(seed)
    sizes = [int(v) for v in rng.choice(ENS_SIZES, size=nruns)]
    eps = [float(v) for v in rng.uniform(-0.35, -0.15, nruns)]
    means = [float(v) for v in rng.uniform(-0.30, -0.15, nruns)]
    return sizes, eps, means


def ensemble_bytes(sizes, esize):
    """Algorithmic bytes of one ensemble step: 2R + 12C per run."""
    return sum(2 * esize * n ** 3 + 12 * 2 * esize * (n // 2 + 1) * n * n for n in sizes)


def ensemble_run(P, torch, steps, warmup, nruns=64, seed=4, acc_steps=100, trace_steps=0):
    """Ensemble of nruns independent plain-PFC runs of uneven cubic
    sizes (dx = pi/4, dt = 0.5), psi0 = mean + U(-0.01, 0.01) (torch's
    device generator, seed 1000*seed + run), FP32-mixed (each field holds
    psi - <psi> in FP32; the mean is kept in FP64).  Autotuned per size
    group (untimed), then `steps` ensemble steps timed, each one hipGraph
    launch; the same steps enqueued without the graph are timed beside it.
    The FP32-mixed ensemble is compared with the FP64 ensemble (plain step)
    from the same psi0 after acc_steps steps."""
    dx, dt = math.pi / 4, 0.5
    sizes, eps, means = ensemble_draw(nruns, seed)
    runs, devs = [], []
    for i, n in enumerate(sizes):
        g = torch.Generator(device="cuda")
        g.manual_seed(1000 * seed + i)
        noise = torch.rand(n ** 3, dtype=torch.float64, device="cuda", generator=g) * 0.02 - 0.01
        m = float(noise.mean().item())
        devs.append(noise - m)
        runs.append(P.Run.cubic(n, dx, eps[i], dt, means[i] + m))
    dof = sum(n ** 3 for n in sizes)
    log(f"ensemble: {nruns} runs, {dof:.3e} DOF")
    ens = P.Ensemble(runs, precision=P.F32, mean_split=True)
    for i in range(nruns):
        ens.field(i).view(-1).copy_(devs[i].float())
    d0 = ens.buffer.clone()
    t0 = time.perf_counter()
    ens.autotune(reps=2)
    torch.cuda.synchronize()
    tune_s = time.perf_counter() - t0
    log(f"ensemble: autotune {tune_s:.1f} s")
    if trace_steps:
        # for a kernel trace: the GPU idles 0.2 s, then trace_steps graph
        # steps run back to back (the window profile_summary --gap-window
        # finds after the last long gap)
        ens.step(1, graph=True)
        torch.cuda.synchronize()
        time.sleep(0.2)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        ens.step(trace_steps, graph=True)
        b.record()
        torch.cuda.synchronize()
        ms = a.elapsed_time(b) / trace_steps
        ens.close()
        return {"trace_steps": trace_steps, "graph_ms_per_step": ms, "dof": dof}

    def timed(graph):
        ens.buffer.copy_(d0)
        ens.step(warmup, graph=graph)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        ens.step(steps, graph=graph)
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / steps

    eager_ms = timed(False)
    graph_ms = timed(True)
    log(f"ensemble: graph {graph_ms:.3f} ms/step, eager {eager_ms:.3f} ms/step")
    alg = ensemble_bytes(sizes, 4)
    # per-group pass times (eager, HIP events)
    groups = ens.groups()
    for g in groups:
        plan, batch, off = ctypes.c_void_p(), ctypes.c_int(), ctypes.c_size_t()
        gi = groups.index(g)
        P.check(ens._lib.pfc_hip_ensemble_group(ens._h, gi, ctypes.byref(plan),
                                                ctypes.byref(batch), ctypes.byref(off)))
        out = (ctypes.c_double * 5)()
        ptr = ctypes.c_void_p(ens.buffer.data_ptr() + 4 * off.value)
        P.check(ens._lib.pfc_hip_pfc_step_timed(
            plan, ptr, 5, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream), out, 5))
        # the transposed layout steps a group's fields one after another and
        # times the passes of field 0; the natural one times all fields
        per_field = g["variants"].get("step_layout") == "transposed"
        step = sum(out) * (batch.value if per_field else 1)
        g["pass_ms"] = [round(v, 4) for v in out]
        g["pass_ms_of"] = "field 0" if per_field else "all fields"
        g["step_ms"] = round(step, 4)
        n = g["ext"][0]
        g["hbm_frac"] = ensemble_bytes([n] * batch.value, 4) / (step * 1e-3) / HBM_PEAK
    # FP32-mixed against the FP64 ensemble from the same psi0
    ens.buffer.copy_(d0)
    ref = P.Ensemble(runs, precision=P.F64, mean_split=False)
    for i in range(nruns):
        ref.field(i).view(-1).copy_(devs[i] + runs[i].mean)
    ens.step(acc_steps)
    ref.step(acc_steps)
    torch.cuda.synchronize()
    l2, mx, drift = [], [], []
    for i in range(nruns):
        a = ref.field(i).view(-1)
        d = ens.field(i).view(-1).double()
        diff = d + runs[i].mean - a
        l2.append(float(diff.norm() / a.norm()))
        mx.append(float(diff.abs().max() / a.abs().max()))
        drift.append(abs(float(d.mean())))
    ref.close()
    ens.close()
    return {
        "metric": ENS_METRIC, "value": dof / (graph_ms * 1e-3), "unit": "DOF-steps/s",
        "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": graph_ms,
        "higher_is_better": True, "dtype": "f32 (FP64 mean)", "data": "synthetic",
        "config": {"workload": f"ensemble of {nruns} plain PFC runs, sizes from {ENS_SIZES}",
                   "seed": seed, "dx": dx, "dt": dt, "dof": dof,
                   "sizes": sizes, "r": [round(v, 6) for v in eps],
                   "mean": [round(v, 6) for v in means]},
        "eager_ms_per_step": eager_ms, "graph_speedup": eager_ms / graph_ms,
        "autotune_s": tune_s,
        "roofline": {"bound": "hbm", "achieved": alg / (graph_ms * 1e-3) / 1e9,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": alg / (graph_ms * 1e-3) / HBM_PEAK, "alg_bytes_per_step": alg},
        "groups": groups,
        "fp32_mixed_vs_fp64": {"steps": acc_steps, "rel_l2_max": max(l2),
                               "rel_max_max": max(mx), "mean_drift_max": max(drift)},
    }


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--model", choices=["pfc", "tungsten"], default="pfc")
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--no-cpp", action="store_true",
                    help="skip the C++ App leg (openpfc_amd_app at the same size)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the (untimed) launch-variant autotuner")
    ap.add_argument("--variants", help="JSON file {role: variant name} to force instead of "
                                       "autotuning (e.g. a previous run's choice)")
    ap.add_argument("--save-variants", help="write the chosen variants to this JSON file")
    ap.add_argument("--variants-fp32", help="as --variants, for the FP32-mixed leg")
    ap.add_argument("--save-variants-fp32", help="as --save-variants, for the FP32-mixed leg")
    ap.add_argument("--variants-fp32xp", help="as --variants, for the FP32-mixed leg with "
                                              "FP64 butterflies")
    ap.add_argument("--save-variants-fp32xp", help="as --save-variants, for that leg")
    ap.add_argument("--traffic", default=None,
                    help="traffic table from openpfc_amd.profile_summary (default: "
                         "profiles/traffic_pfc_step_<n>_f64.json when present)")
    ap.add_argument("--accuracy", type=int, default=0,
                    help="run FP64 and FP32-mixed in lockstep for this many steps and "
                         "report the FP32-mixed error every 100 steps (no timing)")
    ap.add_argument("--adjoint", action="store_true",
                    help="adjoint gradients at 256^3, 200 steps, against "
                         "FP64 central differences (no timing line)")
    ap.add_argument("--ensemble", action="store_true",
                    help="64 uneven-size PFC runs (FP32-mixed) advanced "
                         "together, one hipGraph launch per step; --steps K are timed")
    ap.add_argument("--ens-trace-steps", type=int, default=0, metavar="K",
                    help="with --ensemble: after the autotune, idle 0.2 s, run K graph "
                         "steps and stop (a kernel trace's window for the idle share)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also time a device-to-device copy of 2 x 4 GiB (the plain "
                         "copy kernel whose PMC byte counts scale the traffic figure)")
    args = ap.parse_args(argv)

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: the step runs on one GPU (it is not sharded; no collectives); "
              "use --gpus 1", file=sys.stderr)
        return 2

    import torch

    import openpfc_amd as P

    torch.cuda.set_device(0)
    if args.model == "tungsten":
        return tungsten_main(args, P, torch)
    if args.adjoint:
        print(json.dumps(adjoint_run(P, torch)), flush=True)
        return 0
    if args.ensemble:
        print(json.dumps(ensemble_run(P, torch, args.steps, args.warmup,
                                      trace_steps=args.ens_trace_steps)), flush=True)
        return 0
    n, dx, dt, r = args.n, math.pi / 4, 0.5, -0.25
    if args.accuracy:
        print(json.dumps(accuracy_run(P, torch, n, dx, dt, r, args.accuracy)), flush=True)
        return 0

    # ---- FP64 (value) -----------------------------------------------------
    log(f"fp64 {n}^3: setup")
    fft, st, psi = setup(P, torch, n, P.F64, dx, dt, r)
    psi0 = None if args.no_fp32 else psi.clone()
    chosen64 = choose_variants(P, fft, st, psi, args, "fp64")
    if args.save_variants:
        with open(args.save_variants, "w") as fh:
            json.dump(chosen64, fh, indent=1)
    entry64 = P.cache_entry(fft)
    ms64, wall64 = time_steps(torch, st, psi, args.steps, args.warmup)
    kms = st.step_timed(psi, 3)
    kf = st.step_timed_fused(psi, max(args.steps, 2))
    finite = bool(torch.isfinite(psi).all().item())
    mean64 = psi.mean().item()
    psi64 = psi
    del st, fft
    torch.cuda.empty_cache()

    pb = pass_bytes(n, 8)
    step_bytes = sum(pb.values())
    layout = chosen64.get("step_layout", "natural")
    fused = kf[5] > 0
    if fused:
        # the timed K-step call fuses step boundaries: its steady-state
        # passes are y forward, z, y inverse and the fused boundary
        fb = fused_pass_bytes(n, 8)
        fk = dict(zip(FUSED_NAMES, kf))
        dname = max(fb, key=lambda k: fk[k])
        achieved = fb[dname] / (fk[dname] * 1e-3)
        kern, role = {"y_fwd": ("k_ty_fwd", "t_y_fwd"), "z_fused": ("k_tz", "t_z"),
                      "y_inv": ("k_ty_inv", "t_y_inv"),
                      "x_boundary_fused": ("k_tx_fuse", "t_x_fuse")}[dname]
        alg_launch = fb[dname]
        run_bytes = sum(fb.values()) + 2 * 8 * n ** 3 / args.steps
    else:
        dom = max(range(5), key=lambda i: kms[i])
        dname = PASS_NAMES[dom]
        achieved = pb[dname] / (kms[dom] * 1e-3)
        kern, role = PASS_KERNELS[layout][dom]
        alg_launch = pb[dname]
        run_bytes = step_bytes
    kkey = variant_key(kern, "f64", n, chosen64[role],
                       direction=(-1 if dname == "y_fwd" else 1))
    traffic = None
    tpath = args.traffic or os.path.join(ROOT, "profiles", f"traffic_pfc_step_{n}_f64.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            tk = json.load(fh).get("kernels", {})
        if kkey in tk:
            traffic = (tk[kkey][0] + tk[kkey][1]) * 1e9

    out = {
        "metric": METRIC,
        "value": 1000.0 / ms64,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms64,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: -0.2 + U(-0.01,0.01) noise (splitmix64 seed 1234) + bcc one-mode seed",
        "config": {"workload": f"plain PFC ETD1 step {n}^3 FP64", "grid": [n, n, n],
                   "dx": dx, "dt": dt, "r": r, "psi_mean": -0.2, "noise": 0.01,
                   "seed_radius": 64 * dx, "seed_amplitude": SEED_AMP},
        "ns_per_dof_step": ms64 * 1e6 / n ** 3,
        "step_hbm_frac": step_bytes / (ms64 * 1e-3) / HBM_PEAK,
        "step_alg_bytes": step_bytes,
        "run_bytes_per_step": run_bytes,
        "run_hbm_frac": run_bytes / (ms64 * 1e-3) / HBM_PEAK,
        "wall_s_timed": wall64,
        "kernels_ms": dict(zip(PASS_NAMES, kms)),
        "fused_run_ms": dict(zip(FUSED_NAMES, kf)) if fused else None,
        "roofline": {"bound": "hbm", "kernel": kkey, "pass": dname,
                     "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "alg_bytes_per_launch": alg_launch},
        "variants": chosen64,
        "checks": {"finite": finite, "mean_after": mean64},
    }

    # ---- FP32-mixed (same workload, same step sequence): FP32 butterflies,
    # and FP64 butterflies on the same FP32 fields ---------------------------
    log(f"fp64 done: {ms64:.3f} ms/step")
    if not args.no_fp32:
        legs = [("fp32_mixed", 0, args.variants_fp32, args.save_variants_fp32,
                 "fp32 field and transforms; fp64 conserved mean (mean-split step, k=0 pinned)"),
                ("fp32_mixed_f64_butterflies", P.FLAG_F64_BUTTERFLIES, args.variants_fp32xp,
                 args.save_variants_fp32xp,
                 "fp32 field and spectra, fp64 butterflies (twiddle products and in-register "
                 "DFTs), one fp32 rounding per radix pass; fp64 conserved mean")]
        mean0 = float(psi0.double().mean().item())
        for key, flags, vfile, sfile, mode in legs:
            log(f"{key}: setup")
            fft = P.HipFft(n, n, n, precision=P.F32, flags=flags)
            st = P.PfcStepper(fft, (dx, dx, dx), r)
            st.initialize(dt)
            d = (psi0 - mean0).float()
            st.set_mean(mean0)
            a32 = argparse.Namespace(**vars(args))
            a32.variants = vfile
            chosen32 = choose_variants(P, fft, st, d, a32, key)
            if sfile:
                with open(sfile, "w") as fh:
                    json.dump(chosen32, fh, indent=1)
            ms32, _ = time_steps(torch, st, d, args.steps, args.warmup)
            k32 = st.step_timed(d, 3)
            k32f = st.step_timed_fused(d, max(args.steps, 2))
            got = d.double() + mean0
            diff = got - psi64
            l2 = (torch.linalg.vector_norm(diff) / torch.linalg.vector_norm(psi64)).item()
            mx = (diff.abs().max() / psi64.abs().max()).item()
            drift = abs(got.mean().item() - mean0)
            del d, got, diff, st, fft
            torch.cuda.empty_cache()
            log(f"{key}: {ms32:.3f} ms/step")
            out[key] = {
                "mode": mode,
                "steps_per_s": 1000.0 / ms32, "ms_per_step": ms32,
                "ns_per_dof_step": ms32 * 1e6 / n ** 3,
                "speedup_vs_fp64": ms64 / ms32,
                "step_hbm_frac": sum(pass_bytes(n, 4).values()) / (ms32 * 1e-3) / HBM_PEAK,
                "run_hbm_frac": ((sum(fused_pass_bytes(n, 4).values())
                                  + 2 * 4 * n ** 3 / args.steps)
                                 if k32f[5] > 0 else sum(pass_bytes(n, 4).values()))
                                / (ms32 * 1e-3) / HBM_PEAK,
                "kernels_ms": dict(zip(PASS_NAMES, k32)),
                "fused_run_ms": dict(zip(FUSED_NAMES, k32f)) if k32f[5] > 0 else None,
                "variants": chosen32,
                "error_vs_fp64_after": {"steps": args.warmup + 2 * args.steps + 3
                                        if args.steps >= 2 else args.warmup + args.steps + 5,
                                        "rel_l2": l2, "rel_max": mx, "mean_drift": drift,
                                        "note": "a short run; the 1000-step error is "
                                                "bench.py --accuracy 1000"},
                "accuracy_bar": {"rel_l2_after_1000_steps": 1e-4,
                                 "measured_by": "bench.py --accuracy 1000"}}
        del psi0
    del psi64
    torch.cuda.empty_cache()

    # ---- streaming-copy ceiling on this GPU (arrays the size of one half
    # spectrum) ---------------------------------------------------------------
    sc = stream_copy(P, torch, 16 * (n // 2 + 1) * n * n)
    out["stream_copy"] = sc
    out["roofline"]["frac_of_stream_copy"] = out["roofline"]["achieved"] / sc["GBps"]
    out["step_frac_of_stream_copy"] = step_bytes / (ms64 * 1e-3) / 1e9 / sc["GBps"]

    # ---- the C++ drop-in path (App over PFCModelHIP), same variants -------
    if not args.no_cpp:
        cpp = cpp_app_leg(P, torch, n, dx, dt, r, args.steps, entry64)
        cpp["vs_value"] = ms64 / cpp["ms_per_step"]
        out["cpp_app"] = cpp

    # ---- CPU baseline (oracle, rank 0, N=1) --------------------------------
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle) start")
        out["cpu_baseline"] = cpu_baseline(n, dx, dt, r, args.cpu_steps)

    if args.calibrate_copy:
        out["copy_calibration"] = copy_calibration(torch)

    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
