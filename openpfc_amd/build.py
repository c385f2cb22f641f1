"""In-tree build of the HIP library (gfx950) and the C++ host tests.

    python -m openpfc_amd.build          # builds openpfc_amd/libopenpfc_hip.so

All HIP code is compiled with the packed-FP32 instructions disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libopenpfc_hip.so")
ARCH = os.environ.get("PFC_HIP_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++20", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-pass-failed",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd, quiet=True):
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in l
                    and "-Wunknown-warning" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed ({' '.join(cmd)}):\n{out}")
    if out.strip() and not quiet:
        print(out)


SOURCES = ["pfc_hip.hip", "variants_common.hip", "variants_f64.hip", "variants_f32.hip",
           "variants_t_f64.hip", "variants_t_f32.hip", "variants_t_xp.hip", "field_io.hip",
           "ensemble.hip"]
OBJDIR = os.path.join(HERE, "_build")


def _stale(obj: str, src: str) -> bool:
    """Whether obj is older than its source or any header it may include."""
    if not os.path.exists(obj):
        return True
    deps = [src] + [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hpp")]
    deps.append(os.path.join(ROOT, "include", "pfc_hip.h"))
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) > t for d in deps)


def build_library(verbose: bool = False, incremental: bool = False) -> str:
    """Compile every HIP source (in parallel) and link openpfc_amd/libopenpfc_hip.so.
    incremental=True recompiles only objects older than their sources or
    headers (a development aid; the default compiles everything)."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(OBJDIR, exist_ok=True)
    objs = []
    jobs = []
    for src in SOURCES:
        obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        if incremental and not _stale(obj, os.path.join(CSRC, src)):
            continue
        jobs.append([hipcc(), *HIP_FLAGS, "-c", "-o", obj, os.path.join(CSRC, src)])
    with ThreadPoolExecutor(max_workers=max(1, min(len(jobs), os.cpu_count() or 4))) as ex:
        list(ex.map(lambda c: _run(c, quiet=not verbose), jobs))
    tmp = LIB + ".tmp"
    _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs],
         quiet=not verbose)
    os.replace(tmp, LIB)
    return LIB


TORCH_EXT = os.path.join(HERE, "openpfc_amd_torch.so")


def build_torch_extension(verbose: bool = False) -> str:
    """The libtorch autograd wrapper (csrc/torch_pfc.cpp): host C++ against
    torch's headers and the C ABI, linked to libopenpfc_hip.so next to it
    ($ORIGIN rpath).  Compiled directly with g++ (no JIT cache)."""
    import sysconfig

    import torch
    import torch.utils.cpp_extension as ce
    cxx = shutil.which("g++") or "g++"
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    cmd = [cxx, "-O2", "-std=c++17", "-shared", "-fPIC", "-w",
           f"-D_GLIBCXX_USE_CXX11_ABI={abi}", "-DTORCH_EXTENSION_NAME=openpfc_amd_torch",
           "-DTORCH_API_INCLUDE_EXTENSION_H",
           # torch's current HIP stream (c10/hip/HIPStream.h) needs the HIP headers
           "-D__HIP_PLATFORM_AMD__=1", "-DUSE_ROCM=1"]
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    for d in ce.include_paths() + [sysconfig.get_paths()["include"], os.path.join(rocm, "include")]:
        cmd += ["-isystem", d]
    cmd += [os.path.join(CSRC, "torch_pfc.cpp")]
    for d in ce.library_paths():
        cmd += [f"-L{d}", f"-Wl,-rpath,{d}"]
    cmd += ["-lc10", "-lc10_hip", "-ltorch", "-ltorch_cpu", "-ltorch_python", f"-L{HERE}",
            "-lopenpfc_hip",
            "-Wl,-rpath,$ORIGIN", "-o", TORCH_EXT + ".tmp"]
    _run(cmd, quiet=not verbose)
    os.replace(TORCH_EXT + ".tmp", TORCH_EXT)
    return TORCH_EXT


if __name__ == "__main__":
    print(build_library(verbose="-v" in sys.argv, incremental="-i" in sys.argv))
    print(build_torch_extension(verbose="-v" in sys.argv))
