"""Builds the C++ host-layer test driver openpfc_amd/host/test_host
(hipcc, links openpfc_amd/libopenpfc_hip.so with an $ORIGIN rpath)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
OUT = os.path.join(HERE, "test_host")


def build(verbose: bool = False) -> str:
    from openpfc_amd.build import hipcc
    cmd = [hipcc(), "-std=c++20", "-O2", "-Wall", "-Wextra",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "include"),
           os.path.join(HERE, "tests", "test_host.cpp"),
           "-L", PKG, "-lopenpfc_hip", "-Wl,-rpath,$ORIGIN/..", "-o", OUT + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("host build failed:\n" + r.stdout + r.stderr)
    if verbose and (r.stdout or r.stderr):
        print(r.stdout + r.stderr)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ in ("__main__", "__build__"):
    import sys
    sys.path.insert(0, ROOT)
    build(verbose=__name__ == "__main__")
