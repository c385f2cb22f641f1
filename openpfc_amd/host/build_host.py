"""Builds the C++ host-layer programs (hipcc, linked to
openpfc_amd/libopenpfc_hip.so with an $ORIGIN rpath): the test driver
openpfc_amd/host/test_host and the settings-file application
openpfc_amd/host/openpfc_amd_app."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
OUT = os.path.join(HERE, "test_host")
APP = os.path.join(HERE, "openpfc_amd_app")


def _compile(src: str, out: str, verbose: bool) -> str:
    from openpfc_amd.build import hipcc
    cmd = [hipcc(), "-std=c++20", "-O2", "-Wall", "-Wextra",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "include"),
           src, "-L", PKG, "-lopenpfc_hip", "-Wl,-rpath,$ORIGIN/..", "-o", out + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("host build failed:\n" + r.stdout + r.stderr)
    if verbose and (r.stdout or r.stderr):
        print(r.stdout + r.stderr)
    os.replace(out + ".tmp", out)
    return out


def build(verbose: bool = False) -> str:
    from concurrent.futures import ThreadPoolExecutor
    jobs = [(os.path.join(HERE, "tests", "test_host.cpp"), OUT),
            (os.path.join(HERE, "apps", "openpfc_amd_app.cpp"), APP)]
    with ThreadPoolExecutor(max_workers=2) as ex:
        list(ex.map(lambda j: _compile(j[0], j[1], verbose), jobs))
    return OUT


if __name__ in ("__main__", "__build__"):
    import sys
    sys.path.insert(0, ROOT)
    build(verbose=__name__ == "__main__")
