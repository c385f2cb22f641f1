"""Settings files of the App layer (C++, openpfc_amd/host/include/openpfc_amd/
{json,toml,app}.hpp), checked without a device through the host test driver:

* the JSON reader against Python's json module and the TOML reader against
  tomli (independent parsers): the canonical JSON the C++ side prints must
  load to the same document; integers stay integers (the reference checks
  is_number_integer() for grid sizes);
* the cases of tests/unit/frontend/ui/test_toml_to_json.cpp (basic types,
  arrays, nested tables, arrays of tables);
* World/Time wiring (from_json_world_time.hpp:42-284): centred origin, the
  validation errors, the field-modifier factory and the backend check;
* the oracle's FixedBC against its closed form (S = 1/2 at xpos, rho_high
  left of the band, rho_low right of it, untouched outside).
"""
import json
import math
import os
import subprocess

import numpy as np
import pytest
import tomli

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "openpfc_amd", "host", "test_host")

TUNGSTEN_PARAMS = {"n0": -0.10, "alpha": 0.50, "n_sol": -0.047, "n_vap": -0.464, "T": 3300.0,
                   "T0": 156000.0, "Bx": 0.8582, "alpha_farTol": 0.001, "alpha_highOrd": 4,
                   "lambda": 0.22, "stabP": 0.2, "shift_u": 0.3341, "shift_s": 0.1898,
                   "p2": 1.0, "p3": -0.5, "p4": 0.333333333, "q20": -0.0037, "q21": 1.0,
                   "q30": -12.4567, "q31": 20.0, "q40": 45.0}

TOML_CONFIG = """
# Tungsten, one seed
[model]
name = "tungsten"

[model.params]
n0 = -0.10
alpha = 0.50
n_sol = -0.047
n_vap = -0.464
T = 3300.0
T0 = 156000.0
Bx = 0.8582
alpha_farTol = 0.001
alpha_highOrd = 4
lambda = 0.22
stabP = 0.2
shift_u = 0.3341
shift_s = 0.1898
p2 = 1.0
p3 = -0.5
p4 = 0.333333333
q20 = -0.0037
q21 = 1.0
q30 = -12.4567
q31 = 20.0
q40 = 45.0

[domain]
Lx = 32
Ly = 32
Lz = 32
dx = 1.1107207345395915
dy = 1.1107207345395915
dz = 1.1107207345395915
origin = "center"

[timestepping]
t0 = 0.0
t1 = 10.0
dt = 1.0
saveat = 1.0

[[fields]]
name = "psi"
data = "out/psi_%d.bin"

[[initial_conditions]]
target = "psi"
type = "constant"
n0 = -0.4

[[initial_conditions]]
target = "psi"
type = "single_seed"
amp_eq = 0.215936
rho_seed = -0.047

[[boundary_conditions]]
target = "psi"
type = "fixed"
rho_low = -0.464
rho_high = -0.100
"""


def _run(*args):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} missing: run __graft_entry__.build()")
    return subprocess.run([BIN, *args], capture_output=True, text=True, timeout=60)


def _parse(tmp_path, text, ext):
    p = tmp_path / f"doc{ext}"
    p.write_text(text)
    r = _run("parse", str(p))
    assert r.returncode == 0, r.stderr
    return r.stdout


def _same(a, b):
    """Equality that keeps the int/float distinction and compares NaN as NaN."""
    if isinstance(a, float) and isinstance(b, float) and math.isnan(a) and math.isnan(b):
        return True
    if type(a) is not type(b):
        return False
    if isinstance(a, dict):
        return a.keys() == b.keys() and all(_same(a[k], b[k]) for k in a)
    if isinstance(a, list):
        return len(a) == len(b) and all(_same(x, y) for x, y in zip(a, b))
    return a == b


JSON_DOCS = [
    '{"a": 1, "b": -2.5e-3, "c": [1, 2.0, "x", true, false, null], "d": {"e": {}}}',
    '{"s": "tab\\tquote\\"back\\\\slash\\/ \\u00e9 \\ud83d\\ude00", "big": 123456789012}',
    '[{"Lx": 256}, [], [[1]], 1E+2, -0.0, 0.1]',
    json.dumps({"model": {"name": "tungsten", "params": TUNGSTEN_PARAMS}}),
]


@pytest.mark.parametrize("doc", JSON_DOCS)
def test_json_reader_matches_python_json(tmp_path, doc):
    out = _parse(tmp_path, doc, ".json")
    assert _same(json.loads(out), json.loads(doc))


@pytest.mark.parametrize("bad", ['{"a": 1,}', '{"a" 1}', '[1, 2', '{"a": 1} x', '"\\q"',
                                 '{"a": 1, "a": 2}'])
def test_json_reader_rejects_malformed(tmp_path, bad):
    p = tmp_path / "bad.json"
    p.write_text(bad)
    r = _run("parse", str(p))
    assert r.returncode != 0 and "JSON parse error" in r.stderr


TOML_DOCS = [
    # test_toml_to_json.cpp: basic types
    'name = "test"\npath = "/some/path"\ncount = 42\nnegative = -10\nzero = 0\n'
    'pi = 3.14159\nneg = -2.5\nscientific = 1.23e-4\nenabled = true\ndisabled = false\n',
    # arrays (and a multi-line one with a trailing comma and comments)
    'numbers = [1, 2, 3, 4, 5]\nnames = ["alice", "bob", "charlie"]\nvalues = [1.1, 2.2, 3.3]\n'
    'empty = []\nnested = [[1, 2], ["a"]]\nml = [\n  1, # one\n  2,\n]\n',
    # nested tables and dotted keys
    '[model]\nname = "tungsten"\n[model.params]\nn0 = -0.1\nT = 3300.0\n'
    '[a.b.c]\nvalue = 123\n[x]\ny.z = 1\n"quoted key" = \'literal \\n\'\n',
    # arrays of tables, inline tables, integer forms
    '[[fields]]\nname = "psi"\ndata = "/path/to/psi.bin"\n[[fields]]\nname = "psiMF"\n'
    'data = "/path/to/psimf.bin"\n[opts]\ninline = { a = 1, b = { c = "d" } }\n'
    'hex = 0xff\nbig = 1_000_000\nexp = 5e+3\ninfinite = inf\n',
    TOML_CONFIG,
]


@pytest.mark.parametrize("doc", TOML_DOCS)
def test_toml_reader_matches_tomli(tmp_path, doc):
    out = _parse(tmp_path, doc, ".toml")
    assert _same(json.loads(out), tomli.loads(doc))


@pytest.mark.parametrize("bad", ['a = ', 'a = "x\n', '[t\nb = 1', 'a = 1\na = 2',
                                 'd = 1979-05-27', 'a = 1 b'])
def test_toml_reader_rejects_malformed(tmp_path, bad):
    p = tmp_path / "bad.toml"
    p.write_text(bad + "\n")
    r = _run("parse", str(p))
    assert r.returncode != 0 and "TOML parse error" in r.stderr


def test_toml_and_json_settings_are_the_same_document(tmp_path):
    t = json.loads(_parse(tmp_path, TOML_CONFIG, ".toml"))
    j = json.loads(_parse(tmp_path, json.dumps(tomli.loads(TOML_CONFIG)), ".json"))
    assert _same(t, j)


def _config(tmp_path, doc, name="c.json"):
    p = tmp_path / name
    p.write_text(json.dumps(doc))
    return _run("config", str(p))


def test_world_time_wiring(tmp_path):
    doc = tomli.loads(TOML_CONFIG)
    r = _config(tmp_path, doc)
    assert r.returncode == 0, r.stderr
    w = r.stdout.split("\n")[0].split()
    assert w[1:4] == ["32", "32", "32"]
    dx = 1.1107207345395915
    assert float(w[5]) == dx and float(w[9]) == -0.5 * dx * 32
    t = r.stdout.split("\n")[1].split()
    assert [float(v) for v in t[1:]] == [0.0, 10.0, 1.0, 1.0]
    # corner origin, flat keys
    flat = {"Lx": 8, "Ly": 4, "Lz": 2, "dx": 1.0, "dy": 2.0, "dz": 3.0, "origin": "corner",
            "t0": 0.0, "t1": 1.0, "dt": 0.1, "saveat": 0.5}
    r = _config(tmp_path, flat)
    assert r.returncode == 0, r.stderr
    assert r.stdout.split()[9:12] == ["0", "0", "0"]


@pytest.mark.parametrize("key,value,needle", [
    ("Lx", 0, "Lx"), ("Lx", 32.0, "Lx"), ("dy", -1.0, "dy"), ("origin", "middle", "origin"),
    ("t1", 0.0, "t1"), ("dt", 0.0, "dt"), ("saveat", None, "saveat")])
def test_world_time_validation(tmp_path, key, value, needle):
    doc = tomli.loads(TOML_CONFIG)
    sect = doc["domain"] if key in doc["domain"] else doc["timestepping"]
    if value is None:
        del sect[key]
    else:
        sect[key] = value
    r = _config(tmp_path, doc)
    assert r.returncode != 0 and needle in r.stderr, r.stderr


def test_modifier_factory_and_backend(tmp_path):
    doc = tomli.loads(TOML_CONFIG)
    doc["initial_conditions"].append({"type": "from_file", "filename": "x.bin"})
    doc["initial_conditions"].append({"type": "uniform_noise", "seed": 3, "low": -0.01,
                                      "high": 0.01})
    assert _config(tmp_path, doc).returncode == 0
    bad = json.loads(json.dumps(doc))
    bad["initial_conditions"].append({"type": "vacancy"})
    r = _config(tmp_path, bad)
    assert r.returncode != 0 and "unknown field modifier type" in r.stderr
    bad = json.loads(json.dumps(doc))
    del bad["initial_conditions"][1]["amp_eq"]
    assert _config(tmp_path, bad).returncode != 0
    # the device models transform on the device whatever the base FFT is, as
    # the reference's tungsten_hip does with tungsten_performance.json's fftw
    for name in ("rocm", "hip", "fftw"):
        ok = json.loads(json.dumps(doc))
        ok["plan_options"] = {"backend": name, "use_reorder": True}
        assert _config(tmp_path, ok).returncode == 0, name
    for name, needle in (("cuda", "CUDA FFT backend"), ("opencl", "unknown FFT backend")):
        bad = json.loads(json.dumps(doc))
        bad["plan_options"] = {"backend": name}
        r = _config(tmp_path, bad)
        assert r.returncode != 0 and needle in r.stderr, (name, r.stderr)


def test_seed_grid_random_seeds_and_moving_modifiers(tmp_path):
    """The modifier types of the reference's registry beyond round one's:
    seed_grid, random_seeds and moving, with the reference's field checks
    (from_json_field_modifiers.hpp:65-197)."""
    doc = tomli.loads(TOML_CONFIG)
    doc["initial_conditions"] = [
        {"target": "psi", "type": "constant", "n0": -0.10},
        {"target": "psi", "type": "seed_grid", "X0": 30.0, "Ny": 2, "Nz": 2, "radius": 25.0,
         "amplitude": 0.215936, "rho": -0.047},
        {"target": "psi", "type": "random_seeds", "amplitude": 0.2, "rho": -0.05}]
    doc["boundary_conditions"] = [
        {"target": "psi", "type": "moving", "rho_low": -0.464, "rho_high": -0.10,
         "width": 15.0, "alpha": 1.0, "disp": 40.0, "xpos": 127.17225402106772}]
    r = _config(tmp_path, doc)
    assert r.returncode == 0, r.stderr
    for typ, key in (("seed_grid", "X0"), ("seed_grid", "Nz"), ("random_seeds", "rho"),
                     ("moving", "xpos"), ("moving", "width")):
        bad = json.loads(json.dumps(doc))
        mods = bad["initial_conditions"] + bad["boundary_conditions"]
        m = next(x for x in mods if x["type"] == typ)
        del m[key]
        r = _config(tmp_path, bad)
        assert r.returncode != 0, (typ, key)
        assert f"missing or invalid '{key}' field" in r.stderr, r.stderr


def test_oracle_fixed_bc_closed_form():
    from oracle import oracle as O
    nx, dx = 64, 1.0
    a = np.full((2, 3, nx), 7.0)
    x0 = 0.0
    O.bc_fixed(a, x0, dx, -0.464, -0.1)
    xpos = nx * dx - 20.0
    x = x0 + np.arange(nx) * dx
    band = np.abs(x - xpos) < 20.0
    assert np.all(a[..., ~band] == 7.0)
    S = 1.0 / (1.0 + np.exp(-(x - xpos)))
    assert np.allclose(a[0, 0, band], -0.464 * S[band] - 0.1 * (1 - S[band]), rtol=0,
                       atol=1e-15)
    # x = xpos exactly (i = 44 here): S = 1/2
    i = int(round((xpos - x0) / dx))
    assert a[1, 2, i] == 0.5 * (-0.464) + 0.5 * (-0.1)
    # left end of the band tends to rho_high, right end to rho_low
    assert abs(a[0, 0, np.argmax(band)] - (-0.1)) < 1e-6
    assert abs(a[0, 0, nx - 1] - (-0.464)) < 1e-6
