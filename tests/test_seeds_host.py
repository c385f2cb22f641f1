"""Rotated seeds, SeedGrid, RandomSeeds and MovingBC: the oracle pinned, and
the C++ host layer's paths against it (no device).

Oracle pins:
* uniform_real_distribution<double> over mt19937_64: the oracle's draws equal
  libstdc++'s own (std::uniform_real_distribution compiled into the host test
  driver), bit for bit; mt19937_64 itself is pinned by the standard's KAT
  (test_oracle_ics.py);
* Seed rotation (seed.hpp:58-104): orientation (0,0,0) leaves the six
  <110>/sqrt(2) vectors as they are; a quarter turn about one axis maps them
  as the right-handed rotation about that axis does; yaw and roll together
  apply roll first (R = yaw * pitch * roll), which the reverse order would
  not give; every rotation keeps |q| = 1 and the pairwise dot products;
* the seed field (seed.hpp:124-141, first seed wins, 0.0 outside, X
  absolute) against a direct numpy evaluation;
* MovingBC (moving_bc.hpp:107-219): front index and boundary position from
  their definitions on a profile built to place them, the band's closed form,
  its wrap round the periodic box, and no move back.

C++ host layer: SeedGrid and RandomSeeds lists and fields, and three MovingBC
calls on a moving front, equal the oracle's.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "openpfc_amd", "host", "test_host")
DX = 1.1107207345395915
S2 = 1.0 / math.sqrt(2.0)
Q0 = np.array([[S2, S2, 0], [S2, 0, S2], [0, S2, S2], [S2, 0, -S2], [S2, -S2, 0], [0, S2, -S2]])


def _q(row):
    return np.asarray(row)[3:21].reshape(6, 3)


def test_seed_rotation_closed_forms():
    q = _q(O.seed_make((1, 2, 3), (0, 0, 0), 5.0, -0.1, 0.2))
    assert np.array_equal(q, Q0)
    h = math.pi / 2
    # yaw: about z, (x, y) -> (-y, x)
    q = _q(O.seed_make((0, 0, 0), (h, 0, 0), 1, 0, 0))
    assert np.allclose(q, np.c_[-Q0[:, 1], Q0[:, 0], Q0[:, 2]], atol=1e-15)
    # pitch: about y, (x, z) -> (z, -x)
    q = _q(O.seed_make((0, 0, 0), (0, h, 0), 1, 0, 0))
    assert np.allclose(q, np.c_[Q0[:, 2], Q0[:, 1], -Q0[:, 0]], atol=1e-15)
    # roll: about x, (y, z) -> (-z, y)
    q = _q(O.seed_make((0, 0, 0), (0, 0, h), 1, 0, 0))
    assert np.allclose(q, np.c_[Q0[:, 0], -Q0[:, 2], Q0[:, 1]], atol=1e-15)
    # yaw and roll together: roll acts first, (s, s, 0) -> (s, 0, s) -> (0, s, s)
    q = _q(O.seed_make((0, 0, 0), (h, 0, h), 1, 0, 0))
    assert np.allclose(q[0], [0, S2, S2], atol=1e-15)
    assert not np.allclose(q[0], [-S2, 0, S2], atol=1e-3)  # the other order
    rng = np.random.default_rng(5)
    for _ in range(20):
        q = _q(O.seed_make((0, 0, 0), rng.uniform(0, 2 * math.pi, 3), 1, 0, 0))
        assert np.allclose(np.linalg.norm(q, axis=1), 1.0, atol=1e-15)
        assert np.allclose(q @ q.T, Q0 @ Q0.T, atol=1e-15)


def test_seed_field_against_direct_evaluation():
    n = (24, 20, 16)
    origin, sp = (-6.0, -5.0, -4.0), (0.5, 0.5, 0.5)
    seeds = np.array([O.seed_make((0.3, 0.1, 0.0), (0.4, 1.1, 2.0), 3.0, -0.05, 0.2),
                      O.seed_make((2.0, 0.5, -0.5), (1.0, 0.2, 0.3), 2.5, 0.1, 0.3)])
    a = np.full((n[2], n[1], n[0]), 7.0)
    O.ic_seeds(a, origin, sp, seeds)
    z, y, x = np.meshgrid(*(origin[d] + sp[d] * np.arange(n[d]) for d in (2, 1, 0)),
                          indexing="ij")
    ref = np.zeros_like(a)
    done = np.zeros(a.shape, bool)
    for s in seeds:
        inside = ((x - s[0]) ** 2 + (y - s[1]) ** 2 + (z - s[2]) ** 2 < s[22] ** 2) & ~done
        val = s[21] + sum(2 * s[23] * np.cos(q[0] * x + q[1] * y + q[2] * z) for q in _q(s))
        ref[inside] = val[inside]
        done |= inside
    assert done.sum() > 500 and (~done).sum() > 500
    both = ((x - seeds[1][0]) ** 2 + (y - seeds[1][1]) ** 2 + (z - seeds[1][2]) ** 2
            < seeds[1][22] ** 2) & ((x - seeds[0][0]) ** 2 + (y - seeds[0][1]) ** 2
                                    + (z - seeds[0][2]) ** 2 < seeds[0][22] ** 2)
    assert both.sum() > 10  # the first seed wins where the two overlap
    assert np.abs(a - ref).max() <= 1e-14
    kept = np.full(a.shape, 7.0)
    O.ic_seeds(kept, origin, sp, seeds, keep_outside=True)
    assert np.all(kept[~done] == 7.0) and np.array_equal(kept[done], a[done])


def test_moving_bc_definitions():
    nx, ny, nz, dx, x0 = 64, 4, 3, 0.5, 0.0
    a = np.full((nz, ny, nx), -0.3)
    a[1, 2, :20] = 0.5  # solid up to column 19 (max over y, z > 0.1)
    bc = O.MovingBC(-0.4, -0.1, xwidth=3.0, alpha=2.0, disp=5.0, xpos=0.0)
    xl = bc.apply(a, x0, dx)
    assert np.array_equal(xl, np.where(np.arange(nx) < 20, 0.5, -0.3))
    assert bc.idx == 19 and bc.xpos == 19 * dx + 5.0
    # band: |x - xpos| < 3 written with the sigmoid, the rest untouched
    x = x0 + dx * np.arange(nx)
    d = x - bc.xpos
    band = np.abs(d) < 3.0
    S = 1.0 / (1.0 + np.exp(-2.0 * d))
    want = -0.4 * S + -0.1 * (1 - S)
    assert np.allclose(a[0, 0, band], want[band], rtol=0, atol=1e-16)
    assert np.all(a[0, 0, ~band] == -0.3)
    # second call advances while the column ahead is above threshold
    b = np.full((nz, ny, nx), -0.3)
    b[0, 0, :25] = 0.5
    bc.apply(b, x0, dx)
    assert bc.idx == 25 and bc.xpos == 25 * dx + 5.0
    # a front that went back does not move the boundary back
    c = np.full((nz, ny, nx), -0.3)
    bc.apply(c, x0, dx)
    assert bc.idx == 25 and bc.xpos == 25 * dx + 5.0
    # wrap: xpos within xwidth of the box end writes both ends
    w = O.MovingBC(-0.4, -0.1, xwidth=3.0, alpha=1.0, disp=0.0, xpos=31.0, threshold=10.0)
    e = np.zeros((nz, ny, nx))
    w.apply(e, x0, dx)
    changed = np.nonzero(e[0, 0] != 0.0)[0]
    l = nx * dx
    dd = x - 31.0
    dd = np.where(np.abs(dd) < 3, dd, np.where(np.abs(dd + l) < 3, dd + l, np.nan))
    assert set(changed) == set(np.nonzero(~np.isnan(dd))[0])
    assert 0 in changed and nx - 1 in changed
    # a field above threshold everywhere stops after one period
    f = np.full((nz, ny, nx), 1.0)
    m = O.MovingBC(-0.4, -0.1, threshold=0.1)
    m.apply(f, x0, dx)
    assert m.idx == nx - 1
    f[:] = 1.0  # the band the first call wrote is above threshold again
    m.apply(f, x0, dx)
    assert m.idx == 2 * nx - 1


@pytest.fixture(scope="module")
def host_out(tmp_path_factory):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} missing: run __graft_entry__.build()")
    d = tmp_path_factory.mktemp("seeds")
    r = subprocess.run([BIN, "seeds", str(d)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    return d, r.stdout


def _read(d, name, shape):
    return np.fromfile(os.path.join(d, name), dtype="<f8").reshape(shape)


def test_uniform_real_matches_libstdcxx(host_out):
    d, _ = host_out
    assert np.array_equal(_read(d, "draws.bin", (16,)), O.uniform_real_seq(42, -3.0, 5.0, 16))


def test_host_seed_lists_and_fields_match_oracle(host_out):
    d, _ = host_out
    grid = O.seed_grid((128, 64, 64), (DX,) * 3, 2, 2, 30.0, 25.0, -0.047, 0.215936)
    assert np.array_equal(_read(d, "grid_rows.bin", (4, 24)), grid)
    rnd = O.random_seeds(-0.047, 0.215936)
    assert np.array_equal(_read(d, "random_rows.bin", (150, 24)), rnd)
    ref = O.ic_seeds(np.empty((64, 64, 128)), (0, 0, 0), (DX,) * 3, grid)
    assert np.array_equal(_read(d, "grid_field.bin", (64, 64, 128)), ref)
    ref = O.ic_seeds(np.empty((64, 64, 64)), (-128,) * 3, (4.0,) * 3, rnd)
    got = _read(d, "random_field.bin", (64, 64, 64))
    assert np.array_equal(got, ref) and (ref != 0).sum() > 1000


def test_host_moving_bc_matches_oracle(host_out):
    d, out = host_out
    grid = O.seed_grid((128, 64, 64), (DX,) * 3, 2, 2, 30.0, 25.0, -0.047, 0.215936)
    base = O.ic_seeds(np.full((64, 64, 128), -0.1), (0, 0, 0), (DX,) * 3, grid)
    bc = O.MovingBC(-0.464, -0.1, xwidth=15.0, alpha=1.0, disp=40.0, xpos=0.0)
    lines = [ln.split() for ln in out.splitlines() if ln.startswith("moving ")]
    for call in range(3):
        f = np.ascontiguousarray(np.roll(base, 9 * call, axis=2))
        bc.apply(f, 0.0, DX)
        assert int(lines[call][3]) == bc.idx and float(lines[call][5]) == bc.xpos
        assert np.array_equal(_read(d, f"moving_{call}.bin", f.shape), f), call
