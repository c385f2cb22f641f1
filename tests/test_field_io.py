"""Raw binary field IO (the reference's BinaryWriter / BinaryReader) through
the library's host path; no device needed.

Recipes follow tests/unit/frontend/io/test_binary_writer.cpp: global {4,2,2}
split on x into two bricks (the two-rank case, written by two calls), real
values i + 1000*rank + 0.25*offset_x and complex (i + 10*rank, 2i + offset_x),
round trip within 1e-12, invalid domains, buffer size mismatches.  The file
layout itself is pinned against numpy: the global array (nz, ny, nx) in C
order (x fastest) as little-endian f64, complex as interleaved re/im f64.
"""
import numpy as np
import pytest

from openpfc_amd import io as IO

G = (4, 2, 2)
BRICKS = [((2, 2, 2), (0, 0, 0)), ((2, 2, 2), (2, 0, 0))]


def make_real(local, offset, rank):
    n = local[0] * local[1] * local[2]
    return np.arange(n, dtype=np.float64) + 1000.0 * rank + 0.25 * offset[0]


def make_complex(local, offset, rank):
    n = local[0] * local[1] * local[2]
    i = np.arange(n, dtype=np.float64)
    return (i + 10.0 * rank) + 1j * (2.0 * i + offset[0])


def global_from_bricks(parts, dtype):
    """Assemble the global (nz, ny, nx) array from x-fastest bricks."""
    out = np.zeros((G[2], G[1], G[0]), dtype=dtype)
    for (local, offset), data in parts:
        lx, ly, lz = local
        ox, oy, oz = offset
        out[oz:oz + lz, oy:oy + ly, ox:ox + lx] = data.reshape(lz, ly, lx)
    return out


def test_single_brick_real_round_trip_and_layout(tmp_path):
    data = make_real(G, (0, 0, 0), 0)
    w = IO.BinaryWriter(str(tmp_path / "real_%d.bin"))
    w.set_domain(G, G, (0, 0, 0))
    path = w.write(0, data)
    assert path.endswith("real_0.bin")
    raw = open(path, "rb").read()
    assert raw == data.astype("<f8").tobytes()
    back = np.zeros_like(data)
    r = IO.BinaryReader()
    r.set_domain(G, G, (0, 0, 0))
    r.read(path, back)
    assert np.abs(back - data).max() <= 1e-12


@pytest.mark.parametrize("kind", ["real", "complex"])
def test_two_bricks_split_on_x(tmp_path, kind):
    make = make_real if kind == "real" else make_complex
    dt = np.float64 if kind == "real" else np.complex128
    path = str(tmp_path / f"{kind}.bin")
    parts = []
    for rank, (local, offset) in enumerate(BRICKS):
        data = make(local, offset, rank)
        w = IO.BinaryWriter(path)
        w.set_domain(G, local, offset)
        w.write(0, data, truncate=(rank == 0))
        parts.append(((local, offset), data))
    expect = global_from_bricks(parts, dt)
    raw = np.fromfile(path, dtype="<c16" if kind == "complex" else "<f8")
    assert np.array_equal(raw, expect.ravel())
    for (local, offset), data in parts:
        back = np.zeros_like(data)
        r = IO.BinaryReader()
        r.set_domain(G, local, offset)
        r.read(path, back)
        assert np.abs(back - data).max() <= 1e-12


def test_real_then_complex_increments(tmp_path):
    w = IO.BinaryWriter(str(tmp_path / "f_%d.bin"))
    w.set_domain(G, G, (0, 0, 0))
    p0 = w.write(0, make_real(G, (0, 0, 0), 0))
    p1 = w.write(1, make_complex(G, (0, 0, 0), 0))
    assert p0.endswith("f_0.bin") and p1.endswith("f_1.bin")
    assert len(open(p0, "rb").read()) == 16 * 8
    assert len(open(p1, "rb").read()) == 16 * 16


def test_float32_buffers_widen_and_round(tmp_path):
    g = (8, 4, 2)
    data = (np.arange(64, dtype=np.float32) / 3.0).astype(np.float32)
    path = str(tmp_path / "f32.bin")
    IO.write_field(path, data, g)
    raw = np.fromfile(path, dtype="<f8")
    assert np.array_equal(raw, data.astype(np.float64))
    wide = np.arange(64, dtype=np.float64) / 7.0
    IO.write_field(path, wide, g)
    back = np.zeros(64, dtype=np.float32)
    IO.read_field(path, back, g)
    assert np.array_equal(back, wide.astype(np.float32))


def test_larger_than_one_chunk(tmp_path):
    # 5M f32 values (> one 32 MiB f64 staging chunk) through the widening path
    g = (160, 128, 256)
    data = np.random.default_rng(5).standard_normal(g[0] * g[1] * g[2]).astype(np.float32)
    path = str(tmp_path / "big.bin")
    IO.write_field(path, data, g)
    assert np.array_equal(np.fromfile(path, dtype="<f8"), data.astype(np.float64))
    back = np.empty_like(data)
    IO.read_field(path, back, g)
    assert np.array_equal(back, data)
    # an interior brick of it
    local, offset = (50, 60, 70), (100, 13, 111)
    brick = np.empty(local[0] * local[1] * local[2], dtype=np.float64)
    r = IO.BinaryReader()
    r.set_domain(g, local, offset)
    r.read(path, brick)
    full = data.astype(np.float64).reshape(g[2], g[1], g[0])
    ox, oy, oz = offset
    lx, ly, lz = local
    assert np.array_equal(brick.reshape(lz, ly, lx),
                          full[oz:oz + lz, oy:oy + ly, ox:ox + lx])


@pytest.mark.parametrize("cls", [IO.BinaryWriter, IO.BinaryReader])
def test_invalid_domain_geometry(cls):
    io = cls("x.bin") if cls is IO.BinaryWriter else cls()
    for g, l, o in [((0, 8, 8), (0, 8, 8), (0, 0, 0)), ((8, 8, 8), (0, 8, 8), (0, 0, 0)),
                    ((8, 8, 8), (8, 8, 8), (-1, 0, 0)), ((8, 8, 8), (8, 8, 8), (4, 0, 0))]:
        with pytest.raises(ValueError):
            io.set_domain(g, l, o)
    io.set_domain((8, 8, 8), (8, 8, 8), (0, 0, 0))


def test_buffer_size_mismatch_and_state(tmp_path):
    w = IO.BinaryWriter(str(tmp_path / "m.bin"))
    with pytest.raises(RuntimeError, match="set_domain"):
        w.write(0, np.zeros(16))
    w.set_domain(G, G, (0, 0, 0))
    for n in (15, 17):
        with pytest.raises(RuntimeError, match="buffer size mismatch"):
            w.write(0, np.zeros(n))
    path = w.write(0, np.ones(16))
    r = IO.BinaryReader()
    r.set_domain(G, G, (0, 0, 0))
    for n in (15, 17):
        with pytest.raises(RuntimeError, match="buffer size mismatch"):
            r.read(path, np.zeros(n))


def test_short_file_and_missing_file(tmp_path):
    path = str(tmp_path / "short.bin")
    np.ones(10).astype("<f8").tofile(path)
    r = IO.BinaryReader()
    r.set_domain(G, G, (0, 0, 0))
    with pytest.raises(RuntimeError, match="Short read"):
        r.read(path, np.zeros(16))
    with pytest.raises(RuntimeError, match="cannot open"):
        r.read(str(tmp_path / "nope.bin"), np.zeros(16))


def test_format_with_number():
    assert IO.format_with_number("output_%d.dat", 7) == "output_7.dat"
    assert IO.format_with_number("frame_%04d.vti", 3) == "frame_0003.vti"
    assert IO.format_with_number("plain.bin", 3) == "plain.bin"
    for bad in ("a_%s.bin", "a_%d_%d.bin", "a_%n", "a_%f"):
        with pytest.raises(ValueError):
            IO.format_with_number(bad, 1)
