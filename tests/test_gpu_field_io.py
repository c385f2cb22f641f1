"""Device-resident raw binary field IO (pinned, double-buffered staging) and a
restart through it: a PFC run written after 10 steps, read back and
continued gives bit for bit the uninterrupted 20-step field (the "from_file"
initial condition, initial_conditions/file_reader.hpp:48-92)."""
import math

import numpy as np
import pytest

from openpfc_amd import io as IO

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("g", [(16, 8, 4), (256, 128, 128)])
def test_device_write_matches_host_layout(torch_gpu, tmp_path, g):
    torch = torch_gpu
    n = g[0] * g[1] * g[2]
    host = np.random.default_rng(1).standard_normal(n)
    dev = torch.from_numpy(host).cuda()
    path = str(tmp_path / "d.bin")
    IO.write_field(path, dev, g)
    assert np.array_equal(np.fromfile(path, dtype="<f8"), host)
    back = torch.zeros_like(dev)
    IO.read_field(path, back, g)
    assert torch.equal(back, dev)


def test_device_float32_and_complex(torch_gpu, tmp_path):
    torch = torch_gpu
    g = (128, 64, 32)
    n = g[0] * g[1] * g[2]
    x = torch.randn(n, dtype=torch.float32, device="cuda")
    p = str(tmp_path / "f32.bin")
    IO.write_field(p, x, g)
    assert np.array_equal(np.fromfile(p, dtype="<f8"), x.double().cpu().numpy())
    y = torch.empty_like(x)
    IO.read_field(p, y, g)
    assert torch.equal(x, y)
    c = torch.randn(n, dtype=torch.complex128, device="cuda")
    pc = str(tmp_path / "c.bin")
    IO.write_field(pc, c, g)
    assert np.array_equal(np.fromfile(pc, dtype="<c16"), c.cpu().numpy())
    d = torch.empty_like(c)
    IO.read_field(pc, d, g)
    assert torch.equal(c, d)


def test_device_bricks(torch_gpu, tmp_path):
    torch = torch_gpu
    g = (64, 32, 16)
    full = np.random.default_rng(2).standard_normal(g[0] * g[1] * g[2])
    glob = full.reshape(g[2], g[1], g[0])
    path = str(tmp_path / "b.bin")
    # two z-slabs and one interior brick written from the device
    pieces = [((64, 32, 8), (0, 0, 0)), ((64, 32, 8), (0, 0, 8))]
    for i, (l, o) in enumerate(pieces):
        part = glob[o[2]:o[2] + l[2], o[1]:o[1] + l[1], o[0]:o[0] + l[0]].copy()
        w = IO.BinaryWriter(path)
        w.set_domain(g, l, o)
        w.write(0, torch.from_numpy(part.ravel()).cuda(), truncate=(i == 0))
    assert np.array_equal(np.fromfile(path, dtype="<f8"), full)
    l, o = (20, 10, 5), (33, 7, 9)
    r = IO.BinaryReader()
    r.set_domain(g, l, o)
    brick = torch.zeros(l[0] * l[1] * l[2], dtype=torch.float64, device="cuda")
    r.read(path, brick)
    want = glob[o[2]:o[2] + l[2], o[1]:o[1] + l[1], o[0]:o[0] + l[0]]
    assert np.array_equal(brick.cpu().numpy().reshape(l[2], l[1], l[0]), want)


def test_restart_from_file_is_bitwise(torch_gpu, tmp_path):
    import openpfc_amd as P
    torch = torch_gpu
    n = (64, 64, 64)
    dx = math.pi / 4
    f = P.HipFft(*n)
    st = P.PfcStepper(f, (dx, dx, dx), -0.25)
    st.initialize(0.5)
    psi = torch.empty(n[0] * n[1] * n[2], dtype=torch.float64, device="cuda")
    P.ic_constant(f, psi, -0.2)
    P.ic_add_noise(f, psi, 1234, -0.01, 0.01)
    ref = psi.clone()
    st.step(ref, 20)
    st.step(psi, 10)
    w = IO.BinaryWriter(str(tmp_path / "psi_%04d.bin"))
    w.set_domain(n, n, (0, 0, 0))
    path = w.write(10, psi)
    assert path.endswith("psi_0010.bin")
    again = torch.empty_like(psi)
    r = IO.BinaryReader()
    r.set_domain(n, n, (0, 0, 0))
    r.read(path, again)
    st.step(again, 10)
    assert torch.equal(again, ref)
