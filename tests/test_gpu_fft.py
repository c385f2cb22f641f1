"""GPU parity of the HIP R2C/C2R transforms against the CPU oracle and
numpy.fft, plus the reference's own FFT known-answer tests.

Reference tests mirrored:
  tests/unit/kernel/fft/test_fft.cpp:37-47,59-70      (8-pt DC, 2-pt backward)
  tests/unit/runtime/gpu/test_fft_hip.cpp:105-117     (64^3 DC 1000 -> 1000/64^3)
  tests/unit/runtime/gpu/test_fft_hip.cpp:148-168     (32^3 roundtrip <= 1e-5)
  tests/integration/scenarios/gpu_validation/test_hip_roundtrip.cpp:61,95
                                                      (16^3 roundtrip, 1e-10 / 1e-5)
  tests/unit/kernel/fft/test_fft_buffer_sizes.cpp     (size mismatch -> error)
Tolerances: FP64 transforms within 1e-13 of the oracle relative to max|X|;
FP32 within 2e-6.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

SHAPES = [(2, 1, 1), (8, 1, 1), (4, 2, 2), (16, 16, 16), (32, 8, 4), (8, 32, 64),
          (64, 64, 64), (128, 16, 32), (256, 4, 8), (512, 2, 4), (1024, 8, 4),
          (2048, 2, 2), (4, 1024, 2), (4, 2, 1024), (128, 128, 128)]


def _fft(torch, shape, prec):
    import openpfc_amd as P
    return P.HipFft(*shape, precision=prec)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_r2c_matches_oracle(torch_gpu, shape, prec):
    torch = torch_gpu
    import openpfc_amd as P
    nx, ny, nz = shape
    print(f"r2c shape={shape} prec={prec}", flush=True)
    rng = np.random.default_rng(nx * 7 + ny * 3 + nz)
    a = rng.standard_normal((nz, ny, nx))
    ref = O.r2c(a)
    f = P.HipFft(nx, ny, nz, P.F64 if prec == "f64" else P.F32)
    rdt = torch.float64 if prec == "f64" else torch.float32
    cdt = torch.complex128 if prec == "f64" else torch.complex64
    x = torch.from_numpy(a).to(rdt).cuda()
    X = torch.empty((nz, ny, nx // 2 + 1), dtype=cdt, device="cuda")
    f.forward(x, X)
    got = X.cpu().numpy().astype(np.complex128)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    tol = 1e-13 if prec == "f64" else 2e-6
    assert err <= tol, f"r2c rel err {err:.3e} > {tol}"
    # numpy cross-check (independent FFT)
    assert np.abs(got - np.fft.rfftn(a)).max() / np.abs(ref).max() <= max(tol, 1e-13)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_c2r_matches_oracle_and_keeps_input(torch_gpu, shape, prec):
    torch = torch_gpu
    import openpfc_amd as P
    nx, ny, nz = shape
    print(f"c2r shape={shape} prec={prec}", flush=True)
    rng = np.random.default_rng(nx + 5 * ny + 11 * nz)
    a = rng.standard_normal((nz, ny, nx))
    spec = O.r2c(a)
    ref = O.c2r(spec, nx)
    f = P.HipFft(nx, ny, nz, P.F64 if prec == "f64" else P.F32)
    rdt = torch.float64 if prec == "f64" else torch.float32
    cdt = torch.complex128 if prec == "f64" else torch.complex64
    X = torch.from_numpy(spec).to(cdt).cuda()
    X0 = X.clone()
    y = torch.empty((nz, ny, nx), dtype=rdt, device="cuda")
    f.backward(X, y)
    got = y.cpu().numpy().astype(np.float64)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    tol = 1e-13 if prec == "f64" else 2e-6
    assert err <= tol, f"c2r rel err {err:.3e} > {tol}"
    assert torch.equal(X, X0), "backward modified its const input"


def test_kat_8pt_forward_dc(torch_gpu):
    """test_fft.cpp:37-47: 8x1x1 input, Re X0 = 21.991 +- 0.01."""
    torch = torch_gpu
    import openpfc_amd as P
    f = P.HipFft(8, 1, 1)
    x = torch.tensor([0.000, 0.785, 1.571, 2.356, 3.142, 3.927, 4.712, 5.498],
                     dtype=torch.float64, device="cuda")
    X = torch.empty(5, dtype=torch.complex128, device="cuda")
    f.forward(x, X)
    assert abs(X[0].real.item() - 21.991) <= 0.01


def test_kat_2pt_backward(torch_gpu):
    """test_fft.cpp:59-70: 2x1x1 spectrum {1, 2}, backward -> out[0] = 1.5."""
    torch = torch_gpu
    import openpfc_amd as P
    f = P.HipFft(2, 1, 1)
    X = torch.tensor([1.0, 2.0], dtype=torch.complex128, device="cuda")
    y = torch.empty(2, dtype=torch.float64, device="cuda")
    f.backward(X, y)
    assert abs(y[0].item() - 1.5) <= 0.01


def test_kat_dc_backward_64(torch_gpu):
    """test_fft_hip.cpp:105-117: DC 1000 at 64^3 -> 1000/64^3 everywhere."""
    torch = torch_gpu
    import openpfc_amd as P
    f = P.HipFft(64, 64, 64)
    X = torch.zeros(f.size_outbox(), dtype=torch.complex128, device="cuda")
    X[0] = 1000.0
    y = torch.empty(f.size_inbox(), dtype=torch.float64, device="cuda")
    f.backward(X, y)
    assert torch.allclose(y, torch.full_like(y, 1000.0 / 64 ** 3), atol=1e-6, rtol=0)


@pytest.mark.parametrize("prec,tol", [("f64", 1e-10), ("f32", 1e-5)])
def test_roundtrip_16(torch_gpu, prec, tol):
    """test_hip_roundtrip.cpp:45-61 (double) and :85-95 (float)."""
    torch = torch_gpu
    import openpfc_amd as P
    f = P.HipFft(16, 16, 16, P.F64 if prec == "f64" else P.F32)
    n = f.size_inbox()
    i = np.arange(n)
    h = 0.25 + 0.5 * np.sin(2.0 * np.pi * i / n)
    rdt = torch.float64 if prec == "f64" else torch.float32
    cdt = torch.complex128 if prec == "f64" else torch.complex64
    x = torch.from_numpy(h).to(rdt).cuda()
    X = torch.empty(f.size_outbox(), dtype=cdt, device="cuda")
    y = torch.empty_like(x)
    f.forward(x, X)
    f.backward(X, y)
    assert (y.double().cpu() - x.double().cpu()).abs().max().item() <= tol


def test_roundtrip_32_parity_pattern(torch_gpu):
    """test_fft_hip.cpp:148-168: input i % 2 at 32^3, roundtrip <= 1e-5."""
    torch = torch_gpu
    import openpfc_amd as P
    f = P.HipFft(32, 32, 32)
    x = (torch.arange(f.size_inbox(), device="cuda") % 2).double()
    X = torch.empty(f.size_outbox(), dtype=torch.complex128, device="cuda")
    y = torch.empty_like(x)
    f.forward(x, X)
    f.backward(X, y)
    assert (y - x).abs().max().item() <= 1e-5


def test_size_mismatch_raises(torch_gpu):
    """test_fft_buffer_sizes.cpp: wrong extents -> invalid_argument."""
    torch = torch_gpu
    import openpfc_amd as P
    f = P.HipFft(8, 8, 8)
    x = torch.zeros(8 * 8 * 8 - 1, dtype=torch.float64, device="cuda")
    X = torch.zeros(f.size_outbox(), dtype=torch.complex128, device="cuda")
    with pytest.raises(ValueError, match="size_inbox"):
        f.forward(x, X)
    x = torch.zeros(8 * 8 * 8, dtype=torch.float64, device="cuda")
    X = torch.zeros(f.size_outbox() + 1, dtype=torch.complex128, device="cuda")
    with pytest.raises(ValueError, match="size_outbox"):
        f.forward(x, X)
