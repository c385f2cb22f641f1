"""pfc_hip_column_max_noalloc (the allocation-free column maxima) against
pfc_hip_column_max, through the C ABI: bit for bit (a maximum is exact), on
two calls in a row, the second with a smaller grid, each checked against
numpy as well; sizes beyond the scratch are refused with PFC_HIP_EINVAL."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(P, torch, f, psi):
    out = torch.full((f.n[0],), float("nan"), dtype=torch.float64, device="cuda")
    s = torch.cuda.current_stream()
    P.check(f._lib.pfc_hip_column_max_noalloc(f.handle, ctypes.c_void_p(psi.data_ptr()),
                                              ctypes.c_void_p(out.data_ptr()),
                                              ctypes.c_void_p(s.cuda_stream)))
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_column_max_noalloc_matches_allocating_entry_point(torch_gpu, prec):
    import openpfc_amd as P
    torch = torch_gpu
    f64 = prec == "f64"
    dt = torch.float64 if f64 else torch.float32
    for call, n in enumerate(((2048, 8, 64), (128, 16, 8))):  # the second call is smaller
        print(f"noalloc column max {prec} call {call} {n}", flush=True)
        a = np.random.default_rng(11 + call).standard_normal((n[2], n[1], n[0]))
        f = P.HipFft(*n, precision=P.F64 if f64 else P.F32)
        psi = torch.from_numpy(a.ravel()).to("cuda", dt)
        got = _noalloc(P, torch, f, psi)
        ref = P.column_max(f, psi)
        assert np.array_equal(got, ref)
        want = a.max(axis=(0, 1)) if f64 else a.astype(np.float32).max(axis=(0, 1))
        assert np.array_equal(got, want.astype(np.float64))
        f.close()


def test_column_max_noalloc_refuses_null_buffers(torch_gpu):
    """A null output buffer is refused before any launch (every plan the
    library makes has nx <= 2048, within the scratch)."""
    import openpfc_amd as P
    torch = torch_gpu
    f = P.HipFft(64, 4, 4)
    psi = torch.zeros(64 * 16, dtype=torch.float64, device="cuda")
    rc = f._lib.pfc_hip_column_max_noalloc(f.handle, ctypes.c_void_p(psi.data_ptr()), None, None)
    assert rc == 1 and b"null" in f._lib.pfc_hip_last_error()
