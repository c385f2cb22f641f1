"""pfc_hip_stream_copy, the plain streaming copy bench.py reports as the
bandwidth a pass could reach: copies exactly, and refuses sizes and pointers
that are not 16-byte multiples and overlapping ranges."""
import ctypes

import pytest

pytestmark = pytest.mark.gpu


def _call(lib, src, dst, nbytes):
    return lib.pfc_hip_stream_copy(ctypes.c_void_p(src), ctypes.c_void_p(dst), nbytes, None)


def test_stream_copy_copies_exactly(torch_gpu):
    torch = torch_gpu
    from openpfc_amd._lib import load
    lib = load()
    for n in (2, 1000, 1 << 20, 3 * (1 << 20) + 14):  # doubles; 16 * n/2 bytes
        a = torch.randn(n, dtype=torch.float64, device="cuda")
        b = torch.zeros_like(a)
        assert _call(lib, a.data_ptr(), b.data_ptr(), n * 8) == 0
        torch.cuda.synchronize()
        assert torch.equal(a, b)


def test_stream_copy_rejects_bad_arguments(torch_gpu):
    torch = torch_gpu
    from openpfc_amd._lib import load
    lib = load()
    a = torch.zeros(64, dtype=torch.float64, device="cuda")
    b = torch.zeros(64, dtype=torch.float64, device="cuda")
    assert _call(lib, a.data_ptr(), b.data_ptr(), 24) == 1           # not a multiple of 16
    assert _call(lib, a.data_ptr() + 8, b.data_ptr(), 16) == 1       # misaligned source
    assert _call(lib, a.data_ptr(), a.data_ptr() + 128, 256) == 1    # overlap
    assert _call(lib, 0, b.data_ptr(), 16) == 1                      # null
    assert _call(lib, a.data_ptr(), b.data_ptr(), 0) == 0            # nothing to do
