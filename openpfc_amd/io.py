"""Raw binary field IO: the reference's BinaryWriter / BinaryReader.

``BinaryWriter`` mirrors include/openpfc/frontend/io/binary_writer.hpp:76-205
and ``BinaryReader`` include/openpfc/kernel/simulation/binary_reader.hpp:
79-200: ``set_domain(global, local, offset)`` describes the caller's brick of
the global array, ``write(increment, field)`` writes it into
``format_with_number(filename, increment)`` (the file is truncated first, as
the reference's single writer does), ``read(filename, field)`` fills it.  The
file is the global array, x fastest, native little-endian f64 (complex: re, im
f64), no header.

Fields are torch tensors (device tensors go through the library's pinned
staging, host tensors and numpy arrays through its host path) of float64,
float32, complex128 or complex64 with the brick's element count.  Errors
follow the reference: bad domains raise ValueError (std::invalid_argument),
size mismatches and file errors RuntimeError (std::runtime_error).
"""
from __future__ import annotations

import ctypes
import re

from . import _lib
from ._lib import F32, F64, IO_COMPLEX, IO_REAL, PfcHipError, load

# frontend/utils/utils.hpp:96-104: one %d conversion with flags and width.
# A '*' width would need a second argument the reference never passes, so it
# is rejected here rather than formatted.
_PATTERN = re.compile(r"^[^%]*%[-+ 0]*[0-9]*d[^%]*$")


def format_with_number(filename: str, increment: int) -> str:
    """frontend/utils/utils.hpp:133-139: replace the single %d of a pattern."""
    if "%" not in filename:
        return filename
    if not _PATTERN.match(filename):
        raise ValueError(f"format_with_number: invalid filename pattern '{filename}'. Must "
                         "contain at most one % integer specifier (%d with optional "
                         "flags/width).")
    return filename % int(increment)


def _validate_domain(g, l, o, ctx):
    """kernel/mpi/domain_geometry.hpp:36-60."""
    g, l, o = (tuple(int(v) for v in a) for a in (g, l, o))
    if len(g) != 3 or len(l) != 3 or len(o) != 3:
        raise ValueError(f"{ctx}: domain arrays must have 3 entries")
    for i in range(3):
        if g[i] <= 0 or l[i] <= 0:
            raise ValueError(f"{ctx}: global/local dimensions must be positive")
        if o[i] < 0:
            raise ValueError(f"{ctx}: offsets must be non-negative")
        if o[i] + l[i] > g[i]:
            raise ValueError(f"{ctx}: piece does not lie inside global domain (check "
                             "offset + local_size vs global)")
    return g, l, o


def _field_info(field):
    """(pointer, precision, kind, count, on_device) of a tensor or array."""
    try:
        import torch
    except ImportError:  # pragma: no cover
        torch = None
    if torch is not None and isinstance(field, torch.Tensor):
        if not field.is_contiguous():
            raise ValueError("field must be contiguous")
        table = {torch.float64: (F64, IO_REAL), torch.float32: (F32, IO_REAL),
                 torch.complex128: (F64, IO_COMPLEX), torch.complex64: (F32, IO_COMPLEX)}
        if field.dtype not in table:
            raise ValueError(f"unsupported field dtype {field.dtype}")
        prec, kind = table[field.dtype]
        return (ctypes.c_void_p(field.data_ptr()), prec, kind, field.numel(),
                field.is_cuda)
    import numpy as np
    if not isinstance(field, np.ndarray) or not field.flags["C_CONTIGUOUS"]:
        raise ValueError("field must be a contiguous torch tensor or numpy array")
    table = {np.dtype("<f8"): (F64, IO_REAL), np.dtype("<f4"): (F32, IO_REAL),
             np.dtype("<c16"): (F64, IO_COMPLEX), np.dtype("<c8"): (F32, IO_COMPLEX)}
    if field.dtype not in table:
        raise ValueError(f"unsupported field dtype {field.dtype}")
    prec, kind = table[field.dtype]
    return ctypes.c_void_p(field.ctypes.data), prec, kind, field.size, False


def _arr3(v):
    return (ctypes.c_int * 3)(*v)


def _stream(stream):
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(int(getattr(stream, "cuda_stream", stream)))


class _Brick:
    _what = "BinaryIO"

    def __init__(self):
        self._domain = None

    def set_domain(self, global_size, local_size, offset):
        self._domain = _validate_domain(global_size, local_size, offset,
                                        f"{self._what}::set_domain")

    def _prepare(self, field, op):
        if self._domain is None:
            raise RuntimeError(f"{self._what}::{op}: set_domain() was not called")
        g, l, o = self._domain
        ptr, prec, kind, count, dev = _field_info(field)
        expected = l[0] * l[1] * l[2]
        if count != expected:
            raise RuntimeError(f"{self._what}::{op}: buffer size mismatch (expected "
                               f"{expected} elements from set_domain, got {count})")
        return ptr, prec, kind, dev, _arr3(g), _arr3(l), _arr3(o)


def _check_io(rc):
    if rc == _lib.PFC_HIP_EINVAL:
        raise ValueError(load().pfc_hip_last_error().decode(errors="replace"))
    if rc != _lib.PFC_HIP_OK:
        raise RuntimeError(str(PfcHipError(rc, load().pfc_hip_last_error().decode(
            errors="replace"))))


class BinaryWriter(_Brick):
    """frontend/io/binary_writer.hpp:76-205 for one writer process."""

    _what = "BinaryWriter"

    def __init__(self, filename: str):
        super().__init__()
        self.filename = filename

    def write(self, increment: int, field, stream=None, truncate: bool = True) -> str:
        """Write the brick into format_with_number(filename, increment); returns
        the path.  truncate=False keeps what other bricks wrote to the file."""
        path = format_with_number(self.filename, increment)
        ptr, prec, kind, dev, g, l, o = self._prepare(field, "write")
        lib = load()
        if dev:
            rc = lib.pfc_hip_field_write(path.encode(), ptr, prec, kind, g, l, o,
                                         int(truncate), _stream(stream))
        else:
            rc = lib.pfc_hip_field_write_host(path.encode(), ptr, prec, kind, g, l, o,
                                              int(truncate))
        _check_io(rc)
        return path


class BinaryReader(_Brick):
    """kernel/simulation/binary_reader.hpp:79-200 for one reader process."""

    _what = "BinaryReader"

    def read(self, filename: str, field, stream=None) -> None:
        ptr, prec, kind, dev, g, l, o = self._prepare(field, "read")
        lib = load()
        if dev:
            rc = lib.pfc_hip_field_read(filename.encode(), ptr, prec, kind, g, l, o,
                                        _stream(stream))
        else:
            rc = lib.pfc_hip_field_read_host(filename.encode(), ptr, prec, kind, g, l, o)
        _check_io(rc)


def write_field(path: str, field, global_size, stream=None) -> None:
    """Whole-field convenience: the brick is the global array."""
    w = BinaryWriter(path)
    w.set_domain(global_size, global_size, (0, 0, 0))
    w.write(0, field, stream)


def read_field(path: str, field, global_size, stream=None) -> None:
    r = BinaryReader()
    r.set_domain(global_size, global_size, (0, 0, 0))
    r.read(path, field, stream)
