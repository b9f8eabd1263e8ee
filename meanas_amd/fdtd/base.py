"""FDTD field updates on the GPU, mirroring meanas/fdtd/base.py.

- `maxwell_e(dt, dxes=None)` returns `f(e, h, epsilon) -> e` doing `e += dt curl_back(h) / epsilon`
  in place (`meanas/fdtd/base.py:13-64`);
- `maxwell_h(dt, dxes=None)` returns `f(e, h, mu=None) -> h` doing `h -= dt curl_forward(e) / mu`
  in place (`meanas/fdtd/base.py:67-125`);
- `run(e, h, epsilon, dt, dxes, steps, mu)` does `steps` leapfrog steps (H then E) in one
  call without returning to Python between kernels.

Fields are real `(3, nx, ny, nz)` arrays, float64 or float32.  torch CUDA tensors are
updated in place on the device; numpy arrays are copied to the device, updated and
written back in place (so a caller holding the array sees the new values, as with the
reference).  The kernels run through the C ABI (`mx_fdtd_h`, `mx_fdtd_e`, `mx_fdtd_run`).
"""
from __future__ import annotations

import ctypes
from typing import Callable

import numpy
import torch

from .. import _lib
from ..fdfd.device import torch_stream_handle


class _Grid:
    """Device copies of the real cell widths (or NULL for unit widths)."""

    def __init__(self, dxes, shape, device):
        self.device = torch.device(device)
        self.arrays = []
        ptrs = [[None] * 3, [None] * 3]
        if dxes is not None:
            for s in range(2):
                for a in range(3):
                    t = torch.from_numpy(numpy.ascontiguousarray(numpy.real(dxes[s][a]), dtype=numpy.float64))
                    t = t.to(self.device)
                    if t.numel() != shape[a]:
                        raise ValueError('dxes do not match the field shape')
                    self.arrays.append(t)
                    ptrs[s][a] = t.data_ptr()
        self.desc = _lib.MxFdtdGrid(shape[0], shape[1], shape[2], 0, (ctypes.c_void_p * 3)(*ptrs[0]),
                                    (ctypes.c_void_p * 3)(*ptrs[1]))


def _prec(dtype) -> int:
    if dtype in (torch.float64, numpy.float64):
        return _lib.MX_REAL64
    if dtype in (torch.float32, numpy.float32):
        return _lib.MX_REAL32
    raise TypeError(f'FDTD fields must be float32 or float64, got {dtype}')


class _Field:
    """A field on the device; numpy inputs are written back by `done()`."""

    def __init__(self, f, device, dtype=None):
        self.src = f
        if isinstance(f, torch.Tensor):
            if not f.is_cuda or not f.is_contiguous():
                raise ValueError('torch fields must be contiguous CUDA tensors (updated in place)')
            self.t = f
        else:
            a = numpy.asarray(f)
            self.t = torch.from_numpy(numpy.ascontiguousarray(a)).to(device)
        if self.t.dim() != 4 or self.t.shape[0] != 3:
            raise ValueError('fields must have shape (3, nx, ny, nz)')

    def done(self):
        if not isinstance(self.src, torch.Tensor):
            self.src[...] = self.t.cpu().numpy()
        return self.src


def _material(m, like: torch.Tensor):
    """(device tensor or None, scalar) for a material that may be a scalar or a field."""
    if m is None:
        return None, 1.0
    if numpy.isscalar(m) or (isinstance(m, numpy.ndarray) and m.ndim == 0):
        return None, float(m)
    t = m if isinstance(m, torch.Tensor) else torch.from_numpy(numpy.ascontiguousarray(m))
    t = t.to(device=like.device, dtype=like.dtype).contiguous()
    if t.shape != like.shape:
        raise ValueError('material shape must match the field shape')
    return t, 1.0


def maxwell_e(dt: float, dxes=None, device: str | torch.device = 'cuda') -> Callable:
    """`f(e, h, epsilon) -> e` with `e += dt curl_back(h) / epsilon` (in place)."""
    grids = {}

    def me(e, h, epsilon):
        fe, fh = _Field(e, device), _Field(h, device)
        shape = tuple(fe.t.shape[1:])
        if shape not in grids:
            grids[shape] = _Grid(dxes, shape, device)
        m, ms = _material(epsilon, fe.t)
        if fe.t.dtype != fh.t.dtype:
            raise TypeError('e and h must have the same dtype')
        _lib.check(_lib.lib().mx_fdtd_e(ctypes.byref(grids[shape].desc), _prec(fe.t.dtype), float(dt), fe.t.data_ptr(),
                                        fh.t.data_ptr(), m.data_ptr() if m is not None else None, ms,
                                        torch_stream_handle()), 'mx_fdtd_e')
        return fe.done()
    return me


def maxwell_h(dt: float, dxes=None, device: str | torch.device = 'cuda') -> Callable:
    """`f(e, h, mu=None) -> h` with `h -= dt curl_forward(e) / mu` (in place)."""
    grids = {}

    def mh(e, h, mu=None):
        fe, fh = _Field(e, device), _Field(h, device)
        shape = tuple(fh.t.shape[1:])
        if shape not in grids:
            grids[shape] = _Grid(dxes, shape, device)
        m, ms = _material(mu, fh.t)
        if fe.t.dtype != fh.t.dtype:
            raise TypeError('e and h must have the same dtype')
        _lib.check(_lib.lib().mx_fdtd_h(ctypes.byref(grids[shape].desc), _prec(fh.t.dtype), float(dt), fe.t.data_ptr(),
                                        fh.t.data_ptr(), m.data_ptr() if m is not None else None, ms,
                                        torch_stream_handle()), 'mx_fdtd_h')
        return fh.done()
    return mh


def run(e, h, epsilon, dt: float, dxes=None, steps: int = 1, mu=None, device: str | torch.device = 'cuda'):
    """`steps` leapfrog steps (maxwell_h then maxwell_e) in place; returns (e, h)."""
    fe, fh = _Field(e, device), _Field(h, device)
    if fe.t.dtype != fh.t.dtype:
        raise TypeError('e and h must have the same dtype')
    grid = _Grid(dxes, tuple(fe.t.shape[1:]), device)
    me, mes = _material(epsilon, fe.t)
    mm, mms = _material(mu, fh.t)
    _lib.check(_lib.lib().mx_fdtd_run(ctypes.byref(grid.desc), _prec(fe.t.dtype), float(dt), fe.t.data_ptr(),
                                      fh.t.data_ptr(), me.data_ptr() if me is not None else None, mes,
                                      mm.data_ptr() if mm is not None else None, mms, int(steps),
                                      torch_stream_handle()), 'mx_fdtd_run')
    return fe.done(), fh.done()
