"""FDTD energy and Poynting-flux functions on the GPU, mirroring meanas/fdtd/energy.py.

- `poynting(e, h, dxes=None)`: the width-weighted E x H with E taken at the next cell
  (`energy.py:10-87`), in the field precision;
- `poynting_divergence(s=None, *, e=None, h=None, dxes=None)`: its backward-difference
  divergence with unit spacing (`energy.py:90-125`);
- `energy_hstep`, `energy_estep`, `delta_energy_h2e`, `delta_energy_e2h`,
  `delta_energy_j`, `dxmul` (`energy.py:128-295`): per-cell energies in float64, as the
  reference's float64 widths make them.

One kernel launch per call (`mx_fdtd_poynting`, `mx_fdtd_poynting_divergence`,
`mx_fdtd_energy`), in the reference's expression order without fused multiply-adds.
Fields are real `(3, nx, ny, nz)` float32/float64 arrays of one dtype; materials are
taken in the field precision.  Given any torch CUDA tensor the result is a CUDA tensor;
given numpy arrays only, it is a new numpy array.
"""
from __future__ import annotations

import ctypes

import numpy
import torch

from .. import _lib
from ..fdfd.device import torch_stream_handle
from .base import _Field, _Grid, _material, _prec


_GRIDS: dict = {}


def _grid(dxes, shape, device) -> _Grid:
    """Device widths for `dxes`, reused across calls with the same widths.  The key is
    their content, since callers may change dxes in place between calls."""
    key = (tuple(shape), str(device))
    if dxes is not None:
        key += tuple(numpy.ascontiguousarray(numpy.real(d), dtype=numpy.float64).tobytes() for side in dxes for d in side)
    g = _GRIDS.get(key)
    if g is None:
        if len(_GRIDS) >= 16:
            _GRIDS.clear()
        g = _GRIDS[key] = _Grid(dxes, shape, device)
    return g


def _device_fields(fields, device):
    host = not any(isinstance(f, torch.Tensor) for f in fields)
    ts = [_Field(f, device).t for f in fields]
    for t in ts[1:]:
        if t.dtype != ts[0].dtype or t.shape != ts[0].shape:
            raise ValueError('fields must share one dtype and shape')
    return ts, host


def _out(t: torch.Tensor, host: bool):
    return t.cpu().numpy() if host else t


def _ptr(t):
    return None if t is None else t.data_ptr()


def _energy(form: int, dt: float, fields, epsilon, mu, dxes, device):
    ts, host = _device_fields(fields, device)
    like = ts[0]
    shape = tuple(like.shape[1:])
    grid = _grid(dxes, shape, like.device)
    me, es = _material(epsilon, like)
    mm, ms = _material(mu, like)
    out = torch.empty(shape, dtype=torch.float64, device=like.device)
    table = (ctypes.c_void_p * 4)(*([t.data_ptr() for t in ts] + [None] * (4 - len(ts))))
    _lib.check(_lib.lib().mx_fdtd_energy(ctypes.byref(grid.desc), _prec(like.dtype), form, float(dt), table, _ptr(me),
                                         es, _ptr(mm), ms, out.data_ptr(), torch_stream_handle()), 'mx_fdtd_energy')
    return _out(out, host)


def poynting(e, h, dxes=None, *, device='cuda'):
    """s_x = ey(i+1) hz - ez(i+1) hy and cyclically, fields times their widths
    (energy.py:10-87)."""
    (et, ht), host = _device_fields((e, h), device)
    grid = _grid(dxes, tuple(et.shape[1:]), et.device)
    s = torch.empty_like(et)
    _lib.check(_lib.lib().mx_fdtd_poynting(ctypes.byref(grid.desc), _prec(et.dtype), et.data_ptr(), ht.data_ptr(),
                                           s.data_ptr(), torch_stream_handle()), 'mx_fdtd_poynting')
    return _out(s, host)


def poynting_divergence(s=None, *, e=None, h=None, dxes=None, device='cuda'):
    """sum_a (s_a - s_a at the previous cell along a), unit spacing (energy.py:90-125);
    computes s = poynting(e, h, dxes) first when s is not given."""
    if s is None:
        assert e is not None
        assert h is not None
        assert dxes is not None
        s = poynting(e, h, dxes=dxes, device=device)
    (st,), host = _device_fields((s,), device)
    shape = tuple(st.shape[1:])
    grid = _grid(None, shape, st.device)
    ds = torch.empty(shape, dtype=st.dtype, device=st.device)
    _lib.check(_lib.lib().mx_fdtd_poynting_divergence(ctypes.byref(grid.desc), _prec(st.dtype), st.data_ptr(),
                                                      ds.data_ptr(), torch_stream_handle()),
               'mx_fdtd_poynting_divergence')
    return _out(ds, host)


def energy_hstep(e0, h1, e2, epsilon=None, mu=None, dxes=None, *, device='cuda'):
    """Energy at the time of h1: dxmul(e0 e2, h1 h1) (energy.py:128-153)."""
    return _energy(_lib.MX_ENERGY_HSTEP, 0.0, (e0, h1, e2), epsilon, mu, dxes, device)


def energy_estep(h0, e1, h2, epsilon=None, mu=None, dxes=None, *, device='cuda'):
    """Energy at the time of e1: dxmul(e1 e1, h0 h2) (energy.py:156-181)."""
    return _energy(_lib.MX_ENERGY_ESTEP, 0.0, (h0, e1, h2), epsilon, mu, dxes, device)


def delta_energy_h2e(dt, e0, h1, e2, h3, epsilon=None, mu=None, dxes=None, *, device='cuda'):
    """Energy change from h1 to e2: dxmul(e2 (e2 - e0) / dt, h1 (h3 - h1) / dt)
    (energy.py:184-214)."""
    return _energy(_lib.MX_ENERGY_H2E, dt, (e0, h1, e2, h3), epsilon, mu, dxes, device)


def delta_energy_e2h(dt, h0, e1, h2, e3, epsilon=None, mu=None, dxes=None, *, device='cuda'):
    """Energy change from e1 to h2: dxmul(e1 (e3 - e1) / dt, h2 (h2 - h0) / dt)
    (energy.py:217-247)."""
    return _energy(_lib.MX_ENERGY_E2H, dt, (h0, e1, h2, e3), epsilon, mu, dxes, device)


def delta_energy_j(j0, e1, dxes=None, *, device='cuda'):
    """(j0 . e1) times the E-cell volumes (energy.py:250-270)."""
    return _energy(_lib.MX_ENERGY_J, 0.0, (j0, e1), None, None, dxes, device)


def dxmul(ee, hh, epsilon=None, mu=None, dxes=None, *, device='cuda'):
    """sum_c ee eps dV_e + sum_c hh mu dV_h (energy.py:273-295)."""
    return _energy(_lib.MX_ENERGY_DXMUL, 0.0, (ee, hh), epsilon, mu, dxes, device)
