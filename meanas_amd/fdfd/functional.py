"""Fields after the solve, on the GPU: H from E and the Poynting cross product.

Mirrors the matrix-free forms of the reference:

- `e2h(omega, dxes, mu=None, pmc=None)` returns `f(E) -> H` with
  `H = curl_forward(E) / (-i omega) / mu`, as `meanas/fdfd/functional.py:89-117`
  (`e2h`); the optional `pmc` mask zeroes H components as the sparse form does
  (`meanas/fdfd/operators.py:253-282`).
- `poynting_e_cross_h(dxes)` returns `f(E, H) -> E x H` with E shifted one cell forward
  and both fields weighted by their cell widths, as `meanas/fdfd/functional.py:187-221`.
  The time-averaged Poynting vector is `Re(f(E, H.conj())) / 2`; `poynting(E, H, dxes)`
  returns it, and equals `(operators.poynting_e_cross(E, dxes) @ H.conj()).real / 2`
  (`meanas/fdfd/operators.py:310-336`), the form the reference's `test_poynting_planes`
  uses (`meanas/test/test_fdfd.py:21-55`).

Fields may be given as `(3, nx, ny, nz)` arrays (the functional layout) or as `vec()`
vectors of length 3N.  numpy inputs give numpy outputs of the same shape; torch tensors
stay on the device.  Both kernels run through the C ABI (`mx_e2h`,
`mx_poynting_e_cross_h`).
"""
from __future__ import annotations

import ctypes
from typing import Callable, Sequence

import numpy
import torch

from .. import _lib
from .device import torch_stream_handle


class _Grid:
    """Device copies of dxes and the mx_grid descriptor that points at them."""

    def __init__(self, dxes: Sequence[Sequence[numpy.ndarray]], device):
        self.shape = tuple(len(d) for d in dxes[0])
        if tuple(len(d) for d in dxes[1]) != self.shape:
            raise ValueError('dxes[0] and dxes[1] describe different grids')
        self.device = torch.device(device)
        self.arrays = [torch.from_numpy(numpy.ascontiguousarray(d, dtype=numpy.complex128)).to(self.device)
                       for d in (*dxes[0], *dxes[1])]
        nx, ny, nz = self.shape
        self.desc = _lib.MxGrid(nx, ny, nz, 0,
                                (ctypes.c_void_p * 3)(*[a.data_ptr() for a in self.arrays[:3]]),
                                (ctypes.c_void_p * 3)(*[a.data_ptr() for a in self.arrays[3:]]))
        self.n3 = 3 * nx * ny * nz


def _to_device(f, grid: _Grid, name: str) -> tuple[torch.Tensor, object]:
    """(flat complex128 device tensor, how to shape the result like the input)."""
    if isinstance(f, torch.Tensor):
        t = f.to(device=grid.device, dtype=torch.complex128).reshape(-1).contiguous()
        like = ('torch', tuple(f.shape))
    else:
        a = numpy.asarray(f)
        t = torch.from_numpy(numpy.ascontiguousarray(a.reshape(-1), dtype=numpy.complex128)).to(grid.device)
        like = ('numpy', a.shape)
    if t.numel() != grid.n3:
        raise ValueError(f'{name} has {t.numel()} entries, the grid needs 3N = {grid.n3}')
    return t, like


def _from_device(t: torch.Tensor, like):
    kind, shape = like
    return t.reshape(shape) if kind == 'torch' else t.cpu().numpy().reshape(shape)


def e2h(omega: complex, dxes, mu=None, pmc=None, device: str | torch.device = 'cuda') -> Callable:
    """`f(E) -> H = curl_forward(E) / (-i omega) / mu` (zero on PMC components)."""
    grid = _Grid(dxes, device)
    inv_mu = None
    if mu is not None:
        m = numpy.asarray(mu, dtype=numpy.complex128).reshape(-1)
        if m.size != grid.n3:
            raise ValueError('mu must have 3N entries')
        inv_mu = torch.from_numpy(1.0 / m).to(grid.device)
    mask = None
    if pmc is not None:
        p = numpy.asarray(pmc).reshape(-1)
        if p.size != grid.n3:
            raise ValueError('pmc must have 3N entries')
        mask = torch.from_numpy((p != 0).astype(numpy.uint8)).to(grid.device)
    w = complex(omega)

    def f(e):
        et, like = _to_device(e, grid, 'E')
        h = torch.empty_like(et)
        _lib.check(_lib.lib().mx_e2h(ctypes.byref(grid.desc), w.real, w.imag,
                                     inv_mu.data_ptr() if inv_mu is not None else None,
                                     mask.data_ptr() if mask is not None else None,
                                     et.data_ptr(), h.data_ptr(), torch_stream_handle()), 'mx_e2h')
        return _from_device(h, like)
    return f


def poynting_e_cross_h(dxes, device: str | torch.device = 'cuda') -> Callable:
    """`f(E, H) -> E x H` (E shifted one cell forward, width-weighted fields)."""
    grid = _Grid(dxes, device)

    def f(e, h, conj_h: bool = False):
        et, like = _to_device(e, grid, 'E')
        ht, _ = _to_device(h, grid, 'H')
        s = torch.empty_like(et)
        _lib.check(_lib.lib().mx_poynting_e_cross_h(ctypes.byref(grid.desc), et.data_ptr(), ht.data_ptr(),
                                                    1 if conj_h else 0, s.data_ptr(), torch_stream_handle()),
                   'mx_poynting_e_cross_h')
        return _from_device(s, like)
    return f


def poynting(e, h, dxes, device: str | torch.device = 'cuda'):
    """Time-averaged Poynting vector Re(E x H*) / 2 (real, same layout as `e`)."""
    s = poynting_e_cross_h(dxes, device)(e, h, conj_h=True)
    return s.real / 2
