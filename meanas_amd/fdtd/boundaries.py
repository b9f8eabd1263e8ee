"""Conducting boundary for the FDTD updates on the GPU, mirroring meanas/fdtd/boundaries.py.

`conducting_boundary(direction, polarity)` returns `(update_e, update_h)` for the face at
the low (polarity < 0) or high (polarity > 0) end of axis `direction`
(`boundaries.py:12-78`).  Each takes the field, rewrites the face's planes in place
through `mx_fdtd_conducting_boundary`, and returns it:
- polarity < 0: E normal on plane 0 set to 0 and E tangential copied from plane 1;
  H normal copied from plane 1 and H tangential set to 0;
- polarity > 0: E normal on plane n-2 set to 0 and on plane n-1 to minus plane n-3,
  E tangential on n-1 copied from n-2; H normal on n-1 copied from n-2, H tangential on
  n-2 set to 0 and on n-1 to minus plane n-3.

These are the image values of a mirror plane at cell 1/2 (or n - 3/2) across which
tangential E and normal H are even and normal E and tangential H odd.  The reference
indexes with lists of slices (`e[direction][boundary_slice]`), which numpy >= 1.23
rejects with an IndexError; the updates here are what the same code does with tuple
indices.

Fields follow `meanas_amd.fdtd.base`: real `(3, nx, ny, nz)` float32/float64, torch CUDA
tensors updated in place, numpy arrays written back in place.
"""
from __future__ import annotations

import ctypes
from typing import Callable

import torch

from .. import _lib
from ..fdfd.device import torch_stream_handle
from .base import _Field, _prec


def conducting_boundary(direction: int, polarity: int, *,
                        device: str | torch.device = 'cuda') -> tuple[Callable, Callable]:
    """`(update_e, update_h)` of the conducting face (boundaries.py:12-78).  As in the
    reference, a direction outside 0..2 or a zero polarity raises."""
    if direction not in (0, 1, 2):
        raise Exception(f'Invalid direction: {direction}')
    if not polarity < 0 and not polarity > 0:
        raise Exception(f'Bad polarity: {polarity}')
    pol = -1 if polarity < 0 else 1
    dev = torch.device(device)

    def make(kind: int) -> Callable:
        def update(f):
            ff = _Field(f, dev)
            n = ff.t.shape[1 + direction]
            if n < (2 if pol < 0 else 3):
                raise IndexError(f'axis {direction} has {n} cells; the face needs {2 if pol < 0 else 3}')
            grid = _lib.MxFdtdGrid(*ff.t.shape[1:], 0, (ctypes.c_void_p * 3)(), (ctypes.c_void_p * 3)())
            _lib.check(_lib.lib().mx_fdtd_conducting_boundary(ctypes.byref(grid), _prec(ff.t.dtype), direction, pol,
                                                              kind, ff.t.data_ptr(), torch_stream_handle()),
                       'mx_fdtd_conducting_boundary')
            return ff.done()
        return update

    return make(_lib.MX_FIELD_E), make(_lib.MX_FIELD_H)
