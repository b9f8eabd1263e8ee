"""Source construction on the GPU: TFSF and boundary sources, one-way mode sources.

The step before the solve: these turn a desired E field into the current J that
`generic()` takes.

- `e_tfsf_source(TF_region, omega, dxes, epsilon, mu)` returns `f(E) -> J` with
  `J = (A0 Q E - Q A0 E) / (-i omega)`, A0 = e_full and Q = diag(TF_region)
  (`meanas/fdfd/operators.py:366-394`; the callable form of
  `meanas/fdfd/functional.py:154-184`).
- `e_boundary_source(mask, ...)` keeps that current only on the cells where a one-cell
  shift of the mask (mirror or periodic edges) changes it
  (`meanas/fdfd/operators.py:397-450`).
- `expand_e` and `compute_source` build the one-way source of a mode field
  (`meanas/fdfd/waveguide_3d.py:113-156,206-252`).

A0 is applied on the GPU through the symmetrised operator: with the factored p of the
device operator, A = P A0 P^-1, so A0 v = (1/p) A (p v) -- two diagonal maps around one
FP64 apply (`mx_map`, `mx_apply`).  The masks are computed on the host (a few byte
arrays); the fields and the applies stay on the device.
"""
from __future__ import annotations

from typing import Callable, Sequence

import numpy
import torch

from .. import _lib
from .device import DeviceOperator


def _field_mask(mask, n3: int, ncell: int) -> numpy.ndarray:
    """The 3N float64 values of a mask given as a 3N vector, an (nx, ny, nz) grid mask or a
    (3, nx, ny, nz) field mask.  Values are kept as they are: Q = diag(TF_region) uses
    them (fdfd/operators.py:392-393), so a fractional or weighted region scales E."""
    m = numpy.asarray(mask)
    if numpy.iscomplexobj(m):
        raise ValueError('a mask must be real')
    if m.size == n3:
        return numpy.ascontiguousarray(m.reshape(-1), dtype=numpy.float64)
    if m.size == ncell:
        return numpy.tile(numpy.asarray(m.reshape(-1), dtype=numpy.float64), 3)
    raise ValueError(f'mask must have N = {ncell} or 3N = {n3} entries')


class _A0:
    """y = e_full(omega, dxes, epsilon, mu) @ x on the GPU (complex128)."""

    def __init__(self, omega, dxes, epsilon, mu, device):
        self.op = DeviceOperator(omega, dxes, epsilon, mu=mu, device=device)

    def __call__(self, x: torch.Tensor) -> torch.Tensor:
        px = self.op.map(_lib.MX_MAP_RHS, x, coef=1.0)          # p x
        return self.op.map(_lib.MX_MAP_E, self.op.apply(px))    # (1/p) A (p x)


def _to_dev(e, op: DeviceOperator) -> tuple[torch.Tensor, object]:
    if isinstance(e, torch.Tensor):
        t = e.to(device=op.device, dtype=torch.complex128).reshape(-1).contiguous()
        like = ('torch', tuple(e.shape))
    else:
        a = numpy.asarray(e)
        t = torch.from_numpy(numpy.ascontiguousarray(a.reshape(-1), dtype=numpy.complex128)).to(op.device)
        like = ('numpy', a.shape)
    if t.numel() != op.n3:
        raise ValueError(f'field has {t.numel()} entries, the grid needs 3N = {op.n3}')
    return t, like


def _from_dev(t: torch.Tensor, like):
    kind, shape = like
    return t.reshape(shape) if kind == 'torch' else t.cpu().numpy().reshape(shape)


def _masked_source(A0: _A0, q: torch.Tensor, keep: torch.Tensor | None, omega: complex) -> Callable:
    c = 1.0 / (-1j * complex(omega))

    def f(e):
        et, like = _to_dev(e, A0.op)
        j = (A0(q * et) - q * A0(et)) * c
        if keep is not None:
            j = j * keep
        return _from_dev(j, like)
    return f


def e_tfsf_source(TF_region, omega: complex, dxes, epsilon, mu=None,
                  device: str | torch.device = 'cuda') -> Callable:
    """`f(E) -> J = (A0 Q E - Q A0 E) / (-i omega)` (total-field / scattered-field source)."""
    A0 = _A0(omega, dxes, epsilon, mu, device)
    q = _field_mask(TF_region, A0.op.n3, A0.op.ncell)
    return _masked_source(A0, torch.from_numpy(q).to(A0.op.device), None, omega)


def boundary_mask(mask, shape: Sequence[int], periodic_mask_edges: bool = False) -> numpy.ndarray:
    """3N 0/1 mask of the entries where shifting `mask` by one cell along any axis longer
    than 1 (either way) changes its value; edges mirror unless periodic
    (operators.py:420-445: |(S - I) mask| != 0 per component)."""
    n = int(numpy.prod(shape))
    m = _field_mask(mask, 3 * n, n).reshape(3, *shape)
    out = numpy.zeros(m.shape, dtype=bool)
    for axis in range(3):
        if shape[axis] == 1:
            continue
        for pol in (-1, 1):
            if periodic_mask_edges:
                s = numpy.roll(m, -pol, axis=axis + 1)
            else:
                idx = numpy.arange(shape[axis]) + pol
                idx = numpy.where(idx >= shape[axis], 2 * shape[axis] - idx - 1, idx)
                idx = numpy.where(idx < 0, -1 - idx, idx)
                s = numpy.take(m, idx, axis=axis + 1)
            out |= s != m
    return out.reshape(-1).astype(numpy.float64)


def e_boundary_source(mask, omega: complex, dxes, epsilon, mu=None, periodic_mask_edges: bool = False,
                      device: str | torch.device = 'cuda') -> Callable:
    """`f(E) -> J`: the TFSF current of `mask`, kept only on the mask's edges."""
    A0 = _A0(omega, dxes, epsilon, mu, device)
    shape = A0.op.shape
    q = _field_mask(mask, A0.op.n3, A0.op.ncell)
    keep = boundary_mask(q, shape, periodic_mask_edges)
    dev = A0.op.device
    return _masked_source(A0, torch.from_numpy(q).to(dev), torch.from_numpy(keep).to(dev), omega)


def expand_e(E, wavenumber: complex, dxes, axis: int, polarity: int, slices: Sequence[slice]) -> numpy.ndarray:
    """The slice `slices` of E repeated along `axis` with the propagation phase
    exp(-i polarity k r), r the E-grid position from the slice (waveguide_3d.py:206-252)."""
    slices = tuple(slices)
    E = numpy.asarray(E)
    pos = numpy.real(numpy.asarray(dxes[0][axis])).cumsum()
    r = pos - pos[slices[axis]]
    shape = [1, 1, 1, 1]
    shape[axis + 1] = -1
    phase = numpy.exp(-1j * polarity * wavenumber * r).reshape(shape)
    out = numpy.zeros_like(E, dtype=complex)
    full = list(slices)
    full[axis] = slice(E.shape[axis + 1])
    out[(slice(None), *full)] = phase * E[(slice(None), *slices)]
    return out


def compute_source(E, wavenumber: complex, omega: complex, dxes, axis: int, polarity: int,
                   slices: Sequence[slice], epsilon, mu=None, device: str | torch.device = 'cuda') -> numpy.ndarray:
    """One-way source for the mode field E at `slices` (waveguide_3d.py:113-156): the
    boundary source of the half space that starts at the slice, applied to the field
    expanded along `axis`.  Returns J with E's (3, nx, ny, nz) shape."""
    Ex = expand_e(E, wavenumber, dxes, axis, polarity, slices)
    half = [slice(None)] * 4
    half[axis + 1] = slice(slices[axis].start, None) if polarity > 0 else slice(None, slices[axis].stop)
    mask = numpy.zeros(Ex.shape, dtype=numpy.float64)
    mask[tuple(half)] = 1.0
    eps = numpy.asarray(epsilon).reshape(-1)
    mu_v = None if mu is None else numpy.asarray(mu).reshape(-1)
    f = e_boundary_source(mask.reshape(-1), omega, dxes, eps, mu_v, device=device)
    return f(Ex.reshape(-1)).reshape(Ex.shape)
