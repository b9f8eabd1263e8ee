"""Convolutional PML for the FDTD updates on the GPU, mirroring meanas/fdtd/pml.py.

- `cpml_params(axis, polarity, dt, thickness=8, ln_R_per_layer=-1.6, epsilon_eff=1,
  mu_eff=1, m=3.5, ma=1, cfs_alpha=0)` returns the layer dict of `pml.py:23-93`:
  profiles `param_e` / `param_h` = (p0, p1, p2) at the E and H sample points, shaped to
  broadcast along `axis`, and the `region` of the face the layer covers.
- `updates_with_cpml(cpml_params, dt, dxes, epsilon, *, dtype=numpy.float32)` returns
  `(update_E(e, h, epsilon), update_H(e, h, mu=(1, 1, 1)))` of `pml.py:96-230`.  The
  auxiliary psi fields live on the device in `dtype` (float32, as the reference's
  default, or float64) and persist between calls; the kernels run through
  `mx_fdtd_cpml_e` / `mx_fdtd_cpml_h`.  As in the reference, both difference directions
  use `dxes[1]`, and the updates return None.

Fields follow `meanas_amd.fdtd.base`: real `(3, nx, ny, nz)` float32/float64, torch CUDA
tensors updated in place, numpy arrays written back in place.
"""
from __future__ import annotations

import ctypes
from typing import Any, Callable, Sequence

import numpy
import torch

from .. import _lib
from ..fdfd.device import torch_stream_handle
from .base import _Field, _Grid, _material, _prec


def cpml_params(axis: int, polarity: int, dt: float, thickness: int = 8, ln_R_per_layer: float = -1.6,
                epsilon_eff: float = 1, mu_eff: float = 1, m: float = 3.5, ma: float = 1,
                cfs_alpha: float = 0) -> dict[str, Any]:
    """Profiles of one CPML face (pml.py:23-93).

    Sample depth x (cells from the inner face) of layer cell l: polarity +1 has E at
    l + 1/2 and H at l + 1; polarity -1 is the mirror image (E at t - l, H at t - l - 1/2).
    With s = (x/t)^m: sigma = s sigma_max, kappa = 1 + s (kappa_max - 1), alpha =
    (1 - x/t)^ma cfs_alpha; p0 = exp(-(sigma/kappa + alpha) dt), p1 = sigma / (sigma +
    kappa alpha) (p0 - 1), p2 = 1/kappa; sigma_max = -ln_R_per_layer / 2 (m + 1) and
    kappa_max = sqrt(epsilon_eff mu_eff).
    """
    if axis not in range(3):
        raise ValueError(f'Invalid axis: {axis}')
    if polarity not in (-1, 1):
        raise ValueError(f'Invalid polarity: {polarity}')
    if thickness <= 2:
        raise ValueError('It would be wise to have a pml with 4+ cells of thickness')
    if epsilon_eff <= 0:
        raise ValueError('epsilon_eff must be positive')
    t = thickness
    cells = numpy.arange(t, dtype=float)
    depth_e, depth_h = (cells + 0.5, cells + 1.0) if polarity > 0 else (t - cells, t - cells - 0.5)
    sigma_max = -ln_R_per_layer / 2 * (m + 1)
    kappa_max = numpy.sqrt(epsilon_eff * mu_eff)
    bshape = [1, 1, 1]
    bshape[axis] = t

    def profiles(x):
        frac = (x / t) ** m
        sigma = frac * sigma_max
        kappa = 1 + frac * (kappa_max - 1)
        alpha = ((1 - x / t) ** ma) * cfs_alpha
        p0 = numpy.exp(-(sigma / kappa + alpha) * dt)
        p1 = sigma / (sigma + kappa * alpha) * (p0 - 1)
        return tuple(q.reshape(bshape) for q in (p0, p1, 1 / kappa))

    region = [slice(None)] * 3
    region[axis] = slice(None, t) if polarity < 0 else slice(-t, None)
    return {'param_e': profiles(depth_e), 'param_h': profiles(depth_h), 'region': tuple(region)}


def _face_polarity(region, axis: int, t: int, n: int) -> int:
    """-1 / +1 for a region covering the first / last t cells along `axis` (the regions
    cpml_params makes); anything else has no kernel form."""
    if len(region) != 3 or any(region[a] != slice(None) for a in range(3) if a != axis):
        raise ValueError('a CPML region must span the whole grid across its axis')
    start, stop, step = region[axis].indices(n)
    if step == 1 and stop - start == t:
        if start == 0:
            return -1
        if stop == n:
            return 1
    raise ValueError('a CPML region must be the first or last `thickness` cells of its axis')


def updates_with_cpml(cpml_params: Sequence[Sequence[dict[str, Any] | None]], dt: float, dxes, epsilon, *,
                      dtype=numpy.float32, device: str | torch.device = 'cuda') -> tuple[Callable, Callable]:
    """`(update_E(e, h, epsilon), update_H(e, h, mu=(1, 1, 1)))` with the psi state on the
    device (pml.py:96-230).  cpml_params[axis][p] is a layer dict for polarity -1 (p = 0)
    and +1 (p = 1), or None."""
    dev = torch.device(device)
    shape = tuple(numpy.shape(epsilon)[1:]) if not isinstance(epsilon, torch.Tensor) else tuple(epsilon.shape[1:])
    if len(shape) != 3:
        raise ValueError('epsilon must have shape (3, nx, ny, nz)')
    psi_dtype = numpy.dtype(dtype)
    if psi_dtype == numpy.float32:
        psi_prec, psi_torch = _lib.MX_REAL32, torch.float32
    elif psi_dtype == numpy.float64:
        psi_prec, psi_torch = _lib.MX_REAL64, torch.float64
    else:
        raise TypeError(f'psi dtype must be float32 or float64, got {psi_dtype}')
    grid = _Grid(dxes, shape, dev)
    keep = []   # device arrays the layer table points at
    table = (_lib.MxCpmlLayer * _lib.MX_CPML_MAX_LAYERS)()
    n = 0
    for axis in range(3):
        for pp in range(2):
            c = cpml_params[axis][pp]
            if c is None:
                continue
            pe = [numpy.asarray(q, dtype=numpy.float64).ravel() for q in c['param_e']]
            ph = [numpy.asarray(q, dtype=numpy.float64).ravel() for q in c['param_h']]
            t = pe[0].size
            if any(q.size != t for q in pe + ph):
                raise ValueError('CPML profiles must all have the layer thickness')
            pol = _face_polarity(c['region'], axis, t, shape[axis])
            if pol != (-1, 1)[pp]:
                raise ValueError(f'cpml_params[{axis}][{pp}] covers the other face of axis {axis}')
            cells = int(numpy.prod(shape)) // shape[axis] * t
            arrs = [torch.from_numpy(numpy.concatenate(pe)).to(dev), torch.from_numpy(numpy.concatenate(ph)).to(dev),
                    torch.zeros(2 * cells, dtype=psi_torch, device=dev),
                    torch.zeros(2 * cells, dtype=psi_torch, device=dev)]
            keep.extend(arrs)
            table[n] = _lib.MxCpmlLayer(axis, pol, t, 0, *(a.data_ptr() for a in arrs))
            n += 1

    def check_shape(fe):
        if tuple(fe.t.shape[1:]) != shape:
            raise ValueError(f'field shape {tuple(fe.t.shape[1:])} differs from the CPML grid {shape}')

    def update_E(e, h, epsilon) -> None:
        fe, fh = _Field(e, dev), _Field(h, dev)
        check_shape(fe)
        if fe.t.dtype != fh.t.dtype:
            raise TypeError('e and h must have the same dtype')
        m, ms = _material(epsilon, fe.t)
        _lib.check(_lib.lib().mx_fdtd_cpml_e(ctypes.byref(grid.desc), _prec(fe.t.dtype), psi_prec, float(dt),
                                             fe.t.data_ptr(), fh.t.data_ptr(), m.data_ptr() if m is not None else None,
                                             ms, table, n, torch_stream_handle()), 'mx_fdtd_cpml_e')
        fe.done()

    def update_H(e, h, mu=(1, 1, 1)) -> None:
        fe, fh = _Field(e, dev), _Field(h, dev)
        check_shape(fh)
        if fe.t.dtype != fh.t.dtype:
            raise TypeError('e and h must have the same dtype')
        if isinstance(mu, (tuple, list)) and len(mu) == 3 and all(numpy.isscalar(x) for x in mu):
            if mu[0] == mu[1] == mu[2]:
                mu = float(mu[0])
            else:   # per-component constants: expand to a field
                mu = torch.tensor([float(x) for x in mu], dtype=fh.t.dtype, device=dev).view(3, 1, 1, 1)
                mu = mu.expand(fh.t.shape).contiguous()
        m, ms = _material(mu, fh.t)
        _lib.check(_lib.lib().mx_fdtd_cpml_h(ctypes.byref(grid.desc), _prec(fh.t.dtype), psi_prec, float(dt),
                                             fe.t.data_ptr(), fh.t.data_ptr(), m.data_ptr() if m is not None else None,
                                             ms, table, n, torch_stream_handle()), 'mx_fdtd_cpml_h')
        fh.done()

    update_E.psi_state = update_H.psi_state = keep   # keeps the device arrays alive
    return update_E, update_H
