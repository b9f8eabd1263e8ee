"""Fields after the solve, on the GPU: H from E and the Poynting cross product.

Mirrors the matrix-free forms of the reference:

- `e2h(omega, dxes, mu=None, pmc=None)` returns `f(E) -> H` with
  `H = curl_forward(E) / (-i omega) / mu`, as `meanas/fdfd/functional.py:89-117`
  (`e2h`); the optional `pmc` mask zeroes H components as the sparse form does
  (`meanas/fdfd/operators.py:253-282`).
- `m2j(omega, dxes, mu=None)` returns `f(M) -> J = curl_back(M / mu) / (-i omega)`, as
  `meanas/fdfd/functional.py:120-151`.  The reference's sparse form
  (`meanas/fdfd/operators.py:285-307`) divides by +i omega instead, so its product with
  M is -J.
- `poynting_e_cross_h(dxes)` returns `f(E, H) -> E x H` with E shifted one cell forward
  and both fields weighted by their cell widths, as `meanas/fdfd/functional.py:187-221`.
  The time-averaged Poynting vector is `Re(f(E, H.conj())) / 2`; `poynting(E, H, dxes)`
  returns it, and equals `(operators.poynting_e_cross(E, dxes) @ H.conj()).real / 2`
  (`meanas/fdfd/operators.py:310-336`), the form the reference's `test_poynting_planes`
  uses (`meanas/test/test_fdfd.py:21-55`).

- `poynting_h_cross(h, dxes)` stands in for the sparse matrix of
  `meanas/fdfd/operators.py:339-363`: `poynting_h_cross(h, dxes) @ e` gives
  `-(E x H)` with the same shift and weights (`.apply(e, conj_e=True)` conjugates E on
  the device).

It also gives the wave operators as callables:

- `curl_forward(dx_e)` / `curl_back(dx_h)` (`meanas/fdmath/functional.py`): the forward
  curl runs the e2h kernel with unit scale; the backward curl is the forward curl of the
  index-reversed field on the reversed widths, negated (R S R = S^-1 for the periodic
  shift turns backward differences into negated forward ones);
- `e_full(omega, dxes, epsilon, mu, pec, pmc)`: the sparse `meanas/fdfd/operators.py:40-96`
  operator (1/mu divides, as on the solve path; the reference's functional form multiplies
  by mu instead), applied as (1/p) A (p E) with the solver's symmetrised FP64 apply;
- `h_full(...)` (`meanas/fdfd/operators.py:126-180`) through the exact duality
  h_full(omega, dxes, eps, mu, pec, pmc) = R e_full(omega, [rev dxh, rev dxe], rev mu,
  rev eps, pec = rev pmc, pmc = rev pec) R;
- `eh_full(...)` (`meanas/fdfd/operators.py:183-250`): `f(E, H) -> (J, -M)` from the two
  curls and the diagonal terms.

Fields may be given as `(3, nx, ny, nz)` arrays (the functional layout) or as `vec()`
vectors of length 3N.  numpy inputs give numpy outputs of the same shape; torch tensors
stay on the device.  The kernels run through the C ABI (`mx_e2h`, `mx_m2j`,
`mx_poynting_e_cross_h`, `mx_poynting_h_cross_e`, `mx_map`, `mx_apply`).
"""
from __future__ import annotations

import ctypes
from typing import Callable, Sequence

import numpy
import torch

from .. import _lib
from .device import DeviceOperator, torch_stream_handle


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


def m2j(omega: complex, dxes, mu=None, device: str | torch.device = 'cuda') -> Callable:
    """`f(M) -> J = curl_back(M / mu) / (-i omega)` (fdfd/functional.py:120-151)."""
    grid = _Grid(dxes, device)
    mu_t = None
    if mu is not None:
        m = numpy.asarray(mu, dtype=numpy.complex128).reshape(-1)
        if m.size != grid.n3:
            raise ValueError('mu must have 3N entries')
        mu_t = torch.from_numpy(m.copy()).to(grid.device)
    w = complex(omega)

    def f(m):
        mt, like = _to_device(m, grid, 'M')
        j = torch.empty_like(mt)
        _lib.check(_lib.lib().mx_m2j(ctypes.byref(grid.desc), w.real, w.imag,
                                     mu_t.data_ptr() if mu_t is not None else None, mt.data_ptr(), j.data_ptr(),
                                     torch_stream_handle()), 'mx_m2j')
        return _from_device(j, like)
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


class _HCross:
    """`poynting_h_cross(h, dxes) @ e` on the device; H is uploaded once."""

    def __init__(self, h, dxes, device):
        self.grid = _Grid(dxes, device)
        self.h, _ = _to_device(h, self.grid, 'H')

    def apply(self, e, conj_e: bool = False):
        et, like = _to_device(e, self.grid, 'E')
        s = torch.empty_like(et)
        _lib.check(_lib.lib().mx_poynting_h_cross_e(ctypes.byref(self.grid.desc), self.h.data_ptr(), et.data_ptr(),
                                                    1 if conj_e else 0, s.data_ptr(), torch_stream_handle()),
                   'mx_poynting_h_cross_e')
        return _from_device(s, like)

    def __matmul__(self, e):
        return self.apply(e)


def poynting_h_cross(h, dxes, device: str | torch.device = 'cuda') -> _HCross:
    """Stand-in for the sparse `poynting_h_cross(h, dxes)` (fdfd/operators.py:339-363):
    `P @ e = -(E x H)`, E shifted one cell forward, both fields width-weighted."""
    return _HCross(h, dxes, device)


def poynting(e, h, dxes, device: str | torch.device = 'cuda'):
    """Time-averaged Poynting vector Re(E x H*) / 2 (real, same layout as `e`)."""
    s = poynting_e_cross_h(dxes, device)(e, h, conj_h=True)
    return s.real / 2


# ---- curls and wave operators -----------------------------------------------------
def _flip(t: torch.Tensor, shape) -> torch.Tensor:
    """R: every cell index reversed, components kept (flat 3N tensor)."""
    return t.reshape(3, *shape).flip(1, 2, 3).reshape(-1).contiguous()


def _curl_f(grid: _Grid, scale: complex) -> Callable[[torch.Tensor], torch.Tensor]:
    """t -> scale * curl_forward(t) on grid.dxe (device tensors); the e2h kernel with
    1 / (-i omega) = scale."""
    omega = 1j / complex(scale)          # 1 / (-i omega) = scale

    def f(t: torch.Tensor) -> torch.Tensor:
        out = torch.empty_like(t)
        _lib.check(_lib.lib().mx_e2h(ctypes.byref(grid.desc), omega.real, omega.imag, None, None,
                                     t.data_ptr(), out.data_ptr(), torch_stream_handle()), 'mx_e2h')
        return out
    return f


def _reversed_grid(dx, device) -> _Grid:
    rev = [numpy.asarray(d)[::-1] for d in dx]
    return _Grid([rev, rev], device)


def curl_forward(dx_e, device: str | torch.device = 'cuda') -> Callable:
    """`f(E) -> curl_forward(E)` with forward differences on dx_e (fdmath/functional.py)."""
    grid = _Grid([dx_e, dx_e], device)
    cf = _curl_f(grid, 1.0)

    def f(e):
        et, like = _to_device(e, grid, 'E')
        return _from_device(cf(et), like)
    return f


def curl_back(dx_h, device: str | torch.device = 'cuda') -> Callable:
    """`f(H) -> curl_back(H)` with backward differences on dx_h (fdmath/functional.py):
    -R curl_forward(rev dx_h) R."""
    grid = _reversed_grid(dx_h, device)
    cf = _curl_f(grid, -1.0)

    def f(h):
        ht, like = _to_device(h, grid, 'H')
        return _from_device(_flip(cf(_flip(ht, grid.shape)), grid.shape), like)
    return f


def _vec_or_none(v):
    return None if v is None else numpy.asarray(v).reshape(-1)


def _rev_vec(v, shape):
    return None if v is None else numpy.ascontiguousarray(numpy.asarray(v).reshape(3, *shape)[:, ::-1, ::-1, ::-1]).reshape(-1)


class _A0:
    """E -> e_full(omega, dxes, epsilon, mu, pec, pmc) @ E on the GPU: (1/p) A (p E)."""

    def __init__(self, omega, dxes, epsilon, mu, pec, pmc, device):
        self.op = DeviceOperator(omega, dxes, numpy.asarray(epsilon).reshape(-1), mu=_vec_or_none(mu),
                                 pec=_vec_or_none(pec), pmc=_vec_or_none(pmc), device=device)

    def __call__(self, x: torch.Tensor) -> torch.Tensor:
        px = self.op.map(_lib.MX_MAP_RHS, x, coef=1.0)
        return self.op.map(_lib.MX_MAP_E, self.op.apply(px))


def _dev_field(f, op: DeviceOperator):
    if isinstance(f, torch.Tensor):
        t = f.to(device=op.device, dtype=torch.complex128).reshape(-1).contiguous()
        like = ('torch', tuple(f.shape))
    else:
        a = numpy.asarray(f)
        t = torch.from_numpy(numpy.ascontiguousarray(a.reshape(-1), dtype=numpy.complex128)).to(op.device)
        like = ('numpy', a.shape)
    if t.numel() != op.n3:
        raise ValueError(f'field has {t.numel()} entries, the grid needs 3N = {op.n3}')
    return t, like


def e_full(omega: complex, dxes, epsilon, mu=None, pec=None, pmc=None,
           device: str | torch.device = 'cuda') -> Callable:
    """`f(E) -> e_full @ E` (fdfd/operators.py:40-96)."""
    A0 = _A0(omega, dxes, epsilon, mu, pec, pmc, device)

    def f(e):
        et, like = _dev_field(e, A0.op)
        return _from_device(A0(et), like)
    return f


def h_full(omega: complex, dxes, epsilon, mu=None, pec=None, pmc=None,
           device: str | torch.device = 'cuda') -> Callable:
    """`f(H) -> h_full @ H` (fdfd/operators.py:126-180), through the reversed dual e_full."""
    shape = tuple(len(d) for d in dxes[0])
    eps = numpy.asarray(epsilon).reshape(-1)
    dual_dxes = [[numpy.asarray(d)[::-1] for d in dxes[1]], [numpy.asarray(d)[::-1] for d in dxes[0]]]
    dual_eps = _rev_vec(mu, shape) if mu is not None else numpy.ones_like(eps)
    A0 = _A0(omega, dual_dxes, dual_eps, _rev_vec(eps, shape), _rev_vec(pmc, shape), _rev_vec(pec, shape), device)

    def f(h):
        ht, like = _dev_field(h, A0.op)
        return _from_device(_flip(A0(_flip(ht, shape)), shape), like)
    return f


def eh_full(omega: complex, dxes, epsilon, mu=None, pec=None, pmc=None,
            device: str | torch.device = 'cuda') -> Callable:
    """`f(E, H) -> (J, -M)`: [[-pe (i w eps) pe, pe ch pm], [pm ce pe, pm (i w mu) pm]]
    applied to (E, H) (fdfd/operators.py:183-250)."""
    ge = _Grid([dxes[0], dxes[0]], device)
    gh = _reversed_grid(dxes[1], device)
    ce = _curl_f(ge, 1.0)
    ch_r = _curl_f(gh, -1.0)
    shape = ge.shape
    dev = ge.device
    w = complex(omega)

    def keep(mask):
        if mask is None:
            return None
        return torch.from_numpy((numpy.asarray(mask).reshape(-1) == 0).astype(numpy.float64)).to(dev)

    pe, pm = keep(pec), keep(pmc)
    eps = torch.from_numpy(numpy.asarray(epsilon, dtype=numpy.complex128).reshape(-1)).to(dev)
    mu_t = None if mu is None else torch.from_numpy(numpy.asarray(mu, dtype=numpy.complex128).reshape(-1)).to(dev)

    def m(v, k):
        return v if k is None else v * k

    def f(e, h):
        et, like = _to_device(e, ge, 'E')
        ht, _ = _to_device(h, ge, 'H')
        pe_e, pm_h = m(et, pe), m(ht, pm)
        ch_h = _flip(ch_r(_flip(pm_h, shape)), shape)
        j = m(-1j * w * eps * pe_e + ch_h, pe)
        iwm_h = 1j * w * (pm_h if mu_t is None else mu_t * pm_h)
        mm = m(ce(pe_e) + iwm_h, pm)
        return _from_device(j, like), _from_device(mm, like)
    return f
