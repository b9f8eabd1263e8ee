"""Stretched-coordinate PML cell widths (host side, FP64).

Same interface and semantics as meanas.fdfd.scpml (reference meanas/fdfd/scpml.py):
`prepare_s_function` (:18-38), `uniform_grid_scpml` (:41-95) and `stretch_with_scpml`
(:98-167).  These are tiny 1D computations that prepare the `dxes` input of the solve;
they stay on the host in FP64.
"""
from __future__ import annotations

from typing import Callable, Sequence

import numpy

s_function_t = Callable[[numpy.ndarray], numpy.ndarray]


def prepare_s_function(ln_R: float = -16, m: float = 4) -> s_function_t:
    """Polynomial conductivity profile s(d) = (m + 1) ln_R / 2 * d**m (scpml.py:18-38)."""
    peak = (m + 1) * ln_R / 2       # halved: the grid is periodic, the wave crosses twice

    def s_factor(distance: numpy.ndarray) -> numpy.ndarray:
        return peak * numpy.power(distance, m)
    return s_factor


def uniform_grid_scpml(shape: Sequence[int], thicknesses: Sequence[int], omega: float,
                       epsilon_effective: float = 1.0,
                       s_function: s_function_t | None = None) -> list[list[numpy.ndarray]]:
    """Unit-width grid with a PML of `thicknesses[a]` cells on both ends of axis a (scpml.py:41-95)."""
    if s_function is None:
        s_function = prepare_s_function()
    scale = numpy.sqrt(epsilon_effective) * numpy.real(omega)
    dx_e, dx_h = [], []
    for n, th in zip(tuple(shape), tuple(thicknesses)):
        if th > 0:
            u = numpy.arange(n, dtype=numpy.float64)

            def depth(v):  # normalised distance into the nearer PML
                return (numpy.clip(th - v, 0, None) + numpy.clip(v - (n - th), 0, None)) / th
            dx_e.append(1 + 1j * s_function(depth(u)) / scale)
            dx_h.append(1 + 1j * s_function(depth(u + 0.5)) / scale)
        else:
            dx_e.append(numpy.ones(n))
            dx_h.append(numpy.ones(n))
    return [dx_e, dx_h]


def stretch_with_scpml(dxes: list[list[numpy.ndarray]], axis: int, polarity: int, omega: float,
                       epsilon_effective: float = 1.0, thickness: int = 10,
                       s_function: s_function_t | None = None) -> list[list[numpy.ndarray]]:
    """Add an SC-PML on one face (polarity +1: low end, -1: high end) of one axis.

    Mutates `dxes[0][axis]` and `dxes[1][axis]` and returns `dxes` (scpml.py:98-167).
    """
    if s_function is None:
        s_function = prepare_s_function()
    w_e = numpy.asarray(dxes[0][axis]).astype(complex)
    w_h = numpy.asarray(dxes[1][axis]).astype(complex)
    x = numpy.zeros(w_e.size + 1, dtype=complex)
    numpy.cumsum(w_e, out=x[1:])
    mid = (x[:-1] + x[1:]) / 2          # E-cell centres
    left = x[:-1]                        # H-cell positions
    scale = numpy.sqrt(epsilon_effective) * numpy.real(omega)
    if polarity > 0:
        inner = x[thickness]
        width = inner - x[0]
        sel = slice(thickness)
        frac_e = (inner - mid[sel]) / width
        frac_h = (inner - left[sel]) / width
    else:
        inner = x[-thickness - 1]
        width = x[-1] - inner
        sel = slice(None) if thickness == 0 else slice(-thickness, None)
        frac_e = (mid[sel] - inner) / width
        frac_h = (left[sel] - inner) / width
    w_e[sel] *= 1 + 1j * s_function(frac_e) / width / scale
    w_h[sel] *= 1 + 1j * s_function(frac_h) / width / scale
    dxes[0][axis] = w_e
    dxes[1][axis] = w_h
    return dxes
