"""Seeded synthetic FDFD workloads (host-side numpy inputs for generic()).

  plain_48      48^3, dx = 1, omega = 2 pi / 38.75, 10-cell SC-PML on all faces
                (ln_R = -16, m = 4), eps = 1 with an 8x8-cell core of 12.25 along x,
                J = unit Jy dipole at the centre.
  boxes_128     128^3, same omega / PML, eps = 2.085 + 40 random boxes
                (eps ~ U[2, 12.25], sides U{4..24}), J = Gaussian Jy sheet (sigma 6) at x = 20.
  aniso_multi   128x96x96, diagonal anisotropic eps (independent box sets per
                component), 16 Gaussian-sheet right-hand sides with random complex
                polarisation at random x-planes in [15, 110].
  uneven_batch  32 structures, edges from {48, 64, 80, 96, 128}, box dielectrics
                eps in [1, 12.25], 5% of cells PEC (2-6 cell cubes), omega = 2 pi / U[30, 45],
                one dipole each.
  serving       a Poisson request stream: edges lognormal clipped to [48, 256], 30%
                adjoint, dielectrics and sources as aniso_multi.
Each generator returns plain dicts of numpy arrays in the reference's layout
(`vec` of (3, nx, ny, nz), C order; dxes = [[dx_e, dy_e, dz_e], [dx_h, dy_h, dz_h]]).
"""
from __future__ import annotations

import numpy

from .fdfd.scpml import stretch_with_scpml

OMEGA_1550 = 2 * numpy.pi / 38.75      # lambda = 1550 nm at dx = 40 nm


def pml_dxes(shape, omega: float, thickness: int = 10, epsilon_effective: float = 1.0):
    dxes = [[numpy.ones(n) for n in shape] for _ in range(2)]
    for a, n in enumerate(shape):
        if n > 2 * thickness:
            for p in (-1, 1):
                stretch_with_scpml(dxes, axis=a, polarity=p, omega=omega, thickness=thickness,
                                   epsilon_effective=epsilon_effective)
    return dxes


def random_boxes(rng: numpy.random.Generator, shape, nbox: int, lo: float, hi: float,
                 smin: int = 4, smax: int = 24):
    """nbox axis-aligned boxes (lo/hi corners, clipped to the grid) with values U[lo, hi]."""
    out = []
    for _ in range(nbox):
        val = rng.uniform(lo, hi)
        sides = rng.integers(smin, smax + 1, size=3)
        corner = [int(rng.integers(0, n)) for n in shape]
        sl = tuple(slice(c, min(c + s, n)) for c, s, n in zip(corner, sides, shape))
        out.append((sl, val))
    return out


def paint(shape, background: float, boxes) -> numpy.ndarray:
    eps = numpy.full(shape, background, dtype=numpy.float64)
    for sl, val in boxes:
        eps[sl] = val
    return eps


def gaussian_sheet(shape, x0: int, sigma: float, pol) -> numpy.ndarray:
    nx, ny, nz = shape
    j = numpy.arange(ny)[:, None] - ny // 2
    k = numpy.arange(nz)[None, :] - nz // 2
    prof = numpy.exp(-(j ** 2 + k ** 2) / (2 * sigma ** 2))
    J = numpy.zeros((3, *shape), dtype=complex)
    for c in range(3):
        J[c, x0] = pol[c] * prof
    return J


def plain_48():
    shape = (48, 48, 48)
    omega = OMEGA_1550
    eps = numpy.ones((3, *shape))
    eps[:, :, 20:28, 20:28] = 12.25
    J = numpy.zeros((3, *shape), dtype=complex)
    J[1, 24, 24, 24] = 1.0
    return dict(name='plain_48', shape=shape, omega=omega, dxes=pml_dxes(shape, omega),
                epsilon=eps.astype(complex).ravel(), J=J.ravel(), mu=None, pec=None, pmc=None)


def boxes_128(n: int = 128, seed: int = 2):
    shape = (n, n, n)
    omega = OMEGA_1550
    rng = numpy.random.default_rng(seed)
    e = paint(shape, 2.085, random_boxes(rng, shape, 40, 2.0, 12.25))
    eps = numpy.broadcast_to(e, (3, *shape)).copy()
    J = numpy.zeros((3, *shape), dtype=complex)
    c = n // 2
    j = numpy.arange(n)[:, None] - c
    k = numpy.arange(n)[None, :] - c
    J[1, min(20, n - 1)] = numpy.exp(-(j ** 2 + k ** 2) / (2 * 6.0 ** 2))
    return dict(name=f'boxes_{n}', shape=shape, omega=omega, dxes=pml_dxes(shape, omega),
                epsilon=eps.ravel(), J=J.ravel(), mu=None, pec=None, pmc=None)


def random_polarisation(rng: numpy.random.Generator) -> numpy.ndarray:
    v = rng.standard_normal(3) + 1j * rng.standard_normal(3)
    v /= numpy.linalg.norm(v)
    return v * numpy.exp(1j * rng.uniform(0, 2 * numpy.pi))


def aniso_structure(shape, rng: numpy.random.Generator, nbox: int = 40):
    comps = [paint(shape, 2.085, random_boxes(rng, shape, nbox, 2.0, 12.25)) for _ in range(3)]
    return numpy.stack(comps)


def aniso_multi(shape=(128, 96, 96), nrhs: int = 16, seed: int = 3):
    omega = OMEGA_1550
    rng = numpy.random.default_rng(seed)
    eps = aniso_structure(shape, rng)
    Js = []
    lo, hi = 15, min(110, shape[0] - 16)
    for _ in range(nrhs):
        x0 = int(rng.integers(lo, hi + 1))
        Js.append(gaussian_sheet(shape, x0, 6.0, random_polarisation(rng)).ravel())
    return dict(name='aniso_multi', shape=shape, omega=omega, dxes=pml_dxes(shape, omega),
                epsilon=eps.ravel(), Js=Js, mu=None, pec=None, pmc=None)


def pec_cubes(rng: numpy.random.Generator, shape, fraction: float = 0.05, protect=None):
    mask = numpy.zeros(shape, dtype=bool)
    total = mask.size
    while mask.sum() < fraction * total:
        s = int(rng.integers(2, 7))
        c = [int(rng.integers(0, max(1, n - s))) for n in shape]
        mask[c[0]:c[0] + s, c[1]:c[1] + s, c[2]:c[2] + s] = True
    if protect is not None:
        mask &= ~protect
    return mask


def uneven_batch(count: int = 32, seed: int = 4, edges=(48, 64, 80, 96, 128)):
    rng = numpy.random.default_rng(seed)
    out = []
    for s in range(count):
        shape = tuple(int(rng.choice(edges)) for _ in range(3))
        omega = 2 * numpy.pi / rng.uniform(30, 45)
        e = paint(shape, 1.0, random_boxes(rng, shape, 20, 1.0, 12.25))
        eps = numpy.broadcast_to(e, (3, *shape)).copy()
        pec3 = pec_cubes(rng, shape)
        # a dipole at a random non-PEC cell away from the PML
        while True:
            p = [int(rng.integers(12, n - 12)) for n in shape]
            if not pec3[tuple(p)]:
                break
        J = numpy.zeros((3, *shape), dtype=complex)
        J[int(rng.integers(0, 3)), p[0], p[1], p[2]] = 1.0
        pec = numpy.broadcast_to(pec3, (3, *shape)).astype(numpy.float64).ravel()
        out.append(dict(name=f'uneven_{s}', shape=shape, omega=omega, dxes=pml_dxes(shape, omega),
                        epsilon=eps.ravel(), J=J.ravel(), mu=None, pec=pec, pmc=None))
    return out


def serving_request(rng: numpy.random.Generator, lo: int = 48, hi: int = 256, median: float = 72.0,
                    sigma: float = 0.35, adjoint_fraction: float = 0.3):
    shape = tuple(int(numpy.clip(round(rng.lognormal(numpy.log(median), sigma)), lo, hi)) for _ in range(3))
    omega = OMEGA_1550
    eps = aniso_structure(shape, rng, nbox=max(4, int(40 * numpy.prod(shape) / 128 ** 3)))
    x0 = int(rng.integers(12, shape[0] - 12))
    J = gaussian_sheet(shape, x0, 6.0, random_polarisation(rng))
    return dict(name='serving', shape=shape, omega=omega, dxes=pml_dxes(shape, omega),
                epsilon=eps.ravel(), J=J.ravel(), mu=None, pec=None, pmc=None,
                adjoint=bool(rng.uniform() < adjoint_fraction))


def serving_stream(n_requests: int, qps: float, seed: int = 5, **kw):
    rng = numpy.random.default_rng(seed)
    t = 0.0
    reqs = []
    for _ in range(n_requests):
        t += rng.exponential(1 / qps)
        r = serving_request(rng, **kw)
        r['arrival'] = t
        reqs.append(r)
    return reqs
