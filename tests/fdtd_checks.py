"""The reference's FDTD tests (meanas/test/test_fdtd.py) restated over its fixture grid
(meanas/test/conftest.py), with the time stepping passed in, so the same checks run on
the oracle's updates and on the GPU's.  The energy functions the checks use are the
oracle's unless the simulation names others (`TDResult.en`)."""
from dataclasses import dataclass, field
from types import SimpleNamespace

import numpy

from oracle import fdfd_ref as ref

PRNG = numpy.random.RandomState(12345)


@dataclass
class TDResult:
    shape: tuple
    dt: float
    dxes: list
    epsilon: numpy.ndarray
    es: list = field(default_factory=list)
    hs: list = field(default_factory=list)
    js: list = field(default_factory=list)
    en: object = None   # the energy functions the checks use (the oracle's unless set)


ORACLE_ENERGY = SimpleNamespace(energy_hstep=ref.energy_hstep, energy_estep=ref.energy_estep,
                                delta_energy_j=ref.delta_energy_j, poynting_divergence=ref.poynting_divergence,
                                poynting=ref.fdtd_poynting)


def cases():
    """(shape, eps_kind, bg, fg, dx, dxes_kind, j_kind) as the reference's fixtures combine
    them, with its 3D skips."""
    out = []
    for shape in [(5, 5, 1), (5, 1, 5), (5, 5, 5)]:
        is3d = 1 not in shape
        for eps_kind in ['center', '000', 'random']:
            for bg in [1.0, 1.5]:
                for fg in [1.0, 2.5]:
                    if is3d and (eps_kind == '000' or bg != 1 or fg not in (1.0, 2.0)):
                        continue
                    for dx in [1.0, 1.5]:
                        for dxk in ['uniform', 'centerbig']:
                            for jk in ['center', 'random']:
                                out.append((shape, eps_kind, bg, fg, dx, dxk, jk))
    return out


def make_sim(case, step, dt=0.3, seed=12345):
    """10 steps from zero fields with J added at steps 0, 4 and 8 (test_fdtd.py:183-235)."""
    shape3, eps_kind, bg, fg, dx, dxk, jk = case
    rng = numpy.random.RandomState(seed)
    shape = (3, *shape3)
    eps = numpy.full(shape, bg, dtype=float)
    if eps_kind == 'center':
        eps[:, shape3[0] // 2, shape3[1] // 2, shape3[2] // 2] = fg
    elif eps_kind == '000':
        eps[:, 0, 0, 0] = fg
    else:
        eps[:] = rng.uniform(low=min(bg, fg), high=max(bg, fg), size=shape)
    dxes = [[numpy.full(s, dx) for s in shape3] for _ in range(2)]
    if dxk == 'centerbig':
        for eh in (0, 1):
            for ax in (0, 1, 2):
                dxes[eh][ax][dxes[eh][ax].size // 2] *= 1.1
    j = numpy.zeros(shape)
    if jk == 'center':
        j[:, shape3[0] // 2, shape3[1] // 2, shape3[2] // 2] = 1.0
    else:
        j[:] = rng.uniform(low=-1.0, high=1.0, size=shape)
    sim = TDResult(shape=shape, dt=dt, dxes=dxes, epsilon=eps)
    e = numpy.zeros(shape)
    h = numpy.zeros(shape)
    jz = numpy.zeros(shape)
    for tt in range(10):
        e = e.copy()
        h = h.copy()
        e, h = step(e, h, eps, dt, dxes)
        if tt in (0, 4, 8):
            e += j / eps
            sim.js.append(j)
        else:
            sim.js.append(jz)
        sim.es.append(e)
        sim.hs.append(h)
    return sim


def oracle_step(e, h, eps, dt, dxes):
    ref.maxwell_h(dt, dxes)(e, h)
    ref.maxwell_e(dt, dxes)(e, h, eps)
    return e, h


def close(x, y):
    numpy.testing.assert_allclose(x, y)


def check_initial_fields(sim):
    e0, h0, j0 = sim.es[0], sim.hs[0], sim.js[0]
    mask = j0 != 0
    close(e0[mask], j0[mask] / sim.epsilon[mask])
    assert not e0[~mask].any()
    assert not h0.any()


def E(sim):
    return sim.en or ORACLE_ENERGY


def check_initial_energy(sim):
    j0, e0, h0, h1 = sim.js[0], sim.es[0], sim.hs[0], sim.hs[1]
    dV = numpy.prod(numpy.meshgrid(*sim.dxes[0], indexing='ij'), axis=0)
    u0 = (j0 * j0.conj() / sim.epsilon * dV).sum(axis=0)
    close(E(sim).energy_estep(h0=h0, e1=e0, h2=h1, dxes=sim.dxes, epsilon=sim.epsilon), u0)
    close(E(sim).delta_energy_j(j0=j0, e1=e0, dxes=sim.dxes), u0)


def check_energy_conservation(sim):
    u = E(sim).delta_energy_j(j0=sim.js[0], e1=sim.es[0], dxes=sim.dxes).sum()
    for ii in range(1, 8):
        u_h = E(sim).energy_hstep(e0=sim.es[ii - 1], h1=sim.hs[ii], e2=sim.es[ii], dxes=sim.dxes, epsilon=sim.epsilon)
        u_e = E(sim).energy_estep(h0=sim.hs[ii], e1=sim.es[ii], h2=sim.hs[ii + 1], dxes=sim.dxes, epsilon=sim.epsilon)
        u += E(sim).delta_energy_j(j0=sim.js[ii], e1=sim.es[ii - 1], dxes=sim.dxes).sum()
        close(u_h.sum(), u)
        u += E(sim).delta_energy_j(j0=sim.js[ii], e1=sim.es[ii], dxes=sim.dxes).sum()
        close(u_e.sum(), u)


def check_poynting_divergence(sim):
    u_prev = None
    for ii in range(1, 8):
        u_h = E(sim).energy_hstep(e0=sim.es[ii - 1], h1=sim.hs[ii], e2=sim.es[ii], dxes=sim.dxes, epsilon=sim.epsilon)
        u_e = E(sim).energy_estep(h0=sim.hs[ii], e1=sim.es[ii], h2=sim.hs[ii + 1], dxes=sim.dxes, epsilon=sim.epsilon)
        dj_b = E(sim).delta_energy_j(j0=sim.js[ii], e1=sim.es[ii], dxes=sim.dxes)
        close(u_e - u_h - dj_b, -sim.dt * E(sim).poynting_divergence(e=sim.es[ii], h=sim.hs[ii], dxes=sim.dxes))
        if u_prev is not None:
            dj_a = E(sim).delta_energy_j(j0=sim.js[ii], e1=sim.es[ii - 1], dxes=sim.dxes)
            close(u_h - u_prev - dj_a,
                  -sim.dt * E(sim).poynting_divergence(e=sim.es[ii - 1], h=sim.hs[ii], dxes=sim.dxes))
        u_prev = u_e


def check_poynting_planes(sim):
    mask = (sim.js[0] != 0).any(axis=0)
    if mask.sum() > 1:
        return            # the reference skips distributed sources here
    mx, my, mz = (numpy.roll(mask, -1, axis=a) for a in range(3))
    u_prev = None
    for ii in range(1, 8):
        u_h = E(sim).energy_hstep(e0=sim.es[ii - 1], h1=sim.hs[ii], e2=sim.es[ii], dxes=sim.dxes, epsilon=sim.epsilon)
        u_e = E(sim).energy_estep(h0=sim.hs[ii], e1=sim.es[ii], h2=sim.hs[ii + 1], dxes=sim.dxes, epsilon=sim.epsilon)
        dj_b = E(sim).delta_energy_j(j0=sim.js[ii], e1=sim.es[ii], dxes=sim.dxes)
        s = -E(sim).poynting(e=sim.es[ii], h=sim.hs[ii], dxes=sim.dxes) * sim.dt
        planes = [s[0, mask].sum(), -s[0, mx].sum(), s[1, mask].sum(), -s[1, my].sum(),
                  s[2, mask].sum(), -s[2, mz].sum()]
        close(sum(planes), (u_e - u_h - dj_b)[mask])
        if u_prev is not None:
            dj_a = E(sim).delta_energy_j(j0=sim.js[ii], e1=sim.es[ii - 1], dxes=sim.dxes)
            s = -E(sim).poynting(e=sim.es[ii - 1], h=sim.hs[ii], dxes=sim.dxes) * sim.dt
            planes = [s[0, mask].sum(), -s[0, mx].sum(), s[1, mask].sum(), -s[1, my].sum(),
                      s[2, mask].sum(), -s[2, mz].sum()]
            close(sum(planes), (u_h - u_prev - dj_a)[mask])
        u_prev = u_e


CHECKS = [check_initial_fields, check_initial_energy, check_energy_conservation,
          check_poynting_divergence, check_poynting_planes]


# ---- CPML (meanas/fdtd/pml.py; the reference has no tests of it) ----------------------
def cpml_layers(make, dt, faces, thickness, **kw):
    """cpml_params[axis][p] list with a layer on each (axis, polarity) in `faces`."""
    out = [[None, None] for _ in range(3)]
    for axis, pol in faces:
        out[axis][(pol + 1) // 2] = make(axis, pol, dt, thickness=thickness, **kw)
    return out


ALL_FACES = [(a, p) for a in range(3) for p in (-1, 1)]


def pulse_energy(updates, axis, comp, n=160, thickness=12, steps=500, dt=0.5, layers=True, to_dev=None, to_host=None,
                 make=None):
    """A Gaussian pulse in e[comp] on an n-cell line along `axis` (unit widths, eps = 1),
    stepped H then E `steps` times: (final, max) field energy relative to the initial one."""
    shape = [1, 1, 1]
    shape[axis] = n
    shape = tuple(shape)
    dxes = [[numpy.ones(s) for s in shape] for _ in range(2)]
    eps = numpy.ones((3, *shape))
    faces = [(axis, -1), (axis, 1)] if layers else []
    uE, uH = updates(cpml_layers(make, dt, faces, thickness), dt, dxes, eps)
    e = numpy.zeros((3, *shape))
    h = numpy.zeros_like(e)
    e[comp] = numpy.exp(-((numpy.arange(n).reshape(shape) - n / 2) / 5.0) ** 2)
    u0 = (e ** 2).sum()
    if to_dev is not None:
        e, h, eps = to_dev(e), to_dev(h), to_dev(eps)
    top = 0.0
    for _ in range(steps):
        uH(e, h)
        uE(e, h, eps)
        if to_host is not None and _ % 50:
            continue
        ee, hh = (e, h) if to_host is None else (to_host(e), to_host(h))
        top = max(top, float((ee ** 2).sum() + (hh ** 2).sum()) / u0)
    ee, hh = (e, h) if to_host is None else (to_host(e), to_host(h))
    return float((ee ** 2).sum() + (hh ** 2).sum()) / u0, top


# ---- conducting boundary: the image (mirror) construction ---------------------------
def mirror_fields(direction, n, other=(3, 4), seed=0, dtype=numpy.float64):
    """e, h, eps on a periodic domain of 2 (n - 2) planes along `direction`, symmetric
    about the plane at cell 1/2 (and so, by the period, about n - 3/2).  Along the
    direction, tangential E and normal H sit at whole cells (plane i at i) and are even;
    normal E and tangential H sit at half cells (plane i at i + 1/2) and are odd; eps
    follows its E component with even parity."""
    period = 2 * (n - 2)
    others = iter(other)
    shape = [period if a == direction else next(others) for a in range(3)]
    rng = numpy.random.default_rng(seed)
    e = rng.standard_normal((3, *shape))
    h = rng.standard_normal((3, *shape))
    eps = 1 + rng.random((3, *shape))
    idx = numpy.arange(period)
    image = {True: (1 - idx) % period, False: (-idx) % period}   # whole / half cells

    def sym(f, c, whole, sign):
        f[c] = (f[c] + sign * numpy.take(f[c], image[whole], axis=direction)) / 2

    for c in range(3):
        normal = c == direction
        sym(e, c, not normal, -1 if normal else 1)
        sym(h, c, normal, 1 if normal else -1)
        sym(eps, c, not normal, 1)
    return tuple(a.astype(dtype) for a in (e, h, eps))


def mirror_mismatch(direction, n, step_pair, boundaries, steps=8, seed=0, dtype=numpy.float64, to_dev=None,
                    to_host=None, mutate=None):
    """Largest difference, over `steps` leapfrog steps, between a run on planes 0..n-1
    with both conducting faces of `direction` applied after each update and a plain
    periodic run on the doubled mirror-symmetric domain.  step_pair(e, h, eps, after_h,
    after_e) does one step (H update, after_h(), E update, after_e());
    boundaries(direction, polarity) -> (update_e, update_h); `mutate(e, h)`
    plants a change after the boundaries (for checking that the comparison sees it)."""
    to_dev = to_dev or (lambda a: a)
    to_host = to_host or (lambda a: a)
    e, h, eps = mirror_fields(direction, n, seed=seed, dtype=dtype)
    take = lambda a: numpy.ascontiguousarray(numpy.take(a, range(n), axis=1 + direction))
    eb, hb, epsb = (to_dev(take(a)) for a in (e, h, eps))
    e, h, eps = to_dev(e), to_dev(h), to_dev(eps)
    en, hn = boundaries(direction, -1)
    ep, hp = boundaries(direction, 1)
    worst = 0.0
    for _ in range(steps):
        step_pair(e, h, eps, lambda: None, lambda: None)
        step_pair(eb, hb, epsb, lambda: (hn(hb), hp(hb)), lambda: (en(eb), ep(eb)))
        if mutate is not None:
            mutate(eb, hb)
        for full, part in ((e, eb), (h, hb)):
            diff = numpy.abs(take(to_host(full)) - to_host(part)).max()
            worst = max(worst, float(diff))
    return worst
