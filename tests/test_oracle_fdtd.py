"""The oracle's FDTD updates and energy functions pinned by the reference's FDTD tests
(meanas/test/test_fdtd.py), and its functional curls against the sparse ones."""
import numpy
import pytest

from oracle import fdfd_ref as ref
import fdtd_checks as fc


@pytest.mark.parametrize('case', fc.cases())
def test_reference_fdtd_checks_on_oracle(case):
    sim = fc.make_sim(case, fc.oracle_step)
    for check in fc.CHECKS:
        check(sim)


def test_functional_curls_equal_sparse():
    rng = numpy.random.default_rng(4)
    shape = (4, 5, 6)
    dxes = [[0.5 + rng.random(n) for n in shape] for _ in range(2)]
    f = rng.standard_normal((3, *shape))
    numpy.testing.assert_allclose(ref.vec(ref.f_curl_forward(dxes[0])(f)), ref.curl_forward(dxes[0]) @ ref.vec(f),
                                  rtol=1e-12, atol=1e-12)
    numpy.testing.assert_allclose(ref.vec(ref.f_curl_back(dxes[1])(f)), ref.curl_back(dxes[1]) @ ref.vec(f),
                                  rtol=1e-12, atol=1e-12)


# ---- CPML: the reference ships no test of meanas/fdtd/pml.py, so the oracle is pinned
# by a reduction to the plain Yee updates, by the stretched-grid equivalence of the
# kappa profile, by the orientation of the profiles, and by pulse absorption.
def _random_state(shape, seed):
    rng = numpy.random.default_rng(seed)
    e, h = rng.standard_normal((3, *shape)), rng.standard_normal((3, *shape))
    eps = 1 + rng.random((3, *shape))
    dx = [0.7 + 0.6 * rng.random(n) for n in shape]
    return e, h, eps, dx


def test_cpml_without_layers_is_plain_yee():
    """No layers, equal E and H widths: the plain maxwell_h / maxwell_e updates."""
    shape = (7, 6, 5)
    e, h, eps, dx = _random_state(shape, 1)
    dxes = [dx, dx]
    uE, uH = ref.updates_with_cpml([[None, None]] * 3, 0.3, dxes, eps, dtype=numpy.float64)
    e2, h2 = e.copy(), h.copy()
    for _ in range(3):
        uH(e, h)
        uE(e, h, eps)
        ref.maxwell_h(0.3, dxes)(e2, h2)
        ref.maxwell_e(0.3, dxes)(e2, h2, eps)
    numpy.testing.assert_allclose(e, e2, rtol=1e-12, atol=1e-12)
    numpy.testing.assert_allclose(h, h2, rtol=1e-12, atol=1e-12)


def test_cpml_profile_orientation_and_depth():
    """Polarity -1 mirrors polarity +1 with E and H swapped; at full depth sigma and
    kappa reach their maxima (p0 = exp(-sigma_max/kappa_max dt), p1 = p0 - 1)."""
    t, dt, m, lnr = 7, 0.4, 3.5, -1.6
    kw = dict(thickness=t, epsilon_eff=2.0, mu_eff=1.5, cfs_alpha=0.05)
    lo, hi = ref.cpml_params(1, -1, dt, **kw), ref.cpml_params(1, 1, dt, **kw)
    for a, b in zip(lo['param_e'], hi['param_h']):
        numpy.testing.assert_allclose(a.ravel(), b.ravel()[::-1], rtol=1e-15)
    for a, b in zip(lo['param_h'], hi['param_e']):
        numpy.testing.assert_allclose(a.ravel(), b.ravel()[::-1], rtol=1e-15)
    assert lo['param_e'][0].shape == (1, t, 1)
    assert lo['region'][1] == slice(None, t) and hi['region'][1] == slice(-t, None)
    smax, kmax = -lnr / 2 * (m + 1), numpy.sqrt(3.0)
    p0 = numpy.exp(-smax / kmax * dt)
    numpy.testing.assert_allclose([p.ravel()[-1] for p in hi['param_h']], [p0, p0 - 1, 1 / kmax], rtol=1e-14)
    with pytest.raises(Exception):
        ref.cpml_params(0, 1, dt, thickness=2)
    with pytest.raises(Exception):
        ref.cpml_params(3, 1, dt)


def test_cpml_kappa_only_is_a_stretched_grid():
    """sigma = 0 (ln_R = 0) with alpha > 0 everywhere: psi stays zero and the layers only
    divide the derivatives along their axis by kappa, i.e. the plain updates on widths
    multiplied by kappa inside each layer (E: kappa at the E points on the backward
    differences; H: at the H points on the forward ones; a cell in two layers gets both).
    The short z axis has overlapping layers."""
    shape = (12, 10, 5)
    e, h, eps, dx = _random_state(shape, 2)
    dt = 0.25
    kw = dict(ln_R_per_layer=0.0, epsilon_eff=4.0, cfs_alpha=0.3, ma=0)
    layers = fc.cpml_layers(ref.cpml_params, dt, fc.ALL_FACES, 3, **kw)
    uE, uH = ref.updates_with_cpml(layers, dt, [dx, dx], eps, dtype=numpy.float64)
    dxe = [d.copy() for d in dx]
    dxh = [d.copy() for d in dx]
    for axis in range(3):
        for lay in layers[axis]:
            reg = lay['region'][axis]
            dxe[axis][reg] /= lay['param_e'][2].ravel()
            dxh[axis][reg] /= lay['param_h'][2].ravel()
    assert any(not numpy.allclose(a, b) for a, b in zip(dxe, dx))
    e2, h2 = e.copy(), h.copy()
    for _ in range(4):
        uH(e, h)
        uE(e, h, eps)
        ref.maxwell_h(dt, [dxh, None])(e2, h2)
        ref.maxwell_e(dt, [None, dxe])(e2, h2, eps)
    numpy.testing.assert_allclose(e, e2, rtol=1e-11, atol=1e-12)
    numpy.testing.assert_allclose(h, h2, rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize('axis,comp', [(a, c) for a in range(3) for c in range(3) if c != a])
def test_cpml_absorbs_a_pulse(axis, comp):
    """A pulse on a line terminated by layers at both ends leaves through them (a wrong
    sign, a swapped u/v pairing or a missing decay factor keeps or amplifies it); with no
    layers the periodic line keeps its energy."""
    upd = lambda p, dt, dxes, eps: ref.updates_with_cpml(p, dt, dxes, eps)
    final, top = fc.pulse_energy(upd, axis, comp, make=ref.cpml_params)
    assert final < 1e-8 and top < 1.05
    final0, _ = fc.pulse_energy(upd, axis, comp, make=ref.cpml_params, layers=False, steps=200)
    assert final0 > 0.9


# ---- conducting boundary (meanas/fdtd/boundaries.py): no reference test exists; the
# oracle is pinned by the image construction (a bounded run equals a periodic run on the
# doubled mirror-symmetric domain).
def oracle_pair(dt=0.3):
    mh, me = ref.maxwell_h(dt), ref.maxwell_e(dt)

    def step(e, h, eps, after_h, after_e):
        mh(e, h)
        after_h()
        me(e, h, eps)
        after_e()
    return step


@pytest.mark.parametrize('direction', [0, 1, 2])
@pytest.mark.parametrize('n', [4, 5, 9])
def test_conducting_boundary_is_a_mirror(direction, n):
    assert fc.mirror_mismatch(direction, n, oracle_pair(), ref.conducting_boundary, seed=10 * direction + n) == 0.0


def _plane(d, i):
    sl = [slice(None)] * 3
    sl[d] = i
    return tuple(sl)


@pytest.mark.parametrize('mutant', ['sign', 'plane', 'component', 'h_normal'])
def test_mirror_check_sees_planted_mistakes(mutant):
    """Each plausible mistake in a face update, planted after the oracle's boundaries,
    breaks the mirror equality."""
    d, u = 1, 0

    def mutate(e, h):
        if mutant == 'sign':          # E normal on the last plane not negated
            e[d][_plane(d, -1)] = e[d][_plane(d, -3)]
        elif mutant == 'plane':       # E tangential copied from the wrong plane
            e[u][_plane(d, -1)] = e[u][_plane(d, -3)]
        elif mutant == 'component':   # tangential H zeroed on the normal component too
            h[d][_plane(d, 0)] = 0
        else:                         # H normal not copied at the high face
            h[d][_plane(d, -1)] = h[d][_plane(d, -3)]
    assert fc.mirror_mismatch(d, 6, oracle_pair(), ref.conducting_boundary, mutate=mutate) > 1e-3


def test_reference_indexing_needs_tuples():
    """The reference indexes e[direction] with a list of slices; numpy >= 1.23 rejects
    that, so the oracle uses the tuple the list spells."""
    a = numpy.zeros((3, 4, 5))
    with pytest.raises(IndexError):
        a[[slice(None), slice(None), 0]] = 1
    ep, hp = ref.conducting_boundary(2, 1)
    f = numpy.arange(3 * 2 * 3 * 4, dtype=float).reshape(3, 2, 3, 4)
    g = ep(f.copy())
    assert (g[2][:, :, 2] == 0).all() and (g[2][:, :, 3] == -f[2][:, :, 1]).all()
    assert (g[0][:, :, 3] == f[0][:, :, 2]).all() and (g[:, :, :, :2] == f[:, :, :, :2]).all()
    with pytest.raises(Exception, match='Invalid direction'):
        ref.conducting_boundary(3, 1)
    with pytest.raises(Exception, match='Bad polarity'):
        ref.conducting_boundary(0, 0)


def test_energy_deltas_are_differences_of_step_energies():
    """dt delta_energy_h2e(e0, h1, e2, h3) = energy_estep(h1, e2, h3) - energy_hstep(e0, h1, e2)
    and dt delta_energy_e2h(h0, e1, h2, e3) = energy_hstep(e1, h2, e3) - energy_estep(h0, e1, h2):
    dxmul is linear in each argument.  A swapped field or sign in either delta fails."""
    rng = numpy.random.default_rng(8)
    shape = (4, 5, 3)
    f = [rng.standard_normal((3, *shape)) for _ in range(4)]
    eps, mu = 1 + rng.random((3, *shape)), 1 + rng.random((3, *shape))
    dxes = [[0.5 + rng.random(n) for n in shape] for _ in range(2)]
    dt = 0.37
    kw = dict(epsilon=eps, mu=mu, dxes=dxes)
    h2e = ref.delta_energy_h2e(dt, *f, **kw) * dt
    numpy.testing.assert_allclose(h2e, ref.energy_estep(f[1], f[2], f[3], **kw) - ref.energy_hstep(f[0], f[1], f[2], **kw),
                                  rtol=1e-12, atol=1e-12)
    e2h = ref.delta_energy_e2h(dt, *f, **kw) * dt
    numpy.testing.assert_allclose(e2h, ref.energy_hstep(f[1], f[2], f[3], **kw) - ref.energy_estep(f[0], f[1], f[2], **kw),
                                  rtol=1e-12, atol=1e-12)
    assert ref.fdtd_poynting(f[0].astype(numpy.float32), f[1].astype(numpy.float32), dxes).dtype == numpy.float32
