"""GPU FDTD updates: the reference's FDTD tests (meanas/test/test_fdtd.py) with the time
stepping on the GPU, and the stepped fields against the oracle's."""
import numpy
import pytest
import torch

from oracle import fdfd_ref as ref
import fdtd_checks as fc

pytestmark = pytest.mark.gpu


def gpu_step(e, h, eps, dt, dxes):
    from meanas_amd import fdtd
    fdtd.maxwell_h(dt, dxes)(e, h)
    fdtd.maxwell_e(dt, dxes)(e, h, eps)
    return e, h


@pytest.mark.parametrize('case', fc.cases())
def test_reference_fdtd_checks_on_gpu(case):
    sim = fc.make_sim(case, gpu_step)
    for check in fc.CHECKS:
        check(sim)
    ora = fc.make_sim(case, fc.oracle_step)
    for a, b in zip(sim.es + sim.hs, ora.es + ora.hs):
        numpy.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-13)


def _fields(shape, seed, dtype):
    rng = numpy.random.default_rng(seed)
    e = rng.standard_normal((3, *shape)).astype(dtype)
    h = rng.standard_normal((3, *shape)).astype(dtype)
    eps = (1 + rng.random((3, *shape))).astype(dtype)
    mu = (1 + rng.random((3, *shape))).astype(dtype)
    dxes = [[0.8 + 0.4 * rng.random(n) for n in shape] for _ in range(2)]
    return e, h, eps, mu, dxes


@pytest.mark.parametrize('shape', [(9, 10, 70), (1, 12, 5), (7, 1, 1), (40, 7, 100), (33, 5, 64), (17, 3, 130)])
@pytest.mark.parametrize('steps', [5, 6])
@pytest.mark.parametrize('dtype', [numpy.float64, numpy.float32])
def test_run_matches_steps_and_oracle(shape, steps, dtype):
    """The fused runner (one pass per step over a ping-pong pair; slabs, partial y and z
    tiles, odd step counts) equals the separate H and E updates bit for bit."""
    from meanas_amd import fdtd
    e, h, eps, mu, dxes = _fields(shape, 1, dtype)
    dt = 0.2
    et = torch.from_numpy(e.copy()).cuda()
    ht = torch.from_numpy(h.copy()).cuda()
    fdtd.run(et, ht, eps, dt, dxes, steps=steps, mu=mu)
    e2, h2 = e.copy(), h.copy()
    me, mh = fdtd.maxwell_e(dt, dxes), fdtd.maxwell_h(dt, dxes)
    for _ in range(steps):
        mh(e2, h2, mu)
        me(e2, h2, eps)
    numpy.testing.assert_array_equal(et.cpu().numpy(), e2)   # same expressions, same order
    numpy.testing.assert_array_equal(ht.cpu().numpy(), h2)
    if dtype == numpy.float32:
        return
    e3, h3 = e.copy(), h.copy()
    for _ in range(steps):
        ref.maxwell_h(dt, dxes)(e3, h3, mu)
        ref.maxwell_e(dt, dxes)(e3, h3, eps)
    numpy.testing.assert_allclose(e2, e3, rtol=1e-11, atol=1e-12)
    numpy.testing.assert_allclose(h2, h3, rtol=1e-11, atol=1e-12)


def test_unit_widths_scalar_materials_and_float32():
    from meanas_amd import fdtd
    shape = (8, 6, 66)
    e, h, eps, mu, dxes = _fields(shape, 2, numpy.float64)
    dt = 0.25
    e64, h64 = e.copy(), h.copy()
    fdtd.run(e64, h64, 2.0, dt, None, steps=5, mu=1.5)          # numpy in, updated in place
    e1, h1 = e.copy(), h.copy()                                 # the separate updates, bit for bit
    me, mh = fdtd.maxwell_e(dt), fdtd.maxwell_h(dt)
    for _ in range(5):
        mh(e1, h1, 1.5)
        me(e1, h1, 2.0)
    numpy.testing.assert_array_equal(e64, e1)
    numpy.testing.assert_array_equal(h64, h1)
    eo, ho = e.copy(), h.copy()
    for _ in range(5):
        ref.maxwell_h(dt)(eo, ho, 1.5)
        ref.maxwell_e(dt)(eo, ho, 2.0)
    numpy.testing.assert_allclose(e64, eo, rtol=1e-11, atol=1e-12)
    numpy.testing.assert_allclose(h64, ho, rtol=1e-11, atol=1e-12)
    e32, h32 = e.astype(numpy.float32), h.astype(numpy.float32)
    fdtd.run(e32, h32, 2.0, dt, None, steps=5, mu=1.5)
    assert e32.dtype == numpy.float32
    numpy.testing.assert_allclose(e32, eo, rtol=1e-4, atol=1e-4)
    numpy.testing.assert_allclose(h32, ho, rtol=1e-4, atol=1e-4)


# ---- CPML (meanas/fdtd/pml.py) against the oracle's restatement ---------------------
def _cpml_case(shape, faces, thickness, seed, fdt, with_mu):
    rng = numpy.random.default_rng(seed)
    e = rng.standard_normal((3, *shape)).astype(fdt)
    h = rng.standard_normal((3, *shape)).astype(fdt)
    eps = (1 + rng.random((3, *shape))).astype(fdt)
    mu = (1 + rng.random((3, *shape))).astype(fdt) if with_mu else None
    dxes = [[0.8 + 0.4 * rng.random(n) for n in shape] for _ in range(2)]
    return e, h, eps, mu, dxes


@pytest.mark.parametrize('fdt,pdt,tol', [(numpy.float64, numpy.float64, 1e-11), (numpy.float64, numpy.float32, 1e-5),
                                         (numpy.float32, numpy.float32, 1e-4), (numpy.float32, numpy.float64, 1e-4)])
@pytest.mark.parametrize('shape,faces,thickness', [
    ((14, 12, 9), fc.ALL_FACES, 4),
    ((10, 9, 7), fc.ALL_FACES, 4),                 # z layers overlap (2 * 4 > 7)
    ((16, 1, 11), [(0, 1), (2, -1)], 5),           # two faces, one flat axis
])
@pytest.mark.parametrize('with_mu', [False, True])
def test_cpml_updates_match_oracle(fdt, pdt, tol, shape, faces, thickness, with_mu):
    from meanas_amd import fdtd
    e, h, eps, mu, dxes = _cpml_case(shape, faces, thickness, 5, fdt, with_mu)
    dt = 0.3
    kw = dict(epsilon_eff=2.0, cfs_alpha=0.05)
    uE, uH = ref.updates_with_cpml(fc.cpml_layers(ref.cpml_params, dt, faces, thickness, **kw), dt, dxes, eps,
                                   dtype=pdt)
    gE, gH = fdtd.updates_with_cpml(fc.cpml_layers(fdtd.cpml_params, dt, faces, thickness, **kw), dt, dxes, eps,
                                    dtype=pdt)
    et, ht = torch.from_numpy(e.copy()).cuda(), torch.from_numpy(h.copy()).cuda()
    epst = torch.from_numpy(eps).cuda()
    mut = None if mu is None else torch.from_numpy(mu).cuda()
    for _ in range(6):
        if mu is None:
            uH(e, h)
            gH(et, ht)
        else:
            uH(e, h, mu)
            gH(et, ht, mut)
        uE(e, h, eps)
        gE(et, ht, epst)
    scale = max(numpy.abs(e).max(), numpy.abs(h).max())
    numpy.testing.assert_allclose(et.cpu().numpy(), e, rtol=tol, atol=tol * scale)
    numpy.testing.assert_allclose(ht.cpu().numpy(), h, rtol=tol, atol=tol * scale)


@pytest.mark.parametrize('axis,comp', [(0, 1), (1, 2), (2, 0)])
def test_cpml_absorbs_a_pulse_on_gpu(axis, comp):
    from meanas_amd import fdtd
    upd = lambda p, dt, dxes, eps: fdtd.updates_with_cpml(p, dt, dxes, eps)
    final, top = fc.pulse_energy(upd, axis, comp, make=fdtd.cpml_params,
                                 to_dev=lambda a: torch.from_numpy(a).cuda(), to_host=lambda t: t.cpu().numpy())
    assert final < 1e-8 and top < 1.05


def test_cpml_numpy_fields_and_defaults():
    """numpy fields are written back in place; mu defaults to (1, 1, 1); no layers and
    equal E/H widths give the plain updates."""
    from meanas_amd import fdtd
    e, h, eps, _, dxes = _cpml_case((6, 5, 4), [], 3, 7, numpy.float64, False)
    dxes = [dxes[1], dxes[1]]
    gE, gH = fdtd.updates_with_cpml([[None, None]] * 3, 0.2, dxes, eps, dtype=numpy.float64)
    e1, h1 = e.copy(), h.copy()
    assert gH(e1, h1) is None and gE(e1, h1, eps) is None
    e2, h2 = e.copy(), h.copy()
    ref.maxwell_h(0.2, dxes)(e2, h2)
    ref.maxwell_e(0.2, dxes)(e2, h2, eps)
    numpy.testing.assert_allclose(e1, e2, rtol=1e-12, atol=1e-12)
    numpy.testing.assert_allclose(h1, h2, rtol=1e-12, atol=1e-12)


def test_cpml_rejects_what_has_no_kernel_form():
    import ctypes
    from meanas_amd import _lib, fdtd
    shape = (8, 8, 8)
    eps = numpy.ones((3, *shape))
    lay = fdtd.cpml_params(0, -1, 0.3, thickness=3)
    bad = dict(lay, region=(slice(2, 5), slice(None), slice(None)))
    with pytest.raises(ValueError):
        fdtd.updates_with_cpml([[bad, None], [None, None], [None, None]], 0.3, None, eps)
    with pytest.raises(ValueError):   # a polarity -1 layer in the +1 slot
        fdtd.updates_with_cpml([[None, lay], [None, None], [None, None]], 0.3, None, eps)
    # the C ABI refuses layers out of (axis, polarity) order
    gE, _ = fdtd.updates_with_cpml([[lay, fdtd.cpml_params(0, 1, 0.3, thickness=3)], [None, None], [None, None]],
                                   0.3, None, eps)
    tab = (_lib.MxCpmlLayer * 2)()
    ptrs = [a.data_ptr() for a in gE.psi_state[:4]]
    tab[0] = _lib.MxCpmlLayer(0, 1, 3, 0, *ptrs)
    tab[1] = _lib.MxCpmlLayer(0, -1, 3, 0, *ptrs)
    e = torch.zeros((3, *shape), dtype=torch.float64, device='cuda')
    grid = _lib.MxFdtdGrid(8, 8, 8, 0, (ctypes.c_void_p * 3)(), (ctypes.c_void_p * 3)())
    rc = _lib.lib().mx_fdtd_cpml_e(ctypes.byref(grid), _lib.MX_REAL64, _lib.MX_REAL32, 0.3, e.data_ptr(), e.data_ptr(),
                                   None, 1.0, tab, 2, None)
    assert rc != 0 and b'order' in _lib.lib().mx_last_error()


# ---- conducting boundary (meanas/fdtd/boundaries.py) ------------------------------
@pytest.mark.parametrize('direction', [0, 1, 2])
@pytest.mark.parametrize('polarity', [-1, 1])
@pytest.mark.parametrize('dtype', [numpy.float64, numpy.float32])
@pytest.mark.parametrize('size', ['minimal', 'regular'])
def test_conducting_boundary_matches_oracle(direction, polarity, dtype, size):
    """Both field updates of one face equal the oracle's bit for bit (copies, zeros and
    negations only); the smallest axis each polarity accepts is included."""
    from meanas_amd import fdtd
    shape = [5, 6, 7] if size == 'regular' else [3, 4, 1]
    if size == 'minimal':
        shape[direction] = 2 if polarity < 0 else 3
    rng = numpy.random.default_rng(10 + direction + 3 * polarity)
    for kind in (0, 1):
        f = rng.standard_normal((3, *shape)).astype(dtype)
        want = ref.conducting_boundary(direction, polarity)[kind](f.copy())
        t = torch.from_numpy(f.copy()).cuda()
        got = fdtd.conducting_boundary(direction, polarity)[kind](t)
        assert got is t
        numpy.testing.assert_array_equal(t.cpu().numpy(), want)


@pytest.mark.parametrize('direction', [0, 1, 2])
@pytest.mark.parametrize('dtype,tol', [(numpy.float64, 1e-13), (numpy.float32, 1e-5)])
def test_conducting_boundary_is_a_mirror_on_gpu(direction, dtype, tol):
    """GPU updates with the GPU boundaries on both faces equal a plain periodic GPU run on
    the doubled mirror-symmetric domain."""
    from meanas_amd import fdtd
    mh, me = fdtd.maxwell_h(0.3), fdtd.maxwell_e(0.3)

    def step(e, h, eps, after_h, after_e):
        mh(e, h)
        after_h()
        me(e, h, eps)
        after_e()
    worst = fc.mirror_mismatch(direction, 7, step, fdtd.conducting_boundary, dtype=dtype, seed=direction,
                               to_dev=lambda a: torch.from_numpy(a).cuda(), to_host=lambda t: t.cpu().numpy())
    assert worst <= tol


def test_conducting_boundary_numpy_and_errors():
    from meanas_amd import fdtd
    en, hn = fdtd.conducting_boundary(0, -5)        # any negative polarity is the low face
    f = numpy.random.default_rng(0).standard_normal((3, 4, 3, 2))
    want = ref.conducting_boundary(0, -1)[1](f.copy())
    assert hn(f) is f                                # numpy written back in place
    numpy.testing.assert_array_equal(f, want)
    ep, _ = fdtd.conducting_boundary(2, 1)
    with pytest.raises(IndexError):                  # +1 needs three cells along the axis
        ep(torch.zeros((3, 4, 4, 2), dtype=torch.float64, device='cuda'))
    with pytest.raises(TypeError):
        en(torch.zeros((3, 4, 4, 2), dtype=torch.complex128, device='cuda'))
    with pytest.raises(Exception, match='Invalid direction'):
        fdtd.conducting_boundary(3, 1)
    with pytest.raises(Exception, match='Bad polarity'):
        fdtd.conducting_boundary(1, 0)


# ---- energy and Poynting functions (meanas/fdtd/energy.py) -------------------------
def _gpu_energy():
    from types import SimpleNamespace
    from meanas_amd import fdtd
    return SimpleNamespace(energy_hstep=fdtd.energy_hstep, energy_estep=fdtd.energy_estep,
                           delta_energy_j=fdtd.delta_energy_j, poynting_divergence=fdtd.poynting_divergence,
                           poynting=fdtd.poynting)


@pytest.mark.parametrize('case', fc.cases())
def test_reference_fdtd_checks_with_gpu_energy(case):
    """The reference's FDTD checks (initial energy, energy conservation, Poynting
    divergence and planes) on GPU-stepped fields, with the GPU's energy and Poynting
    functions as the checker."""
    sim = fc.make_sim(case, gpu_step)
    sim.en = _gpu_energy()
    for check in fc.CHECKS:
        check(sim)


@pytest.mark.parametrize('dtype,tol', [(numpy.float64, 1e-14), (numpy.float32, 1e-6)])
@pytest.mark.parametrize('shape', [(5, 6, 7), (1, 4, 9), (8, 1, 3)])
@pytest.mark.parametrize('widths', [True, False])
@pytest.mark.parametrize('mats', ['field', 'scalar', None])
def test_energy_functions_match_oracle(dtype, tol, shape, widths, mats):
    """Every function of energy.py against the oracle's, which follows the reference's
    expression order (products and component sums in the field precision, widths in
    float64, one axis at a time)."""
    from meanas_amd import fdtd
    rng = numpy.random.default_rng(len(shape) * 7 + shape[0])
    f = [rng.standard_normal((3, *shape)).astype(dtype) for _ in range(4)]
    dxes = [[0.5 + rng.random(n) for n in shape] for _ in range(2)] if widths else None
    eps = {'field': (1 + rng.random((3, *shape))).astype(dtype), 'scalar': 2.5, None: None}[mats]
    mu = {'field': (1 + rng.random((3, *shape))).astype(dtype), 'scalar': 1.5, None: None}[mats]
    dt = 0.37
    kw = dict(epsilon=eps, mu=mu, dxes=dxes)
    dxes_given = dxes if dxes is not None else [[numpy.ones(n) for n in shape]] * 2
    pairs = {
        'hstep': (fdtd.energy_hstep(f[0], f[1], f[2], **kw), ref.energy_hstep(f[0], f[1], f[2], **kw)),
        'estep': (fdtd.energy_estep(f[0], f[1], f[2], **kw), ref.energy_estep(f[0], f[1], f[2], **kw)),
        'h2e': (fdtd.delta_energy_h2e(dt, *f, **kw), ref.delta_energy_h2e(dt, *f, **kw)),
        'e2h': (fdtd.delta_energy_e2h(dt, *f, **kw), ref.delta_energy_e2h(dt, *f, **kw)),
        'j': (fdtd.delta_energy_j(f[0], f[1], dxes=dxes), ref.delta_energy_j(f[0], f[1], dxes=dxes)),
        'dxmul': (fdtd.dxmul(f[0], f[1], **kw), ref.dxmul(f[0], f[1], **kw)),
        'poynting': (fdtd.poynting(f[0], f[1], dxes), ref.fdtd_poynting(f[0], f[1], dxes)),
        'div_s': (fdtd.poynting_divergence(f[2]), ref.poynting_divergence(f[2])),
        'div_eh': (fdtd.poynting_divergence(e=f[0], h=f[1], dxes=dxes_given),
                   ref.poynting_divergence(e=f[0], h=f[1], dxes=dxes_given)),
    }
    for name, (got, want) in pairs.items():
        assert isinstance(got, numpy.ndarray) and got.dtype == want.dtype and got.shape == want.shape, name
        scale = numpy.abs(want).max()
        numpy.testing.assert_allclose(got, want, rtol=tol, atol=tol * scale, err_msg=name)


def test_energy_functions_on_device_tensors():
    """CUDA tensors in, CUDA tensors out, equal to the numpy path; mixed dtypes refused."""
    from meanas_amd import fdtd
    rng = numpy.random.default_rng(3)
    e, h, e2 = (rng.standard_normal((3, 6, 5, 4)) for _ in range(3))
    dev = [torch.from_numpy(a).cuda() for a in (e, h, e2)]
    u = fdtd.energy_hstep(*dev, epsilon=2.0)
    assert isinstance(u, torch.Tensor) and u.is_cuda and u.dtype == torch.float64
    numpy.testing.assert_array_equal(u.cpu().numpy(), fdtd.energy_hstep(e, h, e2, epsilon=2.0))
    s = fdtd.poynting(dev[0], dev[1])
    assert isinstance(s, torch.Tensor) and s.shape == (3, 6, 5, 4)
    numpy.testing.assert_array_equal(s.cpu().numpy(), fdtd.poynting(e, h))
    dxes = [[0.5 + rng.random(n) for n in (6, 5, 4)] for _ in range(2)]
    first = fdtd.energy_estep(e, h, e2, dxes=dxes)
    numpy.testing.assert_allclose(first, ref.energy_estep(e, h, e2, dxes=dxes), rtol=1e-14)
    dxes[0][1] *= 1.5                        # changed in place: the device widths follow
    numpy.testing.assert_allclose(fdtd.energy_estep(e, h, e2, dxes=dxes), ref.energy_estep(e, h, e2, dxes=dxes),
                                  rtol=1e-14)
    with pytest.raises(ValueError):
        fdtd.energy_estep(e, h.astype(numpy.float32), e2)
    with pytest.raises(AssertionError):     # as the reference: e, h and dxes, or s
        fdtd.poynting_divergence(e=e, h=h)
