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


@pytest.mark.parametrize('shape', [(9, 10, 70), (1, 12, 5), (7, 1, 1)])
def test_run_matches_steps_and_oracle(shape):
    from meanas_amd import fdtd
    e, h, eps, mu, dxes = _fields(shape, 1, numpy.float64)
    dt = 0.2
    et = torch.from_numpy(e.copy()).cuda()
    ht = torch.from_numpy(h.copy()).cuda()
    fdtd.run(et, ht, eps, dt, dxes, steps=6, mu=mu)
    e2, h2 = e.copy(), h.copy()
    me, mh = fdtd.maxwell_e(dt, dxes), fdtd.maxwell_h(dt, dxes)
    for _ in range(6):
        mh(e2, h2, mu)
        me(e2, h2, eps)
    numpy.testing.assert_array_equal(et.cpu().numpy(), e2)   # same kernels, same order
    numpy.testing.assert_array_equal(ht.cpu().numpy(), h2)
    e3, h3 = e.copy(), h.copy()
    for _ in range(6):
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
