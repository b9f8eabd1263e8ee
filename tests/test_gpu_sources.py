"""GPU source construction against the oracle, and the reference's one-way PML test
(meanas/test/test_fdfd_pml.py:14-28) with the source built and the field solved on the GPU."""
import numpy
import pytest

from oracle import fdfd_ref as ref
from test_oracle import _pml_case, _pml_check

pytestmark = pytest.mark.gpu


def rel(a, b):
    return numpy.linalg.norm(a - b) / numpy.linalg.norm(b)


def _problem(shape, seed, weighted=False):
    rng = numpy.random.default_rng(seed)
    dxes = [[1.0 + 0.3 * rng.random(n) for n in shape] for _ in range(2)]
    for ax in range(3):
        if shape[ax] >= 10:
            for pol in (-1, 1):
                dxes = ref.stretch_with_scpml(dxes, axis=ax, polarity=pol, omega=0.3, thickness=3)
    n3 = 3 * int(numpy.prod(shape))
    eps = (1.0 + 2 * rng.random(n3)) * (1 + 0.05j * rng.random(n3))
    mu = 1.0 + rng.random(n3)
    e = rng.standard_normal(n3) + 1j * rng.standard_normal(n3)
    mask = (rng.random(n3) < 0.5).astype(float)
    if weighted:   # a fractional region: Q = diag(TF_region) uses the values themselves
        mask = mask * rng.choice([0.25, 0.5, 1.0], size=n3)
    return rng, dxes, eps, mu, e, mask


@pytest.mark.parametrize('shape', [(10, 12, 7), (1, 16, 12), (12, 1, 1)])
@pytest.mark.parametrize('with_mu', [False, True])
@pytest.mark.parametrize('weighted', [False, True])
def test_tfsf_source_matches_oracle(shape, with_mu, weighted):
    from meanas_amd.fdfd import sources
    rng, dxes, eps, mu, e, mask = _problem(shape, 7, weighted)
    mu = mu if with_mu else None
    omega = 0.3 - 0.002j
    got = sources.e_tfsf_source(mask, omega, dxes, eps, mu)(e)
    want = ref.e_tfsf_source(mask, omega, dxes, eps, mu) @ e
    assert rel(got, want) < 1e-12


@pytest.mark.parametrize('shape', [(10, 12, 7), (12, 1, 1)])
@pytest.mark.parametrize('periodic', [False, True])
@pytest.mark.parametrize('weighted', [False, True])
def test_boundary_source_matches_oracle(shape, periodic, weighted):
    from meanas_amd.fdfd import sources
    rng, dxes, eps, mu, e, mask = _problem(shape, 8, weighted)
    got = sources.e_boundary_source(mask, 0.25, dxes, eps, periodic_mask_edges=periodic)(e)
    want = ref.e_boundary_source(mask, 0.25, dxes, eps, periodic_mask_edges=periodic) @ e
    assert rel(got, want) < 1e-12


@pytest.mark.parametrize('shape3', [(30, 1, 1), (1, 30, 1), (1, 1, 30)])
@pytest.mark.parametrize('polarity', [1, -1])
@pytest.mark.parametrize('eps_fg', [1.0, 2.5])
@pytest.mark.parametrize('dx', [1.0, 1.5])
def test_one_way_source_gpu(shape3, polarity, eps_fg, dx):
    """test_pml with compute_source and generic both on the GPU; the source also matches
    the oracle's."""
    from meanas_amd.fdfd import solvers, sources
    omega, dxes, epsilon, e, k, dim, slices = _pml_case(shape3, polarity, eps_fg, dx, None)
    J = sources.compute_source(e, k, omega, dxes, dim, polarity, slices, epsilon)
    J_ref = ref.compute_source(e, k, omega, dxes, dim, polarity, slices, epsilon)
    assert rel(J, J_ref) < 1e-12
    E = solvers.generic(omega, dxes, ref.vec(J), ref.vec(epsilon), matrix_solver_opts=dict(atol=1e-15, rtol=1e-11))
    _pml_check(E, shape3, polarity)
