"""GPU curls and wave operators (e_full, h_full, eh_full) against the oracle's sparse
forms (meanas/fdmath/operators.py:222-251, meanas/fdfd/operators.py:40-96,126-250)."""
import numpy
import pytest

from oracle import fdfd_ref as ref

pytestmark = pytest.mark.gpu


def rel(a, b):
    return numpy.linalg.norm(a - b) / numpy.linalg.norm(b)


def _problem(shape, seed, masks):
    rng = numpy.random.default_rng(seed)
    dxes = [[0.8 + 0.4 * rng.random(n) for n in shape] for _ in range(2)]
    for ax in range(3):
        if shape[ax] >= 10:
            for pol in (-1, 1):
                dxes = ref.stretch_with_scpml(dxes, axis=ax, polarity=pol, omega=0.4, thickness=3)
    n3 = 3 * int(numpy.prod(shape))
    eps = (1 + 2 * rng.random(n3)) * (1 + 0.1j * rng.random(n3))
    mu = (1 + rng.random(n3)) * (1 + 0.05j * rng.random(n3))
    pec = (rng.random(n3) < 0.08) if masks else None
    pmc = (rng.random(n3) < 0.08) if masks else None
    v = rng.standard_normal(n3) + 1j * rng.standard_normal(n3)
    w = rng.standard_normal(n3) + 1j * rng.standard_normal(n3)
    return dxes, eps, mu, pec, pmc, v, w


SHAPES = [(10, 9, 12), (1, 14, 10), (11, 1, 1), (6, 7, 70)]


@pytest.mark.parametrize('shape', SHAPES)
def test_curls_match_oracle(shape):
    from meanas_amd.fdfd import functional
    dxes, eps, mu, pec, pmc, v, w = _problem(shape, 1, False)
    assert rel(functional.curl_forward(dxes[0])(v), ref.curl_forward(dxes[0]) @ v) < 1e-13
    assert rel(functional.curl_back(dxes[1])(v), ref.curl_back(dxes[1]) @ v) < 1e-13
    # field layout in = layout out
    out = functional.curl_back(dxes[1])(v.reshape(3, *shape))
    assert out.shape == (3, *shape)


@pytest.mark.parametrize('shape', SHAPES)
@pytest.mark.parametrize('with_mu', [False, True])
@pytest.mark.parametrize('masks', [False, True])
def test_wave_operators_match_oracle(shape, with_mu, masks):
    from meanas_amd.fdfd import functional
    dxes, eps, mu, pec, pmc, v, w = _problem(shape, 2, masks)
    mu = mu if with_mu else None
    omega = 0.41 - 0.003j
    got = functional.e_full(omega, dxes, eps, mu, pec, pmc)(v)
    assert rel(got, ref.e_full(omega, dxes, eps, mu, pec=pec, pmc=pmc) @ v) < 1e-12
    got = functional.h_full(omega, dxes, eps, mu, pec, pmc)(v)
    assert rel(got, ref.h_full(omega, dxes, eps, mu, pec, pmc) @ v) < 1e-12
    j, m = functional.eh_full(omega, dxes, eps, mu, pec, pmc)(v, w)
    want = ref.eh_full(omega, dxes, eps, mu, pec, pmc) @ numpy.concatenate([v, w])
    n3 = v.size
    assert rel(j, want[:n3]) < 1e-12
    assert rel(m, want[n3:]) < 1e-12
