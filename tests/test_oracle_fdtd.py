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
