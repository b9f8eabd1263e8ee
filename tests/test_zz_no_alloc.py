"""mx_residual_noalloc (the allocation-free residual) against mx_residual through the C ABI.

Two calls in a row on grids of different size (the second smaller, so the fixed
workspace still holds the first call's partial sums and state): each gives exactly the
bits and the norm of mx_residual on the same inputs, and the FP64 norm of r.
"""
import ctypes

import numpy
import pytest

pytestmark = pytest.mark.gpu


def operator_and_vectors(shape, seed, pec=False):
    import torch
    from meanas_amd import workloads
    from meanas_amd.fdfd.device import DeviceOperator
    g = numpy.random.default_rng(seed)
    omega = 2 * numpy.pi / 11.0
    dxes = workloads.pml_dxes(shape, omega, thickness=3)
    n3 = 3 * int(numpy.prod(shape))
    eps = g.uniform(1, 4, n3) + 1j * g.uniform(0, 0.1, n3)
    mask = (g.uniform(size=n3) < 0.05).astype(float) if pec else None
    op = DeviceOperator(omega, dxes, eps, pec=mask)
    b = torch.from_numpy(g.standard_normal(n3) + 1j * g.standard_normal(n3)).cuda()
    x = torch.from_numpy(g.standard_normal(n3) + 1j * g.standard_normal(n3)).cuda()
    return op, b, x


def residual(entry, op, b, x):
    import torch
    from meanas_amd import _lib
    from meanas_amd.fdfd.device import torch_stream_handle
    r = torch.empty_like(b)
    n2 = ctypes.c_double(-1.0)
    rc = getattr(_lib.lib(), entry)(ctypes.byref(op.desc), b.data_ptr(), x.data_ptr(), r.data_ptr(),
                                    ctypes.byref(n2), torch_stream_handle())
    _lib.check(rc, entry)
    return r, n2.value


@pytest.mark.parametrize('pec', [False, True])
def test_noalloc_residual_equals_residual_twice(pec):
    import torch
    for shape, seed in (((26, 21, 70), 1), ((9, 12, 5), 2)):
        op, b, x = operator_and_vectors(shape, seed, pec)
        r_ref, n_ref = residual('mx_residual', op, b, x)
        r_na, n_na = residual('mx_residual_noalloc', op, b, x)
        assert torch.equal(r_na, r_ref) and n_na == n_ref, shape
        want = float(torch.sum(torch.abs(r_ref) ** 2))
        assert abs(n_na - want) <= 1e-12 * want
        assert 0.0 < n_na < float('inf')
