"""GPU parity of the matrix-free operator apply against the oracle's sparse matrix.

The HIP kernel applies A = Pl e_full Pr (reference fdfd/solvers.py:124,
fdfd/operators.py:40-123) with p factored per axis; the oracle builds the same
matrix as CSR from the reference's sparse building blocks.  The factored p can differ
from the reference's sqrt(p^2) by a per-cell sign s; the kernel must equal
S A_ref S with S = diag(s) (a +-1 similarity that leaves E unchanged).
Tolerances (norm-wise relative): complex128 1e-13, complex64 2e-6.
"""
import numpy
import pytest
import torch

from oracle import fdfd_ref as ref

pytestmark = pytest.mark.gpu

SHAPES = [(5, 5, 5), (5, 5, 1), (5, 1, 5), (1, 5, 5), (1, 1, 7), (7, 9, 70), (13, 17, 130), (24, 8, 64)]


def make_case(shape, seed, pml=True, eps_kind='aniso_complex', mu=False, masks=False):
    g = numpy.random.default_rng(seed)
    omega = 2 * numpy.pi / 9.5 + 0.01j
    dxes = [[g.uniform(0.8, 1.3, n) for n in shape] for _ in range(2)]
    if pml:
        for a, n in enumerate(shape):
            if n >= 7:
                for p in (-1, 1):
                    ref.stretch_with_scpml(dxes, axis=a, polarity=p, omega=omega.real, thickness=3)
    n3 = 3 * int(numpy.prod(shape))
    if eps_kind == 'iso_real':
        e = g.uniform(1, 4, n3 // 3)
        eps = numpy.concatenate([e, e, e])
    elif eps_kind == 'aniso_real':
        eps = g.uniform(1, 4, n3)
    else:
        eps = g.uniform(1, 4, n3) + 1j * g.uniform(-0.2, 0.2, n3)
    mu_v = (g.uniform(0.8, 1.6, n3) + 1j * g.uniform(-0.1, 0.1, n3)) if mu else None
    pec = (g.uniform(size=n3) < 0.08).astype(float) if masks else None
    pmc = (g.uniform(size=n3) < 0.08).astype(float) if masks else None
    return omega, dxes, eps, mu_v, pec, pmc


def p_separable(dxes):
    se = [numpy.sqrt(numpy.asarray(d, dtype=complex)) for d in dxes[0]]
    sh = [numpy.sqrt(numpy.asarray(d, dtype=complex)) for d in dxes[1]]
    g = lambda a, b, c: (a[:, None, None] * b[None, :, None] * c[None, None, :]).ravel()
    return numpy.concatenate([g(se[0], sh[1], sh[2]), g(sh[0], se[1], sh[2]), g(sh[0], sh[1], se[2])])


def reference_apply(omega, dxes, eps, mu, pec, pmc, x):
    A0 = ref.e_full(omega, dxes, eps, mu=mu, pec=pec, pmc=pmc)
    p_ref = ref.e_full_preconditioner_vector(dxes)
    s = numpy.round(p_separable(dxes) / p_ref).real
    assert numpy.all(numpy.abs(s) == 1)
    return s * (p_ref * (A0 @ ((s * x) / p_ref)))


def rel(a, b):
    return numpy.linalg.norm(a - b) / numpy.linalg.norm(b)


# One variant per material kind of the kernel (real isotropic eps; real per-component eps;
# real isotropic eps with PEC/PMC masks; the general kernel twice)
VARIANTS = dict(plain_iso=dict(eps_kind='iso_real'), aniso_real=dict(eps_kind='aniso_real'),
                iso_masks=dict(eps_kind='iso_real', masks=True), aniso_complex=dict(eps_kind='aniso_complex'),
                mu_masks=dict(eps_kind='aniso_real', mu=True, masks=True))


@pytest.mark.parametrize('shape', SHAPES)
@pytest.mark.parametrize('variant', list(VARIANTS))
@pytest.mark.parametrize('mode', ['c64', 'c128'])
def test_apply_matches_oracle(shape, variant, mode):
    from meanas_amd.fdfd.device import DeviceOperator
    kw = VARIANTS[variant]
    omega, dxes, eps, mu, pec, pmc = make_case(shape, 1 + len(variant) + sum(shape), **kw)
    op = DeviceOperator(omega, dxes, eps, mu=mu, pec=pec, pmc=pmc)
    g = numpy.random.default_rng(7)
    x = g.standard_normal(op.n3) + 1j * g.standard_normal(op.n3)
    if mode == 'c64':
        x = x.astype(numpy.complex64)
        y = op.apply(torch.from_numpy(x).cuda()).cpu().numpy()
        tol = 2e-6
    else:
        y = op.apply(torch.from_numpy(x).cuda()).cpu().numpy()
        tol = 1e-13
    want = reference_apply(omega, dxes, eps, mu, pec, pmc, x.astype(numpy.complex128))
    assert rel(y.astype(numpy.complex128), want) < tol


@pytest.mark.parametrize('tile_rows', [2, 4, 8])
@pytest.mark.parametrize('slab', [1, 3, 16])
@pytest.mark.parametrize('variant', ['plain_iso', 'aniso_real', 'iso_masks', 'mu_masks'])
def test_apply_tiling_invariance(tile_rows, slab, variant):
    """Every tile height / slab depth gives the same complex128 result as the oracle, for
    every material kind."""
    from meanas_amd.fdfd.device import DeviceOperator
    shape = (19, 21, 75)
    kw = dict(VARIANTS[variant], eps_kind='aniso_complex') if variant == 'mu_masks' else VARIANTS[variant]
    omega, dxes, eps, mu, pec, pmc = make_case(shape, 3, **kw)
    op = DeviceOperator(omega, dxes, eps, mu=mu, pec=pec, pmc=pmc)
    g = numpy.random.default_rng(8)
    x = g.standard_normal(op.n3) + 1j * g.standard_normal(op.n3)
    plan = op.plan(mode='c128', tile_rows=tile_rows, slab=slab)
    y = plan.apply(torch.from_numpy(x).cuda()).cpu().numpy()
    want = reference_apply(omega, dxes, eps, mu, pec, pmc, x)
    assert rel(y, want) < 1e-13


def test_scaled_operator_symmetry_on_gpu():
    """x^T A y = y^T A x for the GPU operator (complex symmetry, fdfd/operators.py:105-109)."""
    from meanas_amd.fdfd.device import DeviceOperator
    omega, dxes, eps, mu, pec, pmc = make_case((9, 11, 70), 5, eps_kind='aniso_complex', mu=True, masks=True)
    op = DeviceOperator(omega, dxes, eps, mu=mu, pec=pec, pmc=pmc)
    g = numpy.random.default_rng(9)
    xs = [g.standard_normal(op.n3) + 1j * g.standard_normal(op.n3) for _ in range(2)]
    ax, ay = (op.apply(torch.from_numpy(v).cuda()).cpu().numpy() for v in xs)
    lhs, rhs = xs[0] @ ay, xs[1] @ ax
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_residual_entry_point():
    from meanas_amd.fdfd.device import DeviceOperator
    omega, dxes, eps, mu, pec, pmc = make_case((8, 10, 66), 6)
    op = DeviceOperator(omega, dxes, eps)
    g = numpy.random.default_rng(10)
    x = g.standard_normal(op.n3) + 1j * g.standard_normal(op.n3)
    b = g.standard_normal(op.n3) + 1j * g.standard_normal(op.n3)
    r, n2 = op.residual(torch.from_numpy(b).cuda(), torch.from_numpy(x).cuda())
    want = b - reference_apply(omega, dxes, eps, None, None, None, x)
    assert rel(r.cpu().numpy(), want) < 1e-13
    assert abs(n2 - numpy.vdot(want, want).real) <= 1e-12 * n2


@pytest.mark.gpu
def test_time_kernels_reports_every_kernel():
    """mx_time_kernels times each iteration kernel; the solve that follows still works."""
    from meanas_amd import _lib
    from meanas_amd.fdfd.device import DeviceOperator, time_kernels, solve_batch
    rng = numpy.random.default_rng(5)
    shape = (12, 10, 70)
    dxes = [[numpy.ones(n) for n in shape] for _ in range(2)]
    eps = 1.0 + rng.random(3 * int(numpy.prod(shape)))
    op = DeviceOperator(0.3 + 0.03j, dxes, eps)   # lossy: the solve below converges fast
    b = torch.from_numpy(rng.standard_normal(op.n3) + 1j * rng.standard_normal(op.n3)).cuda()
    for mode in (_lib.MX_FP64, _lib.MX_MIXED):
        kt = time_kernels(op, b, mode, reps=5, slab=4)
        for k in ('apply', 'fin_apply', 'update', 'fin_update', 'pupdate'):
            assert numpy.isfinite(kt[k]) and kt[k] > 0.0, kt
    x = torch.zeros_like(b)
    st = solve_batch([op], [dict(op=0, b=b, x=x, rtol=1e-8)], mode=_lib.MX_FP64)[0]
    assert st['status'] == 'converged', st
