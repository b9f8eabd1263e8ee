"""
Pins the FP64 CPU oracle (oracle/fdfd_ref.py) against what mathematics fixes:
textbook routines (numpy.roll, dense linear algebra), closed forms (plane-wave
eigenvalues, the polynomial PML stretch, the 1D discrete Green's function) and
invariants (div curl = 0, complex symmetry, energy balance).  Each check is
chosen so a plausible mistake -- a dropped term, a wrong sign or index, a
transposed operand -- fails it.  The restated reference tests
(meanas/test/test_fdfd.py:14-55) run over the reference fixture grid
(meanas/test/conftest.py:18-94).
"""
import numpy
import pytest
from scipy import sparse

from oracle import fdfd_ref as ref


def rng(seed=0):
    return numpy.random.default_rng(seed)


def random_dxes(shape, seed, complex_widths=True):
    g = rng(seed)
    dxes = []
    for _ in range(2):
        row = []
        for n in shape:
            d = g.uniform(0.7, 1.4, n)
            if complex_widths:
                d = d + 1j * g.uniform(-0.3, 0.3, n)
            row.append(d)
        dxes.append(row)
    return dxes


# ---------------------------------------------------------------------------
# discrete calculus
# ---------------------------------------------------------------------------
@pytest.mark.parametrize('shape', [(4, 5, 6), (5, 1, 3), (1, 1, 7)])
@pytest.mark.parametrize('dist', [1, 2, -1, -3])
def test_shift_circ_is_numpy_roll(shape, dist):
    f = rng(1).standard_normal(shape)
    for axis in range(3):
        S = ref.shift_circ(axis, shape, dist)
        got = S @ f.ravel()
        # positive distance: row r picks coord+|d|; negative: the inverse permutation
        want = numpy.roll(f, -dist, axis=axis).ravel()
        numpy.testing.assert_array_equal(got, want)


def test_vec_unvec_roundtrip():
    f = rng(2).standard_normal((3, 4, 5, 6))
    v = ref.vec(f)
    assert v.shape == (3 * 4 * 5 * 6,)
    assert v[((1 * 4 + 2) * 5 + 3) * 6 + 4] == f[1, 2, 3, 4]      # z fastest, component slowest
    numpy.testing.assert_array_equal(ref.unvec(v, (4, 5, 6)), f)


@pytest.mark.parametrize('forward', [True, False])
def test_derivatives_on_complex_exponential(forward):
    shape = (6, 5, 8)
    dx = random_dxes(shape, 3)[0]
    ops = ref.deriv_forward(dx) if forward else ref.deriv_back(dx)
    idx = numpy.meshgrid(*[numpy.arange(n) for n in shape], indexing='ij')
    for a, n in enumerate(shape):
        theta = 2 * numpy.pi * 2 / n
        f = numpy.exp(1j * theta * idx[a])
        width = numpy.meshgrid(*dx, indexing='ij')[a]
        if forward:
            want = f * (numpy.exp(1j * theta) - 1) / width
        else:
            want = f * (1 - numpy.exp(-1j * theta)) / width
        numpy.testing.assert_allclose(ops[a] @ f.ravel(), want.ravel(), rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize('shape', [(4, 5, 6), (5, 1, 4)])
def test_div_curl_and_curl_grad_vanish(shape):
    dxe, dxh = random_dxes(shape, 4)
    n = int(numpy.prod(shape))
    for D, C in ((ref.deriv_forward(dxe), ref.curl_forward(dxe)),
                 (ref.deriv_back(dxh), ref.curl_back(dxh))):
        div = sparse.hstack(D)
        grad = sparse.vstack(D)
        assert abs(div @ C).max() < 1e-12
        assert abs(C @ grad).max() < 1e-12
        assert C.shape == (3 * n, 3 * n)


def _plane_wave(shape, widths, modes):
    """Uniform-grid plane wave exp(i k.r) and the per-axis K_a, phase factors."""
    idx = numpy.meshgrid(*[numpy.arange(n) for n in shape], indexing='ij')
    phase = sum(2 * numpy.pi * m * idx[a] / shape[a] for a, m in enumerate(modes))
    theta = numpy.array([numpy.pi * m / shape[a] for a, m in enumerate(modes)])    # k_a d_a / 2
    K = 2 * numpy.sin(theta) / numpy.asarray(widths)
    kf = K * numpy.exp(1j * theta)      # forward difference symbol / i
    kb = K * numpy.exp(-1j * theta)     # backward difference symbol / i
    return numpy.exp(1j * phase), K, kf, kb


def test_curl_curl_plane_wave_eigenvalues():
    shape = (6, 8, 10)
    widths = (0.9, 1.3, 1.1)
    dxes = [[numpy.full(n, w) for n, w in zip(shape, widths)] for _ in range(2)]
    wave, K, kf, kb = _plane_wave(shape, widths, (1, 2, 3))
    CC = ref.curl_back(dxes[1]) @ ref.curl_forward(dxes[0])
    # transverse polarisation: kb . e0 = 0
    e0 = numpy.cross(kb, [0.3, -1.0, 0.7])
    E = numpy.concatenate([c * wave.ravel() for c in e0])
    numpy.testing.assert_allclose(CC @ E, (K ** 2).sum() * E, rtol=1e-12, atol=1e-12)
    # longitudinal polarisation e0 = kf: curl of a discrete gradient, eigenvalue 0
    E = numpy.concatenate([c * wave.ravel() for c in kf])
    assert abs(CC @ E).max() < 1e-12


def test_e_full_plane_wave_eps_mu():
    shape = (6, 8, 10)
    widths = (1.0, 1.2, 0.8)
    omega = 0.37 + 0.01j
    dxes = [[numpy.full(n, w) for n, w in zip(shape, widths)] for _ in range(2)]
    wave, K, kf, kb = _plane_wave(shape, widths, (2, 1, 3))
    e0 = numpy.cross(kb, [1.0, 0.4, -0.2])
    E = numpy.concatenate([c * wave.ravel() for c in e0])
    n3 = E.size
    eps, mu = 2.5 - 0.1j, 1.7
    A = ref.e_full(omega, dxes, numpy.full(n3, eps), mu=numpy.full(n3, mu))
    numpy.testing.assert_allclose(A @ E, ((K ** 2).sum() / mu - omega ** 2 * eps) * E,
                                  rtol=1e-12, atol=1e-12)


def test_e_full_masks_and_mu_divides():
    shape = (4, 5, 3)
    dxes = random_dxes(shape, 5)
    g = rng(6)
    n3 = 3 * int(numpy.prod(shape))
    eps = g.uniform(1, 3, n3) + 1j * g.uniform(0, 0.2, n3)
    omega = 0.6
    A = ref.e_full(omega, dxes, eps)
    # mu divides the curl-curl term (fdfd/operators.py:93)
    A2 = ref.e_full(omega, dxes, eps, mu=numpy.full(n3, 2.0))
    W = sparse.diags_array(omega ** 2 * eps)
    assert abs((A2 + W) - 0.5 * (A + W)).max() < 1e-12
    # PMC everywhere removes the curl-curl term; PMC nowhere is the identity mask
    Ap = ref.e_full(omega, dxes, eps, pmc=numpy.ones(n3))
    assert abs(Ap + W).max() < 1e-14
    assert abs(ref.e_full(omega, dxes, eps, pmc=numpy.zeros(n3)) - A).max() == 0
    # PEC zeroes the masked rows and columns and leaves the rest untouched
    pec = (g.uniform(size=n3) < 0.3).astype(float)
    Ae = ref.e_full(omega, dxes, eps, pec=pec).toarray()
    m = pec != 0
    assert abs(Ae[m]).max() == 0 and abs(Ae[:, m]).max() == 0
    numpy.testing.assert_array_equal(Ae[numpy.ix_(~m, ~m)], A.toarray()[numpy.ix_(~m, ~m)])


@pytest.mark.parametrize('shape', [(4, 5, 3), (5, 1, 4), (1, 6, 1)])
def test_scaled_operator_is_complex_symmetric(shape):
    dxes = random_dxes(shape, 7)
    g = rng(8)
    n3 = 3 * int(numpy.prod(shape))
    eps = g.uniform(1, 3, n3) + 1j * g.uniform(-0.2, 0.2, n3)
    mu = g.uniform(0.8, 1.5, n3) + 1j * g.uniform(-0.1, 0.1, n3)
    pec = (g.uniform(size=n3) < 0.1).astype(float)
    pmc = (g.uniform(size=n3) < 0.1).astype(float)
    A0 = ref.e_full(0.9 - 0.05j, dxes, eps, mu=mu, pec=pec, pmc=pmc)
    Pl, Pr = ref.e_full_preconditioners(dxes)
    A = (Pl @ A0 @ Pr).toarray()
    scale = abs(A).max()
    assert abs(A - A.T).max() < 1e-12 * scale
    # the unscaled operator is not symmetric on a stretched grid: the scaling matters
    assert abs(A0.toarray() - A0.toarray().T).max() > 1e-3 * scale
    numpy.testing.assert_allclose((Pl @ Pr).diagonal(), 1.0, rtol=1e-15)


# ---------------------------------------------------------------------------
# SC-PML
# ---------------------------------------------------------------------------
@pytest.mark.parametrize('delta', [1.0, 1.5])
def test_stretch_with_scpml_closed_form(delta):
    n, th, omega, eps_eff, ln_R, m = 30, 7, 0.21, 2.25, -16, 4
    dxes = [[numpy.full(n, delta) for _ in range(3)] for _ in range(2)]
    s = ref.prepare_s_function(ln_R=ln_R, m=m)
    for axis in (0, 2):
        for pol in (-1, 1):
            dxes = ref.stretch_with_scpml(dxes, axis=axis, polarity=pol, omega=omega,
                                          epsilon_effective=eps_eff, thickness=th, s_function=s)
    i = numpy.arange(n, dtype=float)
    s_max = (m + 1) * ln_R / 2
    corr = numpy.sqrt(eps_eff) * omega

    def profile(u):     # u: distance into the PML in cells, 0 at the inner edge
        return 1 + 1j * s_max * (u / th) ** m / (th * delta) / corr
    e_want = numpy.ones(n, dtype=complex)
    h_want = numpy.ones(n, dtype=complex)
    front, back = i < th, i >= n - th
    e_want[front] = profile(th - i[front] - 0.5)
    h_want[front] = profile(th - i[front])
    e_want[back] = profile(i[back] + 0.5 - (n - th))
    h_want[back] = profile(i[back] - (n - th))
    for axis in (0, 2):
        numpy.testing.assert_allclose(dxes[0][axis], delta * e_want, rtol=1e-14)
        numpy.testing.assert_allclose(dxes[1][axis], delta * h_want, rtol=1e-14)
    numpy.testing.assert_array_equal(dxes[0][1], numpy.full(n, delta))


# ---------------------------------------------------------------------------
# the solve
# ---------------------------------------------------------------------------
def _small_problem(shape, seed, pec_frac=0.0, pmc_frac=0.0, with_mu=False, pml=2):
    g = rng(seed)
    omega = 2 * numpy.pi / 9.5
    dxes = [[numpy.ones(n) for n in shape] for _ in range(2)]
    for a, n in enumerate(shape):
        if n > 2 * pml + 1:
            for p in (-1, 1):
                ref.stretch_with_scpml(dxes, axis=a, polarity=p, omega=omega, thickness=pml)
    n3 = 3 * int(numpy.prod(shape))
    eps = g.uniform(1, 4, n3) + 1j * g.uniform(0, 0.05, n3)
    mu = g.uniform(1, 1.5, n3) if with_mu else None
    pec = (g.uniform(size=n3) < pec_frac).astype(float) if pec_frac else None
    pmc = (g.uniform(size=n3) < pmc_frac).astype(float) if pmc_frac else None
    J = g.standard_normal(n3) + 1j * g.standard_normal(n3)
    if pec is not None:
        J[pec != 0] = 0
    return omega, dxes, J, eps, mu, pec, pmc


@pytest.mark.parametrize('shape', [(5, 5, 5), (5, 5, 1), (5, 1, 5)])
@pytest.mark.parametrize('variant', ['plain', 'mu_pmc', 'pec'])
@pytest.mark.parametrize('adjoint', [False, True])
def test_generic_matches_dense_solve(shape, variant, adjoint):
    kw = dict(plain={}, mu_pmc=dict(with_mu=True, pmc_frac=0.1), pec=dict(pec_frac=0.1))[variant]
    omega, dxes, J, eps, mu, pec, pmc = _small_problem(shape, 11, **kw)
    E = ref.generic(omega, dxes, J, eps, mu, pec=pec, pmc=pmc, adjoint=adjoint,
                    matrix_solver_opts=dict(rtol=1e-13, atol=0, maxiter=20000))
    E_dense = ref.dense_solve(omega, dxes, J, eps, mu, pec=pec, pmc=pmc, adjoint=adjoint)
    err = numpy.linalg.norm(E - E_dense) / numpy.linalg.norm(E_dense)
    assert err < 1e-8, err


def test_adjoint_is_conjugate_of_forward_with_conjugated_data():
    omega, dxes, J, eps, mu, pec, pmc = _small_problem((5, 5, 5), 12)
    Ea = ref.dense_solve(omega, dxes, J, eps, adjoint=True)
    p = ref.e_full_preconditioner_vector(dxes)
    # E_adj = conj(Pl w) with (Pl A0 Pr) w = Pr conj(b0)
    A = sparse.diags_array(p) @ ref.e_full(omega, dxes, eps) @ sparse.diags_array(1 / p)
    w = numpy.linalg.solve(A.toarray(), numpy.conj(-1j * omega * J) / p)
    numpy.testing.assert_allclose(Ea, numpy.conj(p * w), rtol=1e-9, atol=1e-12)


def test_e_guess_writes_scaled_x0_into_opts():
    omega, dxes, J, eps, *_ = _small_problem((5, 5, 1), 13)
    opts = dict(rtol=1e-12)
    guess = numpy.ones_like(J)
    ref.generic(omega, dxes, J, eps, matrix_solver_opts=opts, E_guess=guess)
    numpy.testing.assert_allclose(opts['x0'], ref.e_full_preconditioner_vector(dxes) * guess)


def test_zero_source_returns_zeros():
    omega, dxes, J, eps, *_ = _small_problem((5, 5, 1), 14)
    E = ref.generic(omega, dxes, numpy.zeros_like(J), eps)
    assert not E.any()


def test_1d_greens_function_through_pml():
    """Ex on a z line: (-E[k+1]+2E[k]-E[k-1])/dx^2 - w^2 eps E = b delta_{k,k0}; outside
    the source the field is C exp(-iK|k-k0|) with cos(K) = 1 - w^2 eps/2, and
    C = b / (2i sin K).  The SC-PML (negative s_max) absorbs exp(-iK|z|)."""
    n, th, k0 = 80, 10, 40
    omega, eps_val = 2 * numpy.pi / 38.75, 2.0
    dxes = [[numpy.ones(1), numpy.ones(1), numpy.ones(n)] for _ in range(2)]
    for p in (-1, 1):
        ref.stretch_with_scpml(dxes, axis=2, polarity=p, omega=omega,
                               epsilon_effective=eps_val, thickness=th)
    J = numpy.zeros(3 * n, dtype=complex)
    J[k0] = 1.0                                   # Ex component, z index k0
    eps = numpy.full(3 * n, eps_val)
    E = ref.generic(omega, dxes, J, eps, matrix_solver_opts=dict(rtol=1e-13, atol=0))
    Kd = numpy.arccos(1 - omega ** 2 * eps_val / 2)
    b = -1j * omega
    z = numpy.arange(n)
    want = b / (2j * numpy.sin(Kd)) * numpy.exp(-1j * Kd * abs(z - k0))
    inner = slice(th + 2, n - th - 2)
    rel = abs(E[:n][inner] - want[inner]).max() / abs(want).max()
    assert rel < 2e-3, rel
    assert abs(E[n:]).max() < 1e-12              # Ey, Ez stay zero


# ---------------------------------------------------------------------------
# restated reference tests over the reference fixture grid
# (meanas/test/test_fdfd.py:14-55, conftest.py:18-94)
# ---------------------------------------------------------------------------
def _fixture_cases():
    cases = []
    for shape in [(5, 5, 1), (5, 1, 5), (5, 5, 5)]:
        is3d = 1 not in shape
        for eps_kind in ['center', '000', 'random']:
            for bg in [1.0, 1.5]:
                for fg in [1.0, 2.5]:
                    if is3d and (eps_kind == '000' or bg != 1 or fg not in (1.0, 2.0)):
                        continue
                    for dx in [1.0, 1.5]:
                        for dxk in ['uniform', 'centerbig']:
                            cases.append((shape, eps_kind, bg, fg, dx, dxk))
    return cases


def _fixture_sim(shape, eps_kind, bg, fg, dx, dxk):
    prng = numpy.random.RandomState(12345)
    full = (3, *shape)
    eps = numpy.full(full, bg, dtype=float)
    if eps_kind == 'center':
        eps[:, shape[0] // 2, shape[1] // 2, shape[2] // 2] = fg
    elif eps_kind == '000':
        eps[:, 0, 0, 0] = fg
    else:
        eps[:] = prng.uniform(low=min(bg, fg), high=max(bg, fg), size=full)
    dxes = [[numpy.full(s, dx) for s in shape] for _ in range(2)]
    if dxk == 'centerbig':
        for eh in (0, 1):
            for ax in (0, 1, 2):
                dxes[eh][ax][dxes[eh][ax].size // 2] *= 1.1
    j = numpy.zeros(full, dtype=complex)
    center = numpy.zeros(full, dtype=bool)
    center[:, shape[0] // 2, shape[1] // 2, shape[2] // 2] = True
    j[numpy.roll(center, [1, 1, 1], axis=(1, 2, 3))] = 1 + 1j
    j[numpy.roll(center, [-1, -1, -1], axis=(1, 2, 3))] = 1 - 1j
    omega = 1 / 1500
    e = ref.generic(omega, dxes, ref.vec(j), ref.vec(eps),
                    matrix_solver_opts=dict(atol=1e-15, rtol=1e-11))
    return omega, dxes, eps, j, ref.unvec(e, shape)


@pytest.mark.parametrize('case', _fixture_cases())
def test_reference_residual_and_poynting(case):
    shape = case[0]
    omega, dxes, eps, j, e = _fixture_sim(*case)
    # test_residual (test_fdfd.py:14-18)
    A = ref.e_full(omega, dxes, ref.vec(eps))
    assert numpy.linalg.norm(A @ ref.vec(e) + 1j * omega * ref.vec(j)) < 1e-10
    # test_poynting_planes (test_fdfd.py:21-55)
    mask = (j != 0).any(axis=0)
    assert mask.sum() == 2
    pts = numpy.where(mask)
    mask[pts[0][0], pts[1][0], pts[2][0]] = False
    mx, my, mz = (numpy.roll(mask, -1, axis=a) for a in range(3))
    hv = ref.e2h(omega, dxes) @ ref.vec(e)
    exh = ref.poynting_e_cross(ref.vec(e), dxes) @ hv.conj()
    s = ref.unvec(exh.real / 2, shape)
    planes = [s[0, mask].sum(), -s[0, mx].sum(), s[1, mask].sum(), -s[1, my].sum(),
              s[2, mask].sum(), -s[2, mz].sum()]
    dv = dxes[0][0][:, None, None] * dxes[0][1][None, :, None] * dxes[0][2][None, None, :]
    src = -((e * j.conj()).real * dv)[:, mask] / 2
    numpy.testing.assert_allclose(sum(planes), src.sum(), rtol=1e-6)


def test_poynting_sparse_equals_functional_form():
    """The oracle's sparse poynting_e_cross (operators.py:310-336) applied to h equals the
    roll-based E x H of fdfd/functional.py:187-221, restated here directly; and e2h
    equals curl_forward / (-i omega) assembled from scratch with numpy.roll."""
    rng = numpy.random.default_rng(11)
    shape = (5, 6, 7)
    dxes = [[0.5 + rng.random(n) + 0.1j * rng.random(n) for n in shape] for _ in range(2)]
    e = rng.standard_normal((3, *shape)) + 1j * rng.standard_normal((3, *shape))
    h = rng.standard_normal((3, *shape)) + 1j * rng.standard_normal((3, *shape))
    we = [dxes[0][0][:, None, None], dxes[0][1][None, :, None], dxes[0][2][None, None, :]]
    wh = [dxes[1][0][:, None, None], dxes[1][1][None, :, None], dxes[1][2][None, None, :]]
    ex, ey, ez = (e[c] * we[c] for c in range(3))
    hx, hy, hz = (h[c] * wh[c] for c in range(3))
    s = numpy.stack([numpy.roll(ey, -1, 0) * hz - numpy.roll(ez, -1, 0) * hy,
                     numpy.roll(ez, -1, 1) * hx - numpy.roll(ex, -1, 1) * hz,
                     numpy.roll(ex, -1, 2) * hy - numpy.roll(ey, -1, 2) * hx])
    got = ref.poynting_e_cross(ref.vec(e), dxes) @ ref.vec(h)
    numpy.testing.assert_allclose(got, ref.vec(s), rtol=1e-13, atol=1e-13)
    # e2h from forward differences (f[n+1] - f[n]) / dx_e[n]
    d = [lambda f, a=a: (numpy.roll(f, -1, a) - f) / we[a] for a in range(3)]
    curl = numpy.stack([d[1](e[2]) - d[2](e[1]), d[2](e[0]) - d[0](e[2]), d[0](e[1]) - d[1](e[0])])
    omega = 0.7 + 0.01j
    numpy.testing.assert_allclose(ref.e2h(omega, dxes) @ ref.vec(e), ref.vec(curl / (-1j * omega)),
                                  rtol=1e-12, atol=1e-12)


def test_shift_with_mirror_index_table():
    """fdmath/operators.py:55-99 on a 4-cell axis: +1 -> [1,2,3,3], -1 -> [0,0,1,2]."""
    f = numpy.arange(4.0)
    numpy.testing.assert_array_equal(ref.shift_with_mirror(0, (4, 1, 1), 1) @ f, [1, 2, 3, 3])
    numpy.testing.assert_array_equal(ref.shift_with_mirror(0, (4, 1, 1), -1) @ f, [0, 0, 1, 2])
    g = numpy.arange(12.0).reshape(3, 4)[None].reshape(-1)          # shape (1, 3, 4), z fastest
    numpy.testing.assert_array_equal((ref.shift_with_mirror(2, (1, 3, 4), 1) @ g).reshape(3, 4)[0], [1, 2, 3, 3])


def _pml_case(shape3, polarity, eps_fg, dx, solve):
    """meanas/test/test_fdfd_pml.py:14-28,73-184 with `solve` in place of generic."""
    shape = (3, *shape3)
    omega = 1 / 1500
    epsilon = numpy.full(shape, eps_fg)
    dxes = [[numpy.full(s, dx) for s in shape3] for _ in range(2)]
    dim = int(numpy.where(numpy.array(shape3) > 1)[0][0])
    for pol in (-1, 1):
        dxes = ref.stretch_with_scpml(dxes, axis=dim, polarity=pol, omega=omega, epsilon_effective=eps_fg,
                                      thickness=10)
    other = [0, 1, 2]
    other.remove(dim)
    mid = shape3[dim] // 2
    dx_prop = (dxes[0][dim][mid] + dxes[1][dim][mid]) / 2
    center = numpy.zeros(shape, dtype=bool)
    center[other, shape3[0] // 2, shape3[1] // 2, shape3[2] // 2] = True
    k = omega * numpy.sqrt(epsilon[center].mean())
    k_corr = 2 / dx_prop * numpy.arcsin(k * dx_prop / 2)
    e = numpy.zeros(shape, dtype=complex)
    e[center] = 1 / numpy.linalg.norm(center)
    slices = [slice(None)] * 3
    slices[dim] = slice(mid, mid + 1)
    return omega, dxes, epsilon, e, k_corr, dim, slices


def _pml_check(E, shape3, polarity):
    e = ref.unvec(E, shape3)
    e_sqr = numpy.squeeze((e.conj() * e).sum(axis=0))
    tgt, rev = e_sqr[16:19], e_sqr[10:13]
    if polarity < 0:
        tgt, rev = rev, tgt
    numpy.testing.assert_allclose(rev, 0, atol=1e-12)
    numpy.testing.assert_allclose(tgt, 1, rtol=3e-6)


@pytest.mark.parametrize('shape3', [(30, 1, 1), (1, 30, 1), (1, 1, 30)])
@pytest.mark.parametrize('polarity', [1, -1])
@pytest.mark.parametrize('eps_fg', [1.0, 2.5])
@pytest.mark.parametrize('dx', [1.0, 1.5])
def test_one_way_source_oracle(shape3, polarity, eps_fg, dx):
    """The reference's test_pml on the oracle: compute_source launches one way only."""
    omega, dxes, epsilon, e, k, dim, slices = _pml_case(shape3, polarity, eps_fg, dx, None)
    J = ref.compute_source(e, k, omega, dxes, dim, polarity, slices, epsilon)
    E = ref.generic(omega, dxes, ref.vec(J), ref.vec(epsilon), matrix_solver_opts=dict(atol=1e-15, rtol=1e-11))
    _pml_check(E, shape3, polarity)


def _rev(v, shape):
    """R: reverse every cell index of a 3N field (components kept)."""
    return numpy.asarray(v).reshape(3, *shape)[:, ::-1, ::-1, ::-1].reshape(-1)


def test_h_full_is_reversed_e_full():
    """h_full(omega, dxes, eps, mu, pec, pmc) = R e_full(omega, [rev dxh, rev dxe], rev mu,
    rev eps, pec=rev pmc, pmc=rev pec) R (R S R = S^-1 for the periodic shift)."""
    rng = numpy.random.default_rng(21)
    shape = (5, 4, 6)
    n3 = 3 * 120
    dxes = [[0.7 + rng.random(k) + 0.2j * rng.random(k) for k in shape] for _ in range(2)]
    eps = 1 + rng.random(n3) + 0.1j * rng.random(n3)
    mu = 1 + rng.random(n3)
    pec = rng.random(n3) < 0.1
    pmc = rng.random(n3) < 0.1
    omega = 0.37 - 0.01j
    H = ref.h_full(omega, dxes, eps, mu, pec, pmc)
    dual = [[d[::-1] for d in dxes[1]], [d[::-1] for d in dxes[0]]]
    E = ref.e_full(omega, dual, _rev(mu, shape), _rev(eps, shape), pec=_rev(pmc, shape), pmc=_rev(pec, shape))
    v = rng.standard_normal(n3) + 1j * rng.standard_normal(n3)
    numpy.testing.assert_allclose(H @ v, _rev(E @ _rev(v, shape), shape), rtol=1e-12, atol=1e-12)
    # mu = None is mu = 1 (then the dual problem has eps = 1)
    numpy.testing.assert_allclose(ref.h_full(omega, dxes, eps) @ v,
                                  _rev(ref.e_full(omega, dual, numpy.ones(n3), _rev(eps, shape)) @ _rev(v, shape),
                                       shape), rtol=1e-12, atol=1e-12)


def test_eh_full_reduces_to_e_full():
    """With H = e2h(E) the second block row of eh_full vanishes and -i omega times the
    first equals e_full @ E (no M, no masks)."""
    rng = numpy.random.default_rng(22)
    shape = (4, 5, 3)
    n3 = 3 * 60
    dxes = [[0.8 + rng.random(k) for k in shape] for _ in range(2)]
    eps = 1 + rng.random(n3)
    mu = 1 + rng.random(n3)
    omega = 0.45
    e = rng.standard_normal(n3) + 1j * rng.standard_normal(n3)
    h = ref.e2h(omega, dxes, mu) @ e
    out = ref.eh_full(omega, dxes, eps, mu) @ numpy.concatenate([e, h])
    numpy.testing.assert_allclose(out[n3:], 0, atol=1e-12)
    numpy.testing.assert_allclose(-1j * omega * out[:n3], ref.e_full(omega, dxes, eps, mu) @ e, rtol=1e-12)


def test_m2j_plane_wave_closed_form():
    """On uniform widths d, M = a e^{i k.r} gives J = (kb x a / mu) e^{i k.r} / (-i omega)
    with the backward-difference factor kb_c = (1 - e^{-i k_c d_c}) / d_c; the sparse
    form (fdfd/operators.py:285-307) is minus the functional one."""
    shape = (6, 5, 4)
    d = (0.7, 1.3, 0.9)
    n = (1, 2, 3)                                            # wave numbers in cells
    k = [2 * numpy.pi * n[a] / (shape[a] * d[a]) for a in range(3)]
    dxes = [[numpy.full(shape[a], d[a]) for a in range(3)] for _ in range(2)]
    grids = numpy.meshgrid(*[numpy.arange(shape[a]) * d[a] for a in range(3)], indexing='ij')
    phase = numpy.exp(1j * sum(k[a] * grids[a] for a in range(3)))
    amp = numpy.array([0.3 - 0.2j, 1.1, -0.4 + 0.9j])
    mu0 = 1.7 - 0.1j
    omega = 0.83 + 0.02j
    m = amp[:, None, None, None] * phase
    kb = numpy.array([(1 - numpy.exp(-1j * k[a] * d[a])) / d[a] for a in range(3)])
    want = numpy.cross(kb, amp / mu0)[:, None, None, None] * phase / (-1j * omega)
    got = ref.m2j_functional(omega, dxes, numpy.full((3, *shape), mu0))(m)
    numpy.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)
    sp = ref.m2j(omega, dxes, numpy.full(3 * numpy.prod(shape), mu0)) @ ref.vec(m)
    numpy.testing.assert_allclose(sp, -ref.vec(got), rtol=1e-12, atol=1e-12)
    numpy.testing.assert_allclose(ref.m2j_functional(omega, dxes)(m), got * mu0, rtol=1e-12, atol=1e-12)


def test_poynting_h_cross_is_minus_e_cross():
    """poynting_h_cross(h) @ e = -(poynting_e_cross(e) @ h): both hold the same width-weighted
    products, E shifted forward, with the roles of the two fields exchanged."""
    rng = numpy.random.default_rng(21)
    shape = (4, 3, 5)
    n3 = 3 * int(numpy.prod(shape))
    dxes = [[0.5 + rng.random(n) + 0.1j * rng.random(n) for n in shape] for _ in range(2)]
    e = rng.standard_normal(n3) + 1j * rng.standard_normal(n3)
    h = rng.standard_normal(n3) + 1j * rng.standard_normal(n3)
    numpy.testing.assert_allclose(ref.poynting_h_cross(h, dxes) @ e, -(ref.poynting_e_cross(e, dxes) @ h),
                                  rtol=1e-13, atol=1e-13)
