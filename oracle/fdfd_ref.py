"""
FP64 CPU restatement of meanas' 3D FDFD solve path -- TEST INFRASTRUCTURE ONLY
(used by `tests/`, `__graft_entry__.smoke()` and `bench.py`'s cpu_baseline leg
as the checker; never imported by the `meanas_amd` product package).

Restated from the reference's text (paths relative to the reference root):
  meanas/fdmath/vectorization.py:51-125   vec / unvec (C order, component-major)
  meanas/fdmath/operators.py:15-52        shift_circ
  meanas/fdmath/operators.py:99-150       deriv_forward / deriv_back
  meanas/fdmath/operators.py:153-171      cross
  meanas/fdmath/operators.py:222-251      curl_forward / curl_back
  meanas/fdfd/operators.py:40-96          e_full
  meanas/fdfd/operators.py:99-123         e_full_preconditioners
  meanas/fdfd/operators.py:253-282        e2h
  meanas/fdfd/operators.py:310-336        poynting_e_cross
  meanas/fdfd/operators.py:339-363        poynting_h_cross
  meanas/fdfd/scpml.py:18-38,98-167       prepare_s_function / stretch_with_scpml
  meanas/fdfd/solvers.py:21-141           _scipy_qmr / generic
  meanas/fdfd/operators.py:126-250        h_full / eh_full
  meanas/fdmath/operators.py:55-99        shift_with_mirror
  meanas/fdmath/functional.py             curl_forward / curl_back (callables, dx None = 1)
  meanas/fdfd/functional.py:120-151       m2j (and the sparse form, fdfd/operators.py:285-307)
  meanas/fdtd/base.py:13-125              maxwell_e / maxwell_h
  meanas/fdtd/energy.py:10-295            poynting, poynting_divergence, energy_hstep,
                                          energy_estep, delta_energy_h2e, delta_energy_e2h,
                                          delta_energy_j, dxmul (the same expression order)
  meanas/fdfd/operators.py:366-450        e_tfsf_source / e_boundary_source
  meanas/fdfd/waveguide_3d.py:113-156,206-252  compute_source / expand_e
  meanas/fdtd/pml.py:23-230               cpml_params / updates_with_cpml
  meanas/fdtd/boundaries.py:12-78         conducting_boundary (tuple indexing; see there)
Third-party arithmetic: scipy.sparse (CSR products) and scipy.sparse.linalg.qmr
(reference pins scipy~=1.14, pyproject.toml:44; this container has 1.15.3).

How each function is pinned (tests/test_oracle.py):
  shift_circ        -- equals the permutation numpy.roll performs (textbook routine),
                       negative distance is the inverse permutation.
  deriv_*           -- closed form on a periodic complex exponential
                       ((e^{ik dx}-1)/dx factor, backward (1-e^{-ik dx})/dx).
  cross / curl_*    -- discrete div(curl)=0 identity (backward divergence of the
                       forward curl vanishes); plane-wave eigenvalues of curl_b curl_f
                       (transverse: sum_a (2/dx_a sin(k_a dx_a/2))^2, longitudinal: 0).
  e_full            -- plane-wave closed form with eps, 1/mu scaling; PEC/PMC rows and
                       columns vanish exactly; mu divides (checked by uniform mu=2 giving
                       half the curl-curl term).
  e_full_precond.   -- Pl A0 Pr is complex-symmetric (A^T = A) on random stretched
                       grids with anisotropic eps/mu (fails for a wrong index in p).
  stretch_with_scpml-- closed form of the polynomial stretch on a uniform grid.
  generic           -- dense numpy.linalg.solve brute force on 5x5x5 grids (forward
                       and adjoint, the latter against the conjugate-transposed dense
                       system); the restated reference tests test_residual and
                       test_poynting_planes (meanas/test/test_fdfd.py:14-55); the 1D
                       discrete Green's function through a PML-terminated line.
  e2h / poynting    -- pinned through test_poynting_planes (energy balance), and
                       against roll-based restatements of fdfd/functional.py:89-117,187-221.
  h_full            -- equals R e_full(omega, [rev dxh, rev dxe], rev mu, rev eps,
                       pec=rev pmc, pmc=rev pec) R, R the reversal of every cell index
                       (R S R = S^-1 turns backward differences into negated forward
                       ones); a swapped role of eps/mu or pec/pmc fails it.
  eh_full           -- with H = e2h(E) (no M, no masks) its second block row vanishes
                       and -i omega times the first equals e_full @ E.
  poynting_h_cross  -- P_h(h) @ e = -(P_e(e) @ h) against the pinned poynting_e_cross, on
                       random complex widths and fields.
  shift_with_mirror -- explicit index table on a small grid (mirror at both edges).
  m2j               -- plane-wave closed form: with uniform widths, M = a e^{ik.r} gives
                       J = (kb x a / mu) e^{ik.r} / (-i omega), kb_c = (1 - e^{-i k_c d_c}) / d_c;
                       the sparse form equals minus the functional one.
  fdtd              -- the functional curls equal the sparse curl_forward/curl_back; the
                       updates and the energy functions are pinned together by the
                       restated reference tests of meanas/test/test_fdtd.py (initial
                       fields and energy, energy conservation, Poynting divergence and
                       Poynting planes) over the reference's fixture grid.
                       delta_energy_h2e / _e2h, which those tests do not call: dt times
                       each equals the difference of the energies at its two ends
                       (energy_estep - energy_hstep, and the reverse), by linearity of dxmul.
  sources           -- e_tfsf_source, e_boundary_source, expand_e and compute_source
                       pinned together by the restated one-way PML test
                       (meanas/test/test_fdfd_pml.py:14-28): the source launches a wave
                       in one direction only (reverse |E|^2 < 1e-12, forward 1 +- 3e-6).
  cpml              -- no reference test exists (fdtd/pml.py:8 "TODO retest pmls");
                       pinned by: no layers and equal widths give the plain Yee update;
                       the profiles' mirror symmetry between polarities and their
                       closed-form sigma_max, kappa_max at full depth; sigma = 0 equals
                       the plain update on widths multiplied by kappa inside each layer
                       (overlapping layers included); a pulse along each axis and in each
                       polarization leaves through the layers without growth.
  conducting_bound. -- a bounded leapfrog run with both faces of an axis equals, plane for
                       plane, a periodic run on the doubled domain whose fields are
                       mirror-symmetric about cells 1/2 and n - 3/2 (tangential E and
                       normal H even, normal E and tangential H odd); a wrong plane, sign
                       or component breaks the equality.
Deviation, by necessity: the adjoint branch uses `.conj().T` where the reference
writes `.H`, an attribute scipy sparse arrays do not have (solvers.py:121).
"""
from __future__ import annotations

import logging
from typing import Any, Callable, Sequence

import numpy
from scipy import sparse
import scipy.sparse.linalg

logger = logging.getLogger(__name__)


# ---------------------------------------------------------------------------
# vectorization  (meanas/fdmath/vectorization.py:51-125)
# ---------------------------------------------------------------------------
def vec(f):
    """Flatten a (3, X, Y, Z) field in C order (vectorization.py:51-68)."""
    if f is None:
        return None
    return numpy.ravel(f, order='C')


def unvec(v, shape: Sequence[int], nvdim: int = 3):
    """Inverse of `vec` (vectorization.py:103-125)."""
    if v is None:
        return None
    return v.reshape((nvdim, *shape), order='C')


def _grid_shape(dx_lists) -> tuple[int, ...]:
    return tuple(int(numpy.size(d)) for d in dx_lists)


# ---------------------------------------------------------------------------
# discrete calculus  (meanas/fdmath/operators.py)
# ---------------------------------------------------------------------------
def shift_circ(axis: int, shape: Sequence[int], shift_distance: int = 1) -> sparse.csr_array:
    """Periodic shift operator (fdmath/operators.py:15-52).

    Row r of the operator selects the grid point whose coordinate along `axis` is
    (coord + |shift_distance|) mod n; a negative distance returns the transpose
    (the inverse shift).
    """
    shape = tuple(int(s) for s in shape)
    if len(shape) not in (2, 3):
        raise Exception(f'Invalid shape: {shape}')
    if axis not in range(len(shape)):
        raise Exception(f'Invalid direction: {axis}, shape is {shape}')
    n = int(numpy.prod(shape))
    flat = numpy.arange(n).reshape(shape)
    source = numpy.roll(flat, -abs(shift_distance), axis=axis).ravel()
    op = sparse.csr_array((numpy.ones(n), (numpy.arange(n), source)), shape=(n, n))
    if shift_distance < 0:
        op = op.T.tocsr()
    return op


def _expand(dx: Sequence[numpy.ndarray], axis: int) -> numpy.ndarray:
    """Broadcast the 1D widths of `axis` over the grid, flattened in C order."""
    grids = numpy.meshgrid(*dx, indexing='ij')
    return grids[axis].ravel(order='C')


def deriv_forward(dx_e: Sequence[numpy.ndarray]) -> list[sparse.csr_array]:
    """Forward differences (f[i+1]-f[i])/dx_e[i], periodic (fdmath/operators.py:99-123)."""
    shape = _grid_shape(dx_e)
    n = int(numpy.prod(shape))
    eye = sparse.eye_array(n)
    return [sparse.diags_array(1 / _expand(dx_e, a)) @ (shift_circ(a, shape, 1) - eye)
            for a in range(len(shape))]


def deriv_back(dx_h: Sequence[numpy.ndarray]) -> list[sparse.csr_array]:
    """Backward differences (f[i]-f[i-1])/dx_h[i], periodic (fdmath/operators.py:126-150)."""
    shape = _grid_shape(dx_h)
    n = int(numpy.prod(shape))
    eye = sparse.eye_array(n)
    return [sparse.diags_array(-1 / _expand(dx_h, a)) @ (shift_circ(a, shape, -1) - eye)
            for a in range(len(shape))]


def cross(B: Sequence[sparse.sparray]) -> sparse.csr_array:
    """Block operator of (B x .) (fdmath/operators.py:153-171)."""
    bx, by, bz = B
    return sparse.block_array([[None, -bz, by],
                               [bz, None, -bx],
                               [-by, bx, None]], format='csr')


def curl_forward(dx_e: Sequence[numpy.ndarray]) -> sparse.csr_array:
    """Curl acting on E (fdmath/operators.py:222-235)."""
    return cross(deriv_forward(dx_e))


def curl_back(dx_h: Sequence[numpy.ndarray]) -> sparse.csr_array:
    """Curl acting on H (fdmath/operators.py:238-251)."""
    return cross(deriv_back(dx_h))


# ---------------------------------------------------------------------------
# FDFD operators  (meanas/fdfd/operators.py)
# ---------------------------------------------------------------------------
def _mask_keep(mask, n: int):
    """diag(1 where mask == 0 else 0), identity when mask is None."""
    if mask is None:
        return sparse.eye_array(n)
    return sparse.diags_array(numpy.where(numpy.asarray(mask) != 0, 0.0, 1.0))


def e_full(omega: complex, dxes, epsilon, mu=None, pec=None, pmc=None) -> sparse.csr_array:
    """Wave operator pe (ch pm mu^-1 ce - omega^2 eps) pe (fdfd/operators.py:40-96)."""
    epsilon = numpy.asarray(epsilon)
    n = epsilon.size
    ch = curl_back(dxes[1])
    ce = curl_forward(dxes[0])
    pe = _mask_keep(pec, n)
    pm = _mask_keep(pmc, n)
    m_div = sparse.eye_array(n) if mu is None else sparse.diags_array(1 / numpy.asarray(mu))
    op = pe @ (ch @ pm @ m_div @ ce - omega ** 2 * sparse.diags_array(epsilon)) @ pe
    return op.tocsr()


def h_full(omega: complex, dxes, epsilon, mu=None, pec=None, pmc=None) -> sparse.csr_array:
    """pm (ce pe diag(1/eps) ch - omega^2 diag(mu)) pm (fdfd/operators.py:126-180)."""
    eps = numpy.asarray(epsilon)
    n = eps.size
    ch = curl_back(dxes[1])
    ce = curl_forward(dxes[0])
    pe = _mask_keep(pec, n)
    pm = _mask_keep(pmc, n)
    m = sparse.eye_array(n) if mu is None else sparse.diags_array(numpy.asarray(mu))
    return (pm @ (ce @ pe @ sparse.diags_array(1 / eps) @ ch - omega ** 2 * m) @ pm).tocsr()


def eh_full(omega: complex, dxes, epsilon, mu=None, pec=None, pmc=None) -> sparse.csr_array:
    """[[-pe (i w eps) pe, pe ch pm], [pm ce pe, pm (i w mu) pm]] acting on [E; H]
    (fdfd/operators.py:183-250)."""
    eps = numpy.asarray(epsilon)
    n = eps.size
    pe = _mask_keep(pec, n)
    pm = _mask_keep(pmc, n)
    iwe = pe @ (1j * omega * sparse.diags_array(eps)) @ pe
    iwm = 1j * omega * (sparse.eye_array(n) if mu is None else sparse.diags_array(numpy.asarray(mu)))
    iwm = pm @ iwm @ pm
    a1 = pe @ curl_back(dxes[1]) @ pm
    a2 = pm @ curl_forward(dxes[0]) @ pe
    return sparse.block_array([[-iwe, a1], [a2, iwm]], format='csr')


def e_full_preconditioner_vector(dxes) -> numpy.ndarray:
    """p = sqrt(p^2) with p^2 the per-component products of widths (fdfd/operators.py:116-120)."""
    (ex, ey, ez), (hx, hy, hz) = dxes
    p_squared = [ex[:, None, None] * hy[None, :, None] * hz[None, None, :],
                 hx[:, None, None] * ey[None, :, None] * hz[None, None, :],
                 hx[:, None, None] * hy[None, :, None] * ez[None, None, :]]
    return numpy.sqrt(vec(numpy.array(p_squared, dtype=complex)))


def e_full_preconditioners(dxes) -> tuple[sparse.sparray, sparse.sparray]:
    """(Pl, Pr) = (diag(p), diag(1/p)) (fdfd/operators.py:99-123)."""
    p = e_full_preconditioner_vector(dxes)
    return sparse.diags_array(p), sparse.diags_array(1 / p)


def e2h(omega: complex, dxes, mu=None, pmc=None) -> sparse.csr_array:
    """E -> H conversion, curl_f / (-i omega) then 1/mu then the PMC mask (fdfd/operators.py:253-282)."""
    op = curl_forward(dxes[0]) / (-1j * omega)
    if mu is not None:
        op = sparse.diags_array(1 / numpy.asarray(mu)) @ op
    if pmc is not None:
        op = sparse.diags_array(numpy.where(numpy.asarray(pmc) != 0, 0.0, 1.0)) @ op
    return op.tocsr()


def poynting_e_cross(e, dxes) -> sparse.csr_array:
    """(E x) part of the Poynting vector with face-area weights (fdfd/operators.py:310-336)."""
    shape = _grid_shape(dxes[0])
    fx, fy, fz = (shift_circ(a, shape, 1) for a in range(3))
    dxa = [d.ravel(order='C') for d in numpy.meshgrid(*dxes[0], indexing='ij')]
    dxb = [d.ravel(order='C') for d in numpy.meshgrid(*dxes[1], indexing='ij')]
    ex, ey, ez = (c * w for c, w in zip(numpy.split(numpy.asarray(e), 3), dxa))
    dg = sparse.diags_array
    # each block is the diagonal of a shifted (width-weighted) E component
    blocks = [[None, dg(fx @ -ez), dg(fx @ ey)],
              [dg(fy @ ez), None, dg(fy @ -ex)],
              [dg(fz @ -ey), dg(fz @ ex), None]]
    return (sparse.block_array(blocks) @ dg(numpy.concatenate(dxb))).tocsr()


def poynting_h_cross(h, dxes) -> sparse.csr_array:
    """(H x) part of the Poynting vector (fdfd/operators.py:339-363): P @ e = -(E x H)
    with E shifted one cell forward and both fields width-weighted."""
    shape = _grid_shape(dxes[0])
    fx, fy, fz = (shift_circ(a, shape, 1) for a in range(3))
    dxa = [d.ravel(order='C') for d in numpy.meshgrid(*dxes[0], indexing='ij')]
    dxb = [d.ravel(order='C') for d in numpy.meshgrid(*dxes[1], indexing='ij')]
    dg = sparse.diags_array
    hx, hy, hz = (dg(c * w) for c, w in zip(numpy.split(numpy.asarray(h), 3), dxb))
    # each block is a width-weighted H component times a forward shift of E
    blocks = [[None, -hz @ fx, hy @ fx],
              [hz @ fy, None, -hx @ fy],
              [-hy @ fz, hx @ fz, None]]
    return (sparse.block_array(blocks) @ dg(numpy.concatenate(dxa))).tocsr()


def shift_with_mirror(axis: int, shape: Sequence[int], shift_distance: int = 1) -> sparse.csr_array:
    """Shift by `shift_distance` along `axis`, mirroring past the edges
    (fdmath/operators.py:55-99): (S f)[n] = f[m(n + d)], m reflecting at both ends."""
    if shift_distance >= shape[axis]:
        raise ValueError('shift too large for the axis')

    def mirrored(n: int, d: int) -> numpy.ndarray:
        v = numpy.arange(n) + d
        v = numpy.where(v >= n, 2 * n - v - 1, v)
        return numpy.where(v < 0, -1 - v, v)

    idx = [mirrored(n, shift_distance if a == axis else 0) for a, n in enumerate(shape)]
    cols = numpy.ravel_multi_index(numpy.meshgrid(*idx, indexing='ij'), shape).ravel()
    n = int(numpy.prod(shape))
    return sparse.csr_array((numpy.ones(n), (numpy.arange(n), cols)), shape=(n, n))


def e_tfsf_source(TF_region, omega: complex, dxes, epsilon, mu=None) -> sparse.csr_array:
    """(A Q - Q A) / (-i omega), A = e_full, Q = diag(TF_region) (fdfd/operators.py:366-394)."""
    A = e_full(omega, dxes, epsilon, mu)
    Q = sparse.diags_array(numpy.asarray(TF_region, dtype=float))
    return ((A @ Q - Q @ A) / (-1j * omega)).tocsr()


def e_boundary_source(mask, omega: complex, dxes, epsilon, mu=None,
                      periodic_mask_edges: bool = False) -> sparse.csr_array:
    """e_tfsf_source restricted to the cells where a one-cell shift of the mask (along
    any axis longer than 1, either way; mirror or periodic edges) changes it
    (fdfd/operators.py:397-450)."""
    full = e_tfsf_source(mask, omega, dxes, epsilon, mu)
    shape = _grid_shape(dxes[0])
    n = int(numpy.prod(shape))
    m = numpy.asarray(mask, dtype=float)
    jmask = numpy.zeros(m.shape, dtype=bool)
    for axis in range(3):
        if shape[axis] == 1:
            continue
        for pol in (-1, 1):
            s = shift_circ(axis, shape, pol) if periodic_mask_edges else shift_with_mirror(axis, shape, pol)
            r = s - sparse.eye_array(n)
            jmask |= numpy.abs(sparse.block_diag((r, r, r)) @ m) != 0
    return (sparse.diags_array(jmask.astype(float)) @ full).tocsr()


def expand_e(E, wavenumber: complex, dxes, axis: int, polarity: int, slices) -> numpy.ndarray:
    """Copy the slice `slices` of E along `axis` with the propagation phase
    exp(-i polarity k r), r the E-grid position relative to the slice
    (fdfd/waveguide_3d.py:206-252)."""
    slices = tuple(slices)
    a_shape = numpy.roll([1, -1, 1, 1], axis)
    pos = numpy.real(dxes[0][axis]).cumsum()
    r = pos - pos[slices[axis]]
    phase = numpy.exp(polarity * -1j * wavenumber * r).reshape(a_shape)
    out = numpy.zeros_like(E)
    full = list(slices)
    full[axis] = slice(E.shape[axis + 1])
    out[(slice(None), *full)] = phase * numpy.asarray(E)[(slice(None), *slices)]
    return out


def compute_source(E, wavenumber: complex, omega: complex, dxes, axis: int, polarity: int, slices,
                   epsilon, mu=None) -> numpy.ndarray:
    """One-way source: e_boundary_source of the half-space mask starting at the slice,
    applied to the expanded field (fdfd/waveguide_3d.py:113-156)."""
    Ex = expand_e(E, wavenumber, dxes, axis, polarity, slices)
    smask = [slice(None)] * 4
    smask[axis + 1] = slice(slices[axis].start, None) if polarity > 0 else slice(None, slices[axis].stop)
    mask = numpy.zeros_like(Ex, dtype=int)
    mask[tuple(smask)] = 1
    eps_v = vec(epsilon)
    mu_v = None if mu is None else vec(mu)
    J = e_boundary_source(vec(mask), omega, dxes, eps_v, mu_v) @ vec(Ex)
    return unvec(J, E.shape[1:])


# ---------------------------------------------------------------------------
# SC-PML  (meanas/fdfd/scpml.py)
# ---------------------------------------------------------------------------
def prepare_s_function(ln_R: float = -16, m: float = 4) -> Callable[[numpy.ndarray], numpy.ndarray]:
    """s(d) = (m+1) ln_R / 2 * d^m (scpml.py:18-38)."""
    s_max = (m + 1) * ln_R / 2

    def s_factor(distance):
        return s_max * distance ** m
    return s_factor


def stretch_with_scpml(dxes, axis: int, polarity: int, omega: float,
                       epsilon_effective: float = 1.0, thickness: int = 10,
                       s_function=None):
    """Stretch one face of the grid with an SC-PML (scpml.py:98-167); mutates and returns dxes."""
    if s_function is None:
        s_function = prepare_s_function()
    dx_e = numpy.array(dxes[0][axis], dtype=complex)
    dx_h = numpy.array(dxes[1][axis], dtype=complex)

    edges = numpy.concatenate(([0], numpy.cumsum(dx_e)))
    centers = 0.5 * (edges[:-1] + edges[1:])
    lefts = edges[:-1]
    correction = numpy.sqrt(epsilon_effective) * numpy.real(omega)

    if polarity > 0:
        bound = edges[thickness]
        depth = bound - edges[0]
        region = slice(thickness)

        def frac(u):
            return (bound - u) / (bound - edges[0])
    else:
        bound = edges[-thickness - 1]
        depth = edges[-1] - bound
        region = slice(None) if thickness == 0 else slice(-thickness, None)

        def frac(u):
            return (u - bound) / (edges[-1] - bound)

    dx_e[region] *= 1 + 1j * s_function(frac(centers[region])) / depth / correction
    dx_h[region] *= 1 + 1j * s_function(frac(lefts[region])) / depth / correction
    dxes[0][axis] = dx_e
    dxes[1][axis] = dx_h
    return dxes


# ---------------------------------------------------------------------------
# solver  (meanas/fdfd/solvers.py)
# ---------------------------------------------------------------------------
def scipy_qmr(A, b, **kwargs: Any) -> numpy.ndarray:
    """scipy QMR with the reference's every-100-iterations residual log (solvers.py:21-63).

    Returns the iterate whether or not it converged; the info code is dropped as in
    the reference.  The last info code is kept in `scipy_qmr.last_info` for tests.
    """
    count = 0

    def log_residual(xk):
        nonlocal count
        count += 1
        if count % 100 == 0:
            res = numpy.linalg.norm(A @ xk - b) / numpy.linalg.norm(b)
            logger.info('Solver residual at iteration %d : %g', count, res)

    user_cb = kwargs.get('callback')
    if user_cb is not None:
        def both(xk):
            log_residual(xk)
            user_cb(xk)
        kwargs['callback'] = both
    else:
        kwargs['callback'] = log_residual
    x, info = scipy.sparse.linalg.qmr(A, b, **kwargs)
    scipy_qmr.last_info = info
    scipy_qmr.last_iterations = count
    return x


scipy_qmr.last_info = None
scipy_qmr.last_iterations = 0


def generic(omega: complex, dxes, J, epsilon, mu=None, *, pec=None, pmc=None,
            adjoint: bool = False, matrix_solver: Callable[..., Any] = scipy_qmr,
            matrix_solver_opts: dict | None = None, E_guess=None) -> numpy.ndarray:
    """FDFD solve through a sparse matrix solver (solvers.py:66-141)."""
    if matrix_solver_opts is None:
        matrix_solver_opts = {}
    b0 = -1j * omega * numpy.asarray(J)
    A0 = e_full(omega, dxes, epsilon=epsilon, mu=mu, pec=pec, pmc=pmc)
    Pl, Pr = e_full_preconditioners(dxes)
    if adjoint:
        A = (Pl @ A0 @ Pr).conj().T
        b = Pr.conj().T @ b0
    else:
        A = Pl @ A0 @ Pr
        b = Pl @ b0
    if E_guess is not None:
        x0 = (Pr.conj().T @ E_guess) if adjoint else (Pl @ E_guess)
        matrix_solver_opts['x0'] = x0
    x = matrix_solver(A.tocsr(), b, **matrix_solver_opts)
    return (Pl.conj().T @ x) if adjoint else (Pr @ x)


def dense_solve(omega: complex, dxes, J, epsilon, mu=None, pec=None, pmc=None,
                adjoint: bool = False) -> numpy.ndarray:
    """Brute-force dense solve of the same system, for tiny grids only.

    PEC rows/columns of e_full are zero; like a Krylov solver started from zero,
    the minimum-norm solution leaves those entries at zero (lstsq).
    """
    A0 = e_full(omega, dxes, epsilon=epsilon, mu=mu, pec=pec, pmc=pmc).toarray()
    b0 = -1j * omega * numpy.asarray(J)
    if adjoint:
        Pl, Pr = (m.toarray() for m in e_full_preconditioners(dxes))
        A = (Pl @ A0 @ Pr).conj().T
        z = numpy.linalg.lstsq(A, Pr.conj().T @ b0, rcond=None)[0]
        return Pl.conj().T @ z
    return numpy.linalg.lstsq(A0, b0, rcond=None)[0]


# ---------------------------------------------------------------------------
# FDTD  (meanas/fdmath/functional.py curls; meanas/fdtd/base.py, energy.py)
# ---------------------------------------------------------------------------
def _unit_dx(dx):
    return [numpy.ones(1)] * 3 if dx is None else list(dx)


def f_curl_forward(dx_e=None) -> Callable:
    """f -> curl with forward differences (f[n+1] - f[n]) / dx_e[n] (fdmath/functional.py)."""
    dx = _unit_dx(dx_e)

    def d(f, a):
        w = dx[a].reshape([-1 if b == a else 1 for b in range(3)])
        return (numpy.roll(f, -1, axis=a) - f) / w

    def curl(e):
        return numpy.stack([d(e[2], 1) - d(e[1], 2), d(e[0], 2) - d(e[2], 0), d(e[1], 0) - d(e[0], 1)])
    return curl


def f_curl_back(dx_h=None) -> Callable:
    """h -> curl with backward differences (f[n] - f[n-1]) / dx_h[n] (fdmath/functional.py)."""
    dx = _unit_dx(dx_h)

    def d(f, a):
        w = dx[a].reshape([-1 if b == a else 1 for b in range(3)])
        return (f - numpy.roll(f, 1, axis=a)) / w

    def curl(h):
        return numpy.stack([d(h[2], 1) - d(h[1], 2), d(h[0], 2) - d(h[2], 0), d(h[1], 0) - d(h[0], 1)])
    return curl


def m2j_functional(omega, dxes, mu=None) -> Callable:
    """J = curl_back(M / mu) / (-i omega) on (3, nx, ny, nz) arrays (fdfd/functional.py:120-151)."""
    ch = f_curl_back(dxes[1])
    if mu is None:
        return lambda m: ch(m) / (-1j * omega)
    return lambda m: ch(m / mu) / (-1j * omega)


def m2j(omega, dxes, mu=None):
    """Sparse form (fdfd/operators.py:285-307): curl_back / (i omega), times diag(1/mu)
    on the right.  Its sign is opposite to the functional form's."""
    op = curl_back(dxes[1]) / (1j * omega)
    if mu is not None:
        op = op @ sparse.diags_array(1 / mu)
    return op


def maxwell_e(dt: float, dxes=None) -> Callable:
    """e += dt curl_back(h) / epsilon, in place (fdtd/base.py:13-64)."""
    ch = f_curl_back(None if dxes is None else dxes[1])

    def me(e, h, epsilon):
        e += dt * ch(h) / epsilon
        return e
    return me


def maxwell_h(dt: float, dxes=None) -> Callable:
    """h -= dt curl_forward(e) / mu, in place (fdtd/base.py:67-125)."""
    ce = f_curl_forward(None if dxes is None else dxes[0])

    def mh(e, h, mu=None):
        h -= dt * ce(e) / (1 if mu is None else mu)
        return h
    return mh


def _weigh(u, dx):
    """u * dx_x * dx_y * dx_z, one axis at a time, as the reference's expressions multiply."""
    return u * dx[0][:, None, None] * dx[1][None, :, None] * dx[2][None, None, :]


def _unit_dxes(dxes):
    return [[numpy.ones(1)] * 3] * 2 if dxes is None else dxes


def fdtd_poynting(e, h, dxes=None):
    """E x H with E shifted forward, width-weighted (fdtd/energy.py:10-87); the products
    are float64 (the widths promote) and the result has e's dtype."""
    dxes = _unit_dxes(dxes)
    ex, ey, ez = (e[c] * dxes[0][c].reshape([-1 if b == c else 1 for b in range(3)]) for c in range(3))
    hx, hy, hz = (h[c] * dxes[1][c].reshape([-1 if b == c else 1 for b in range(3)]) for c in range(3))
    s = numpy.empty_like(e)
    s[0] = numpy.roll(ey, -1, 0) * hz - numpy.roll(ez, -1, 0) * hy
    s[1] = numpy.roll(ez, -1, 1) * hx - numpy.roll(ex, -1, 1) * hz
    s[2] = numpy.roll(ex, -1, 2) * hy - numpy.roll(ey, -1, 2) * hx
    return s


def poynting_divergence(s=None, *, e=None, h=None, dxes=None):
    """Backward-difference divergence of S with unit spacing, as the reference computes it
    (deriv_back() without widths, fdtd/energy.py:90-125)."""
    if s is None:
        s = fdtd_poynting(e, h, dxes)
    d = [s[a] - numpy.roll(s[a], 1, axis=a) for a in range(3)]
    return d[0] + d[1] + d[2]


def dxmul(ee, hh, epsilon=None, mu=None, dxes=None):
    """sum_c ee eps dV_e + sum_c hh mu dV_h (fdtd/energy.py:273-295)."""
    epsilon = 1 if epsilon is None else epsilon
    mu = 1 if mu is None else mu
    dxes = _unit_dxes(dxes)
    return _weigh((ee * epsilon).sum(axis=0), dxes[0]) + _weigh((hh * mu).sum(axis=0), dxes[1])


def energy_hstep(e0, h1, e2, epsilon=None, mu=None, dxes=None):
    """Energy at an H step: dxmul(e0 e2, h1 h1) (fdtd/energy.py:128-153)."""
    return dxmul(e0 * e2, h1 * h1, epsilon, mu, dxes)


def energy_estep(h0, e1, h2, epsilon=None, mu=None, dxes=None):
    """Energy at an E step: dxmul(e1 e1, h0 h2) (fdtd/energy.py:156-181)."""
    return dxmul(e1 * e1, h0 * h2, epsilon, mu, dxes)


def delta_energy_h2e(dt, e0, h1, e2, h3, epsilon=None, mu=None, dxes=None):
    """Energy change from h1 to e2: dxmul(e2 (e2 - e0) / dt, h1 (h3 - h1) / dt)
    (fdtd/energy.py:184-214)."""
    return dxmul(e2 * (e2 - e0) / dt, h1 * (h3 - h1) / dt, epsilon, mu, dxes)


def delta_energy_e2h(dt, h0, e1, h2, e3, epsilon=None, mu=None, dxes=None):
    """Energy change from e1 to h2: dxmul(e1 (e3 - e1) / dt, h2 (h2 - h0) / dt)
    (fdtd/energy.py:217-247)."""
    return dxmul(e1 * (e3 - e1) / dt, h2 * (h2 - h0) / dt, epsilon, mu, dxes)


def delta_energy_j(j0, e1, dxes=None):
    """(j0 . e1) dV_e (fdtd/energy.py:250-270)."""
    return _weigh((j0 * e1).sum(axis=0), _unit_dxes(dxes)[0])


# ---------------------------------------------------------------------------
# FDTD convolutional PML  (meanas/fdtd/pml.py:23-230)
# ---------------------------------------------------------------------------
def cpml_params(axis: int, polarity: int, dt: float, thickness: int = 8, ln_R_per_layer: float = -1.6,
                epsilon_eff: float = 1, mu_eff: float = 1, m: float = 3.5, ma: float = 1,
                cfs_alpha: float = 0) -> dict:
    """CPML profiles of one layer (pml.py:23-93).

    Slab index l = 0..t-1 (cells [0, t) for polarity -1, [n - t, n) for +1).  Depth x of
    the sample points, in cells from the inner face: polarity +1 has E at l + 1/2 and H
    at l + 1; polarity -1 mirrors them (E at t - l, H at t - l - 1/2).  With s = (x/t)^m:
    sigma = s sigma_max, kappa = 1 + s (kappa_max - 1), alpha = (1 - x/t)^ma alpha_max,
    p0 = exp(-(sigma/kappa + alpha) dt), p1 = sigma / (sigma + kappa alpha) (p0 - 1),
    p2 = 1/kappa; sigma_max = -ln_R_per_layer (m + 1) / 2, kappa_max = sqrt(eps_eff mu_eff).
    """
    if axis not in (0, 1, 2):
        raise ValueError(f'Invalid axis: {axis}')
    if polarity not in (-1, 1):
        raise ValueError(f'Invalid polarity: {polarity}')
    if thickness <= 2:
        raise ValueError('It would be wise to have a pml with 4+ cells of thickness')
    if epsilon_eff <= 0:
        raise ValueError('epsilon_eff must be positive')
    t = thickness
    l = numpy.arange(t, dtype=float)
    if polarity > 0:
        xe, xh = l + 0.5, l + 1.0
    else:
        xe, xh = t - l, t - l - 0.5
    sigma_max = -ln_R_per_layer / 2 * (m + 1)
    kappa_max = numpy.sqrt(epsilon_eff * mu_eff)
    shape = [1, 1, 1]
    shape[axis] = t

    def profiles(x):
        s = (x / t) ** m
        sigma = sigma_max * s
        kappa = 1 + (kappa_max - 1) * s
        alpha = cfs_alpha * (1 - x / t) ** ma
        p0 = numpy.exp(-(sigma / kappa + alpha) * dt)
        p1 = sigma / (sigma + kappa * alpha) * (p0 - 1)
        return tuple(p.reshape(shape) for p in (p0, p1, 1 / kappa))

    region = [slice(None)] * 3
    region[axis] = slice(None, t) if polarity < 0 else slice(-t, None)
    return {'param_e': profiles(xe), 'param_h': profiles(xh), 'region': tuple(region)}


def updates_with_cpml(cpml_params, dt: float, dxes, epsilon, *, dtype=numpy.float32):
    """(update_E(e, h, epsilon), update_H(e, h, mu=(1, 1, 1))) with CPML layers
    (pml.py:96-230).  cpml_params[axis][p] (p = 0: polarity -1, 1: +1) is a layer dict or
    None.  Both difference directions use dxes[1] (pml.py:106-107).  Per update the six
    derivatives d_a f_c (a != c) are formed; each layer on axis a, in the order axis 0..2,
    polarity -1 then +1, scales d_a f_u and d_a f_v (u < v the components other than a) by
    p2 inside its region, then advances psi_u = p0 psi_u + p1 d_a f_v and psi_v = p0 psi_v +
    p1 d_a f_u (psi in `dtype`), and adds s psi_u to the auxiliary term of component u and
    -s psi_v to that of v, s = (-1, 1, -1)[a].  Then e_c += dt / eps_c (curl_c + pE_c) and
    h_c -= dt / mu_c (curl_c + pH_c) with the scaled derivatives."""
    w = [numpy.asarray(dxes[1][a]).reshape([-1 if b == a else 1 for b in range(3)]) for a in range(3)]
    layers = []
    for axis in range(3):
        for pp in range(2):
            c = cpml_params[axis][pp]
            if c is None:
                continue
            rshape = epsilon[0][c['region']].shape
            layers.append(dict(axis=axis, region=c['region'], pe=c['param_e'], ph=c['param_h'],
                               psi_e=[numpy.zeros(rshape, dtype=dtype) for _ in range(2)],
                               psi_h=[numpy.zeros(rshape, dtype=dtype) for _ in range(2)]))
    acc = numpy.empty_like(epsilon, dtype=dtype)

    def absorb(d, key_p, key_psi):
        acc.fill(0)
        for lay in layers:
            a = lay['axis']
            u, v = [c for c in range(3) if c != a]
            p0, p1, p2 = lay[key_p]
            psi_u, psi_v = lay[key_psi]
            reg = lay['region']
            d[a, u][reg] *= p2
            d[a, v][reg] *= p2
            psi_u *= p0
            psi_v *= p0
            psi_u += p1 * d[a, v][reg]
            psi_v += p1 * d[a, u][reg]
            s = (-1, 1, -1)[a]
            acc[u][reg] += s * psi_u
            acc[v][reg] += -s * psi_v

    def update_E(e, h, epsilon):
        d = {(a, c): (h[c] - numpy.roll(h[c], 1, axis=a)) / w[a] for a in range(3) for c in range(3) if a != c}
        absorb(d, 'pe', 'psi_e')
        for c in range(3):
            c1, c2 = (c + 1) % 3, (c + 2) % 3
            e[c] += dt / epsilon[c] * (d[c1, c2] - d[c2, c1] + acc[c])

    def update_H(e, h, mu=(1, 1, 1)):
        d = {(a, c): (numpy.roll(e[c], -1, axis=a) - e[c]) / w[a] for a in range(3) for c in range(3) if a != c}
        absorb(d, 'ph', 'psi_h')
        for c in range(3):
            c1, c2 = (c + 1) % 3, (c + 2) % 3
            h[c] -= dt / mu[c] * (d[c1, c2] - d[c2, c1] + acc[c])

    return update_E, update_H


# ---------------------------------------------------------------------------
# FDTD conducting boundary  (meanas/fdtd/boundaries.py:12-78)
# ---------------------------------------------------------------------------
def conducting_boundary(direction: int, polarity: int) -> tuple[Callable, Callable]:
    """(update_e, update_h) of boundaries.py:12-78 with tuple indices.  The reference
    indexes `e[direction][boundary_slice]` with a list of slices, which numpy >= 1.23
    rejects; the updates are what the same assignments do with `tuple(slice_list)`."""
    if direction not in (0, 1, 2):
        raise Exception(f'Invalid direction: {direction}')
    u, v = [a for a in range(3) if a != direction]

    def plane(index):
        sl = [slice(None)] * 3
        sl[direction] = index
        return tuple(sl)

    if polarity < 0:
        b0, s1 = plane(0), plane(1)

        def en(e):
            e[direction][b0] = 0
            e[u][b0] = e[u][s1]
            e[v][b0] = e[v][s1]
            return e

        def hn(h):
            h[direction][b0] = h[direction][s1]
            h[u][b0] = 0
            h[v][b0] = 0
            return h

        return en, hn
    if polarity > 0:
        b0, s1, s2 = plane(-1), plane(-2), plane(-3)

        def ep(e):
            e[direction][b0] = -e[direction][s2]
            e[direction][s1] = 0
            e[u][b0] = e[u][s1]
            e[v][b0] = e[v][s1]
            return e

        def hp(h):
            h[direction][b0] = h[direction][s1]
            h[u][b0] = -h[u][s2]
            h[u][s1] = 0
            h[v][b0] = -h[v][s2]
            h[v][s1] = 0
            return h

        return ep, hp
    raise Exception(f'Bad polarity: {polarity}')
