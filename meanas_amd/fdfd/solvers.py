"""Drop-in `generic()` FDFD solver running on one MI355X.

Same call signature, argument meaning and return type as
meanas.fdfd.solvers.generic (reference meanas/fdfd/solvers.py:66-141):

    E = generic(omega, dxes, J, epsilon, mu=None, *, pec=None, pmc=None, adjoint=False,
                matrix_solver=..., matrix_solver_opts=None, E_guess=None)

It solves the same symmetrised system (solvers.py:115-125): b = Pl (-i omega J),
A = Pl e_full Pr, E = Pr x (adjoint: A^H, b = Pr^H b0, E = Pl^H x, solvers.py:120-137),
but matrix-free on the GPU with a complex-symmetric Krylov method (COCG) instead of
scipy's qmr on an assembled CSR matrix.  Because A = A^T, the adjoint system is the
forward operator applied to conjugated data: E_adj = conj(Pl w) with A w = Pr conj(b0).

matrix_solver selects the GPU method: 'cocg_fp64' (complex128 COCG, the default) or
'cocg_mixed' (complex64 iterations with FP64 residual replacement and FP64 solution
accumulation).  Complex64 rounding in the COCG recurrence makes it diverge on PML
workloads of 48^3 and up; such a solve stalls and finishes by an FP64 continuation from
its best iterate, and says so in last_info (fp64_fallback, mixed_iterations).  A
CSR-based callable, as the reference accepts, cannot be honoured by a matrix-free solver
and raises TypeError.
matrix_solver_opts honours scipy qmr's keys: rtol (default 1e-5), atol (default 0),
maxiter (default 10 * 3N) and x0 (scaled initial guess); as in the reference, passing
E_guess writes the scaled guess into matrix_solver_opts['x0'] (solvers.py:127-132).
One key beyond scipy's: form ('auto', 'fused' or 'classic', the COCG iteration form of
mx_options; 'auto' is fused for FP64).
Convergence means ||b - A x|| <= max(rtol ||b||, atol) for the FP64 true residual of
the scaled system.  Like the reference, a solve that does not converge still returns
an iterate -- here the best one (lowest FP64 true residual seen), where the reference
returns qmr's last; the status is logged and kept in `last_info`.  A J with a part on
PEC entries (where e_full has zero rows, so no E matches it) is solved in the
least-squares sense and reported as 'inconsistent' with its residual floor.  Near-static
problems (omega^2 eps below complex64 resolution of the curl-curl term) go to FP64 up
front even when 'cocg_mixed' is asked for.
"""
from __future__ import annotations

import logging
from typing import Any

import numpy
import torch

from .. import _lib
from .device import DeviceOperator, grid_shape, solve_batch

logger = logging.getLogger(__name__)

METHODS = {'cocg_mixed': _lib.MX_MIXED, 'cocg_fp64': _lib.MX_FP64}

# complex64 cannot resolve the operator when omega^2 eps falls below this fraction of the
# curl-curl scale 4/dx^2: the gradient (longitudinal) subspace of A has eigenvalues
# -omega^2 eps, and complex64 rounding of the curl-curl term is ~1e-7 of its scale.
NEAR_STATIC = 1e-5


def near_static(omega: complex, dxes, epsilon) -> bool:
    """True when complex64 arithmetic cannot represent the operator's gradient subspace."""
    dmin = min(float(numpy.min(numpy.abs(d))) for d in dxes[0] + dxes[1])
    cc = 4.0 / dmin ** 2
    w2 = abs(complex(omega)) ** 2 * float(numpy.min(numpy.abs(epsilon)))
    return w2 < NEAR_STATIC * cc

#: stats of the most recent generic() call (status, iterations, residual, ...)
last_info: dict[str, Any] = {}


def _method(matrix_solver) -> int:
    if matrix_solver is None:
        return _lib.MX_FP64
    if isinstance(matrix_solver, str):
        if matrix_solver not in METHODS:
            raise ValueError(f'unknown GPU method {matrix_solver!r}; choose one of {sorted(METHODS)}')
        return METHODS[matrix_solver]
    raise TypeError('meanas_amd.fdfd.solvers.generic is matrix-free: matrix_solver must be None or one of '
                    f'{sorted(METHODS)}, not a callable taking an assembled sparse matrix')


def generic(omega: complex, dxes, J, epsilon, mu=None, *, pec=None, pmc=None, adjoint: bool = False,
            matrix_solver=None, matrix_solver_opts: dict[str, Any] | None = None, E_guess=None,
            device: str | torch.device = 'cuda') -> numpy.ndarray:
    """Solve the FDFD wave equation for E on the GPU (see the module docstring)."""
    global last_info
    if matrix_solver_opts is None:
        matrix_solver_opts = dict()
    mode = _method(matrix_solver)
    unknown = set(matrix_solver_opts) - {'rtol', 'atol', 'maxiter', 'x0', 'callback', 'tol', 'form'}
    if unknown:
        raise TypeError(f'unsupported matrix_solver_opts keys: {sorted(unknown)}')
    shape = grid_shape(dxes)
    n3 = 3 * int(numpy.prod(shape))
    J = numpy.asarray(J)
    if J.size != n3:
        raise ValueError(f'J has {J.size} entries, expected {n3}')
    rtol = float(matrix_solver_opts.get('rtol', matrix_solver_opts.get('tol', 1e-5)))
    atol = float(matrix_solver_opts.get('atol', 0.0))
    maxiter = int(matrix_solver_opts.get('maxiter', 10 * n3))
    if 'callback' in matrix_solver_opts:
        logger.warning('matrix_solver_opts["callback"] is ignored: iterates stay on the GPU')

    fp64_reason = None
    if mode == _lib.MX_MIXED and near_static(omega, dxes, epsilon):
        mode, fp64_reason = _lib.MX_FP64, 'near-static: omega^2 eps below complex64 resolution'
    op = DeviceOperator(omega, dxes, epsilon, mu=mu, pec=pec, pmc=pmc, device=device)
    dev = op.device
    J_d = torch.from_numpy(numpy.ascontiguousarray(J, dtype=numpy.complex128)).to(dev)
    b = op.map(_lib.MX_MAP_RHS_ADJ if adjoint else _lib.MX_MAP_RHS, J_d, coef=-1j * complex(omega))
    del J_d

    # Scaled guesses cross the boundary in the reference's convention (p = sqrt(p^2),
    # principal root); internally p is factored per axis and may differ by a sign per
    # entry, which MX_MAP_SIGN applies (an involution).
    if E_guess is not None:
        g = torch.from_numpy(numpy.ascontiguousarray(E_guess, dtype=numpy.complex128)).to(dev)
        x = op.map(_lib.MX_MAP_RHS_ADJ if adjoint else _lib.MX_MAP_RHS, g, coef=1.0)
        x0_host = op.map(_lib.MX_MAP_SIGN, x).cpu().numpy()
        # the reference stores Pl E_guess (forward) or Pr^H E_guess (adjoint)
        matrix_solver_opts['x0'] = numpy.conj(x0_host) if adjoint else x0_host
    elif matrix_solver_opts.get('x0') is not None:
        x0 = numpy.asarray(matrix_solver_opts['x0'], dtype=numpy.complex128)
        if adjoint:
            x0 = numpy.conj(x0)
        x = op.map(_lib.MX_MAP_SIGN, torch.from_numpy(numpy.ascontiguousarray(x0)).to(dev))
    else:
        x = torch.zeros(n3, dtype=torch.complex128, device=dev)

    if float(torch.linalg.vector_norm(b).item()) == 0.0:
        # scipy qmr returns b itself when ||b|| = 0 (sparse/linalg/_isolve/iterative.py:909-910)
        last_info = dict(status='converged', iterations=0, replacements=0, rel_residual=0.0, bnorm=0.0)
        return numpy.zeros(n3, dtype=numpy.complex128)

    form = matrix_solver_opts.get('form', 'auto')
    if form not in _lib.FORMS:
        raise ValueError(f"matrix_solver_opts['form'] must be one of {sorted(_lib.FORMS)}")
    if mode == _lib.MX_MIXED and form == 'fused':
        form = 'auto'   # complex64 inner vectors always use the classic form
    stats = solve_batch([op], [dict(op=0, b=b, x=x, rtol=rtol, atol=atol, maxiter=maxiter)], mode=mode,
                        options=dict(form=form))[0]
    E = op.map(_lib.MX_MAP_E_ADJ if adjoint else _lib.MX_MAP_E, x)
    last_info = dict(stats, method='cocg_fp64' if mode == _lib.MX_FP64 else 'cocg_mixed', adjoint=adjoint)
    if fp64_reason:
        last_info['fp64_reason'] = fp64_reason
    if stats['status'] != 'converged':
        logger.warning('FDFD solve did not converge: %s', last_info)
    else:
        logger.info('FDFD solve: %s', last_info)
    return E.cpu().numpy()
