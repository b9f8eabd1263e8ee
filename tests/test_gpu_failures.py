"""Forced failure paths of the solve (the conditions are set up on the host; every kernel
runs to its normal end):

  * a request larger than the server's HBM budget is refused before anything is
    allocated for it, and the server stays usable;
  * a source inside PEC cells makes A x = b inconsistent (e_full has zero rows there,
    fdfd/operators.py:84-95): the solve returns the least-squares solution, equal to the
    oracle's dense lstsq solution, with status 'inconsistent' and the residual floor it
    cannot go below.  The reference's qmr runs on silently (fdfd/solvers.py:34-36,62).
"""
import numpy
import pytest

from oracle import fdfd_ref as ref

pytestmark = pytest.mark.gpu


def pec_problem(shape=(6, 7, 9), seed=50):
    g = numpy.random.default_rng(seed)
    omega = 2 * numpy.pi / 9.5
    dxes = [[numpy.ones(n) for n in shape] for _ in range(2)]
    for a, n in enumerate(shape):
        for p in (-1, 1):
            ref.stretch_with_scpml(dxes, axis=a, polarity=p, omega=omega, thickness=2)
    n3 = 3 * int(numpy.prod(shape))
    eps = g.uniform(1, 4, n3)
    pec = (g.uniform(size=n3) < 0.1).astype(float)
    J = g.standard_normal(n3) + 1j * g.standard_normal(n3)    # nonzero on PEC entries too
    return omega, dxes, J, eps, pec


def test_oversized_request_is_refused_before_allocation():
    import torch
    from meanas_amd import _lib, workloads
    from meanas_amd.serve import Request, Server, footprint_bytes
    g = numpy.random.default_rng(1)

    def request(n, tag):
        shape = (n, n, n)
        omega = 2 * numpy.pi / 12.0
        J = workloads.gaussian_sheet(shape, n // 2, 2.0, workloads.random_polarisation(g)).ravel()
        return Request(omega=omega, dxes=workloads.pml_dxes(shape, omega, thickness=4), J=J,
                       epsilon=workloads.aniso_structure(shape, g, nbox=3).ravel(), rtol=1e-9, tag=tag)
    req_small, req_big = request(16, 's'), request(40, 'b')
    budget = footprint_bytes(req_small, _lib.MX_FP64) + 1024
    assert footprint_bytes(req_big, _lib.MX_FP64) > budget
    srv = Server(capacity=2, mode=_lib.MX_FP64, hbm_budget=budget)
    torch.cuda.synchronize()
    before = torch.cuda.memory_allocated()
    with pytest.raises(_lib.MxError, match='budget'):
        srv.run([req_big], realtime=False)
    torch.cuda.synchronize()
    assert torch.cuda.memory_allocated() == before and srv.used == 0 and not srv.slots
    res = srv.run([req_small], realtime=False)          # and the server still serves
    srv.close()
    assert res[0].stats['status'] == 'converged' and srv.used == 0


@pytest.mark.parametrize('adjoint', [False, True])
def test_source_inside_pec_is_flagged(adjoint):
    from meanas_amd.fdfd import solvers
    omega, dxes, J, eps, pec = pec_problem()
    E = solvers.generic(omega, dxes, J, eps, pec=pec, adjoint=adjoint, matrix_solver_opts=dict(rtol=1e-11))
    info = solvers.last_info
    assert info['status'] == 'inconsistent' and info['solved_status'] == 'converged', info
    # the least-squares solution: the oracle's dense lstsq (zero on PEC entries)
    E_dense = ref.dense_solve(omega, dxes, J, eps, pec=pec, adjoint=adjoint)
    assert numpy.linalg.norm(E - E_dense) <= 1e-8 * numpy.linalg.norm(E_dense)
    assert not E[pec != 0].any()
    # the reported residual is that of the returned field, floor included
    A0 = ref.e_full(omega, dxes, eps, pec=pec)
    b0 = -1j * omega * J
    p = ref.e_full_preconditioner_vector(dxes)
    if adjoint:
        Pl, Pr = ref.e_full_preconditioners(dxes)
        A = (Pl @ A0 @ Pr).conj().T
        b = Pr.conj().T @ b0
        z = E / numpy.conj(p)                        # E = Pl^H z (fdfd/solvers.py:136-137)
        res = numpy.linalg.norm(A @ z - b) / numpy.linalg.norm(b)
        floor = numpy.linalg.norm(b[pec != 0]) / numpy.linalg.norm(b)
    else:
        res = numpy.linalg.norm(p * (A0 @ E - b0)) / numpy.linalg.norm(p * b0)
        floor = numpy.linalg.norm((p * b0)[pec != 0]) / numpy.linalg.norm(p * b0)
    assert abs(info['rel_residual'] - res) <= 1e-9 and abs(info['residual_floor'] - floor) <= 1e-12, (info, res, floor)


def test_server_flags_source_inside_pec():
    from meanas_amd import _lib
    from meanas_amd.fdfd import solvers
    from meanas_amd.serve import Request, Server
    omega, dxes, J, eps, pec = pec_problem(seed=51)
    srv = Server(capacity=2, mode=_lib.MX_FP64)
    res = srv.run([Request(omega=omega, dxes=dxes, J=J, epsilon=eps, pec=pec, rtol=1e-11, tag=0)], realtime=False)[0]
    srv.close()
    assert res.stats['status'] == 'inconsistent' and res.stats['residual_floor'] > 0
    E = solvers.generic(omega, dxes, J, eps, pec=pec, matrix_solver_opts=dict(rtol=1e-11))
    assert numpy.linalg.norm(res.E - E) <= 1e-9 * numpy.linalg.norm(E)
