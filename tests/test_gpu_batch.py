"""Batched and served solves on the GPU.

  * many right-hand sides on one structure (anisotropic eps) in one batch: every
    solution meets its residual target against the oracle's CSR operator and equals
    the single-system solve bit for bit;
  * an uneven batch of structures (different grid sizes, PEC blocks): every solution
    meets the physical residual target with the oracle's operator;
  * one finishing launch per fused iteration gives the same bits as two;
  * a solve plan kept from the previous call gives the same bits as a fresh one;
  * the continuous-batching server: each result is the same alone and in the mix
    (bit for bit), forward and adjoint requests, with admission while others iterate;
    a fused request that loses track continues in the classic form as mx_solve's does.
"""
import numpy
import pytest

from oracle import fdfd_ref as ref

pytestmark = pytest.mark.gpu


def small_aniso(shape=(24, 20, 22), nrhs=4, seed=7):
    from meanas_amd import workloads
    g = numpy.random.default_rng(seed)
    omega = 2 * numpy.pi / 12.0
    dxes = workloads.pml_dxes(shape, omega, thickness=4)
    eps = workloads.aniso_structure(shape, g, nbox=6).ravel()
    Js = [workloads.gaussian_sheet(shape, int(g.integers(6, shape[0] - 6)), 3.0,
                                   workloads.random_polarisation(g)).ravel() for _ in range(nrhs)]
    return omega, dxes, eps, Js


def physical_residual(omega, dxes, eps, J, E, pec=None):
    A0 = ref.e_full(omega, dxes, eps, pec=pec)
    b0 = -1j * omega * J
    return numpy.linalg.norm(A0 @ E - b0) / numpy.linalg.norm(b0)


@pytest.mark.parametrize('method', ['fp64', 'mixed'])
def test_multi_rhs_batch(method):
    import torch
    from meanas_amd import _lib
    from meanas_amd.fdfd.device import DeviceOperator, solve_batch
    mode = _lib.MX_FP64 if method == 'fp64' else _lib.MX_MIXED
    omega, dxes, eps, Js = small_aniso()
    op = DeviceOperator(omega, dxes, eps)
    bs = [op.map(_lib.MX_MAP_RHS, torch.from_numpy(J).cuda(), coef=-1j * omega) for J in Js]
    xs = [torch.zeros_like(b) for b in bs]
    st = solve_batch([op], [dict(op=0, b=b, x=x, rtol=1e-9) for b, x in zip(bs, xs)], mode=mode)
    for k, (J, x) in enumerate(zip(Js, xs)):
        assert st[k]['status'] == 'converged', st[k]
        if mode == _lib.MX_MIXED:   # finished by the complex64 iterations themselves
            assert st[k]['fp64_fallback'] is False, st[k]
        E = op.map(_lib.MX_MAP_E, x).cpu().numpy()
        assert physical_residual(omega, dxes, eps, J, E) < 1e-8
    # the first system alone gives the same bits
    x1 = torch.zeros_like(bs[0])
    solve_batch([op], [dict(op=0, b=bs[0], x=x1, rtol=1e-9)], mode=mode)
    assert torch.equal(x1, xs[0])


@pytest.mark.parametrize('case', ['aniso_batch', 'near_static'])
def test_merged_finishing_equals_separate(case, monkeypatch):
    """Fused batches finish the update and the next apply in one launch; separate
    launches (MEANAS_AMD_MERGED_FIN=0) must give the same bits and statistics, with
    even and odd chunk lengths."""
    import torch
    from meanas_amd import _lib
    from meanas_amd.fdfd.device import DeviceOperator, solve_batch
    if case == 'aniso_batch':
        omega, dxes, eps, Js = small_aniso(nrhs=3)
        rtol = 1e-9
    else:   # the reference's near-static fixture (meanas/test/test_fdfd.py): fused
        shape = (5, 5, 5)      # systems lose track and finish for the classic re-solve
        omega = 1 / 1500
        eps = numpy.ones((3, *shape))
        eps[:, 2, 2, 2] = 2.5
        dxes = [[numpy.full(5, 1.5) for _ in range(3)] for _ in range(2)]
        J = numpy.zeros((3, *shape), dtype=complex)
        J[:, 3, 3, 3] = 1 + 1j
        J[:, 1, 1, 1] = 1 - 1j
        eps, Js, rtol = eps.ravel(), [J.ravel()], 1e-11
    op = DeviceOperator(omega, dxes, eps)
    bs = [op.map(_lib.MX_MAP_RHS, torch.from_numpy(J).cuda(), coef=-1j * omega) for J in Js]
    runs = []
    for merged, chunk in (('1', 16), ('0', 16), ('1', 15), ('0', 15)):
        monkeypatch.setenv('MEANAS_AMD_MERGED_FIN', merged)
        xs = [torch.zeros_like(b) for b in bs]
        st = solve_batch([op], [dict(op=0, b=b, x=x, rtol=rtol) for b, x in zip(bs, xs)], mode=_lib.MX_FP64,
                         options=dict(form='fused', chunk=chunk, check_every=chunk))
        runs.append((xs, st))
    for (xm, sm), (xs_, ss) in (runs[:2], runs[2:]):
        for a, b in zip(xm, xs_):
            assert torch.equal(a, b)
        assert sm == ss
    if case == 'aniso_batch':   # the near-static system reaches the iteration cap, identically
        assert all(s['status'] == 'converged' for s in sm), sm


def test_kept_plans_equal_fresh_plans(monkeypatch):
    """mx_solve keeps the last call's tables, maps and captured graph per thread and reuses
    them when the next call has the same inputs.  Consecutive calls on the same buffers --
    changed stopping rules, a batch whose maps are compacted as its systems finish, and a
    fused system continued by a nested classic solve -- must give the same bits and
    statistics as calls that build every plan afresh, and must reuse the kept plans."""
    import torch
    from meanas_amd import _lib
    from meanas_amd.fdfd.device import DeviceOperator, solve_batch
    omega, dxes, eps, Js = small_aniso(nrhs=3)
    op = DeviceOperator(omega, dxes, eps)
    bs = [op.map(_lib.MX_MAP_RHS, torch.from_numpy(J).cuda(), coef=-1j * omega) for J in Js]
    xs = [torch.zeros_like(b) for b in bs]
    works = [torch.empty(op.work_bytes(_lib.MX_FP64), dtype=torch.uint8, device=b.device) for b in bs]
    # the reference's near-static fixture: its fused solve loses track and finishes in a
    # nested classic-form solve
    shape = (5, 5, 5)
    eps_s = numpy.ones((3, *shape))
    eps_s[:, 2, 2, 2] = 2.5
    dxes_s = [[numpy.full(5, 1.5) for _ in range(3)] for _ in range(2)]
    J_s = numpy.zeros((3, *shape), dtype=complex)
    J_s[:, 3, 3, 3] = 1 + 1j
    J_s[:, 1, 1, 1] = 1 - 1j
    op_s = DeviceOperator(1 / 1500, dxes_s, eps_s.ravel())
    b_s = op_s.map(_lib.MX_MAP_RHS, torch.from_numpy(J_s.ravel()).cuda(), coef=-1j / 1500)
    x_s = torch.zeros_like(b_s)
    work_s = torch.empty(op_s.work_bytes(_lib.MX_FP64), dtype=torch.uint8, device=b_s.device)

    def calls():
        out = []
        for rtols in ((1e-4, 1e-7, 1e-9), (1e-4, 1e-7, 1e-9), (1e-9, 1e-9, 1e-9), (1e-6, 1e-8, 1e-5)):
            for x in xs:
                x.zero_()
            st = solve_batch([op], [dict(op=0, b=b, x=x, work=w, rtol=r) for b, x, w, r in zip(bs, xs, works, rtols)],
                             mode=_lib.MX_FP64, options=dict(form='fused'))
            out.append(([x.clone() for x in xs], st))
        for rtol in (1e-11, 1e-11):
            x_s.zero_()
            st = solve_batch([op_s], [dict(op=0, b=b_s, x=x_s, work=work_s, rtol=rtol)], mode=_lib.MX_FP64,
                             options=dict(form='fused'))
            out.append(([x_s.clone()], st))
        return out

    monkeypatch.setenv('MEANAS_AMD_PLAN_CACHE', '0')
    fresh = calls()
    monkeypatch.setenv('MEANAS_AMD_PLAN_CACHE', '1')
    before = _lib.lib().mx_solve_plan_reuses()
    kept = calls()
    reuses = _lib.lib().mx_solve_plan_reuses() - before
    # the batch: calls 2-4 reuse the plan of call 1 (call 1 compacts it); the
    # near-static system: its second call reuses the first call's plan
    assert reuses >= 4, reuses
    # a released plan is not reused
    _lib.check(_lib.lib().mx_solve_plan_release(), 'mx_solve_plan_release')
    before = _lib.lib().mx_solve_plan_reuses()
    x_s.zero_()
    solve_batch([op_s], [dict(op=0, b=b_s, x=x_s, work=work_s, rtol=1e-11)], mode=_lib.MX_FP64, options=dict(form='fused'))
    assert _lib.lib().mx_solve_plan_reuses() == before
    assert any(s['iterations'] != fresh[0][1][0]['iterations'] for s in fresh[0][1])   # uneven finish
    for (xa, sa), (xb, sb) in zip(kept, fresh):
        for a, b in zip(xa, xb):
            assert torch.equal(a, b)
        assert sa == sb


@pytest.mark.parametrize('rnd', range(3))   # the threads interleave differently each time
def test_kept_plans_are_per_thread(rnd):
    """Two host threads solving different systems at the same time, twice each (the
    second call reuses each thread's kept plan), give the bits of the same solves run one
    after the other.  Each thread queues its own torch work on a stream of its own: an
    operation on the legacy stream while the other thread captures its chunk graph is
    refused by HIP, and invalidates that capture (the library itself issues none)."""
    import threading
    import torch
    from meanas_amd import _lib
    from meanas_amd.fdfd.device import DeviceOperator, solve_batch
    cases = []
    for shape, seed in (((24, 20, 22), 7), ((20, 26, 18), 8)):
        omega, dxes, eps, Js = small_aniso(shape=shape, nrhs=1, seed=seed)
        op = DeviceOperator(omega, dxes, eps)
        b = op.map(_lib.MX_MAP_RHS, torch.from_numpy(Js[0]).cuda(), coef=-1j * omega)
        work = torch.empty(op.work_bytes(_lib.MX_FP64), dtype=torch.uint8, device=b.device)
        cases.append((op, b, work))

    def run(case, x, out, key):
        op, b, work = case
        stream = torch.cuda.Stream()
        with torch.cuda.stream(stream):
            for rtol in (1e-6, 1e-9):
                x.zero_()
                st = solve_batch([op], [dict(op=0, b=b, x=x, work=work, rtol=rtol)], mode=_lib.MX_FP64,
                                 stream=stream)[0]
                out[(key, rtol)] = (x.clone(), st)
        stream.synchronize()

    done = threading.Barrier(2)

    def run_and_release(case, x, out, key):
        try:
            run(case, x, out, key)
        finally:
            # the thread ends: release its plan once neither thread captures any more (the
            # release frees device memory, which may synchronise the device)
            done.wait(timeout=120)
            _lib.check(_lib.lib().mx_solve_plan_release(), 'mx_solve_plan_release')

    seq = {}
    for k, case in enumerate(cases):
        run(case, torch.zeros_like(case[1]), seq, k)
    par = {}
    xs = [torch.zeros_like(case[1]) for case in cases]
    torch.cuda.synchronize()            # inputs and outputs ready before the threads start
    threads = [threading.Thread(target=run_and_release, args=(case, x, par, k))
               for k, (case, x) in enumerate(zip(cases, xs))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert set(par) == set(seq)
    for key in seq:
        assert torch.equal(par[key][0], seq[key][0]), key
        assert par[key][1] == seq[key][1], key


def test_uneven_structures_with_pec():
    import torch
    from meanas_amd import _lib, workloads
    from meanas_amd.fdfd.device import DeviceOperator, solve_batch
    probs = workloads.uneven_batch(count=4, seed=11, edges=(26, 30, 34))
    ops, systems = [], []
    for k, w in enumerate(probs):
        op = DeviceOperator(w['omega'], w['dxes'], w['epsilon'], pec=w['pec'])
        ops.append(op)
        b = op.map(_lib.MX_MAP_RHS, torch.from_numpy(w['J']).cuda(), coef=-1j * w['omega'])
        systems.append(dict(op=k, b=b, x=torch.zeros_like(b), rtol=1e-9))
    st = solve_batch(ops, systems, mode=_lib.MX_FP64)
    for k, w in enumerate(probs):
        assert st[k]['status'] == 'converged', st[k]
        E = ops[k].map(_lib.MX_MAP_E, systems[k]['x']).cpu().numpy()
        assert physical_residual(w['omega'], w['dxes'], w['epsilon'], w['J'], E, pec=w['pec']) < 1e-8
        # PEC entries of E stay zero (the source avoids PEC cells)
        assert numpy.abs(E[w['pec'] != 0]).max() == 0


def _requests():
    from meanas_amd import workloads
    from meanas_amd.serve import Request
    g = numpy.random.default_rng(5)
    out = []
    for k, shape in enumerate([(24, 20, 22), (20, 26, 18), (28, 18, 20), (22, 22, 22)]):
        omega = 2 * numpy.pi / 12.0
        dxes = workloads.pml_dxes(shape, omega, thickness=4)
        eps = workloads.aniso_structure(shape, g, nbox=5).ravel()
        J = workloads.gaussian_sheet(shape, 8, 3.0, workloads.random_polarisation(g)).ravel()
        out.append(Request(omega=omega, dxes=dxes, J=J, epsilon=eps, adjoint=(k % 2 == 1), rtol=1e-9,
                           arrival=0.0, tag=k))
    return out


@pytest.mark.parametrize('method', ['fp64', 'mixed'])
def test_server_alone_equals_mixed(method):
    from meanas_amd import _lib
    from meanas_amd.serve import Server
    mode = _lib.MX_FP64 if method == 'fp64' else _lib.MX_MIXED
    reqs = _requests()
    alone = {}
    for r in reqs:
        srv = Server(capacity=1, mode=mode, chunk=8)
        res = srv.run([r], realtime=False)
        srv.close()
        alone[r.tag] = res[0]
    srv = Server(capacity=3, mode=mode, chunk=8)        # fewer slots than requests: admission mid-run
    mixed = {res.tag: res for res in srv.run(reqs, realtime=False)}
    srv.close()
    assert set(mixed) == set(alone)
    for r in reqs:
        a, m = alone[r.tag], mixed[r.tag]
        assert m.stats['status'] == 'converged', m.stats
        assert numpy.array_equal(a.E, m.E), r.tag
        # and the solution is right: forward vs the oracle operator, adjoint vs the
        # conjugate-transposed system (fdfd/solvers.py:120-137)
        if r.adjoint:
            A0 = ref.e_full(r.omega, r.dxes, r.epsilon)
            Pl, Pr = ref.e_full_preconditioners(r.dxes)
            Aadj = (Pl @ A0 @ Pr).conj().T
            z = Pl.diagonal().conj() ** -1 * m.E          # E = Pl^H z
            bb = Pr.conj().T @ (-1j * r.omega * r.J)
            assert numpy.linalg.norm(Aadj @ z - bb) / numpy.linalg.norm(bb) < 1e-8
        else:
            assert physical_residual(r.omega, r.dxes, r.epsilon, r.J, m.E) < 1e-8


def test_session_steps_give_small_groups_equal_time():
    """A session step gives each launch group about the same time: the chunk of a group of
    small systems repeats (about cost_max / cost times, its cells plus a fixed launch cost
    of 4e5 cells).  A 16x14x12 request (real isotropic eps) beside a 96^3 one (per-component
    eps: another material kind, so another group) advances 3 chunks per step, the large one
    1, and the small request's field has the bits it has alone."""
    from meanas_amd import _lib, workloads
    from meanas_amd.serve import Request, Server
    g = numpy.random.default_rng(11)
    omega = 2 * numpy.pi / 12.0
    reqs = []
    for tag, shape, aniso in (('small', (16, 14, 12), False), ('large', (96, 96, 96), True)):
        dxes = workloads.pml_dxes(shape, omega, thickness=4)
        if aniso:
            eps = workloads.aniso_structure(shape, g, nbox=5).ravel()
        else:
            eps = numpy.tile(workloads.paint(shape, 1.0, workloads.random_boxes(g, shape, 3, 1.5, 4.0)).ravel(), 3)
        J = workloads.gaussian_sheet(shape, shape[0] // 2, 2.0, workloads.random_polarisation(g)).ravel()
        reqs.append(Request(omega=omega, dxes=dxes, J=J, epsilon=eps, rtol=1e-9, arrival=0.0, tag=tag))
    srv = Server(capacity=2, mode=_lib.MX_FP64, chunk=8)
    idx = {r.tag: srv.admit(srv.prepare(r), 0.0) for r in reqs}
    srv.step()
    stats, _ = srv.poll(srv.lane)
    srv.close()
    assert stats[idx['large']].iterations == 8
    assert stats[idx['small']].iterations == 8 * 3
    alone = Server(capacity=1, mode=_lib.MX_FP64, chunk=8)
    a = alone.run([reqs[0]], realtime=False)[0]
    alone.close()
    srv = Server(capacity=2, mode=_lib.MX_FP64, chunk=8)
    both = {res.tag: res for res in srv.run(reqs, realtime=False)}
    srv.close()
    assert both['small'].stats['status'] == 'converged' and both['large'].stats['status'] == 'converged'
    assert both['small'].stats['iterations'] == a.stats['iterations']
    assert numpy.array_equal(both['small'].E, a.E)
    assert physical_residual(reqs[0].omega, reqs[0].dxes, reqs[0].epsilon, reqs[0].J, a.E) < 1e-8


def _near_static_request(tag, adjoint=False):
    """The reference's near-static fixture (meanas/test/test_fdfd.py, omega = 1/1500): its
    fused FP64 recursion loses track and the solve continues in the classic form."""
    from meanas_amd.serve import Request
    shape = (5, 5, 5)
    eps = numpy.ones((3, *shape))
    eps[:, 2, 2, 2] = 2.5
    dxes = [[numpy.full(5, 1.5) for _ in range(3)] for _ in range(2)]
    J = numpy.zeros((3, *shape), dtype=complex)
    J[:, 3, 3, 3] = 1 + 1j
    J[:, 1, 1, 1] = 1 - 1j
    return Request(omega=1 / 1500, dxes=dxes, J=J.ravel(), epsilon=eps.ravel(), adjoint=adjoint, rtol=1e-11,
                   maxiter=3000, arrival=0.0, tag=tag)


def test_server_continues_switched_systems_in_classic_form():
    """FP64 sessions run the fused form; a request whose fused recursion loses track
    finishes there (MX_FLAG_SWITCHED) and continues in a classic-form session.  Its result
    equals mx_solve's (which continues such systems itself) bit for bit, with the same
    statistics, alone and in a mix with other requests."""
    import torch
    from meanas_amd import _lib
    from meanas_amd.fdfd.device import DeviceOperator, solve_batch
    from meanas_amd.serve import Server
    r = _near_static_request('ns')
    op = DeviceOperator(r.omega, r.dxes, r.epsilon)
    b = op.map(_lib.MX_MAP_RHS, torch.from_numpy(r.J).cuda(), coef=-1j * r.omega)
    x = torch.zeros_like(b)
    ref_st = solve_batch([op], [dict(op=0, b=b, x=x, rtol=r.rtol, maxiter=r.maxiter)], mode=_lib.MX_FP64,
                         options=dict(chunk=16))[0]
    E_ref = op.map(_lib.MX_MAP_E, x).cpu().numpy()
    srv = Server(capacity=1, mode=_lib.MX_FP64, chunk=16)
    alone = srv.run([r], realtime=False)[0]
    srv.close()
    assert 0 < alone.stats.get('fused_iterations', 0) < alone.stats['iterations'], alone.stats
    assert numpy.array_equal(alone.E, E_ref)
    for k in ('status', 'iterations', 'replacements', 'rel_residual'):
        assert alone.stats[k] == ref_st[k], (k, alone.stats, ref_st)
    srv = Server(capacity=2, mode=_lib.MX_FP64, chunk=16)
    mixed = {res.tag: res for res in srv.run([r] + _requests()[:3], realtime=False)}
    srv.close()
    assert numpy.array_equal(mixed['ns'].E, alone.E)
    assert mixed['ns'].stats == alone.stats


def test_server_matches_generic():
    from meanas_amd import _lib
    from meanas_amd.fdfd import solvers
    from meanas_amd.serve import Server
    r = _requests()[0]
    srv = Server(capacity=2, mode=_lib.MX_FP64, chunk=16)
    res = srv.run([r], realtime=False)[0]
    srv.close()
    E = solvers.generic(r.omega, r.dxes, r.J, r.epsilon, matrix_solver='cocg_fp64',
                        matrix_solver_opts=dict(rtol=1e-9))
    assert numpy.linalg.norm(res.E - E) <= 1e-7 * numpy.linalg.norm(E)
