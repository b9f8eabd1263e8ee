#!/usr/bin/env python3
"""Benchmark of the 3D FDFD solve path on one MI355X.

Metric: solves/s to 1e-8 relative residual on the 128^3 workload (eps = 2.085 + 40
random dielectric boxes, 10-cell SC-PML on all faces, Gaussian Jy sheet at x = 20),
plus p50/p99 solve latency and the operator-apply HBM bandwidth.

A step is one complete solve (FP64 COCG by default, `--mode mixed` for complex64 COCG
with FP64 residual replacement), stopped when the FP64 true residual ||b - A x|| / ||b||
<= 1e-8, with the operator, materials and right-hand side already resident in HBM.
Beside it: the same solve in the other precision mode (same GPU), the field error
against an FP64 solve to 1e-11, each kernel of one iteration timed with HIP events on
its launch stream (the roofline of the dominant one, and of the operator apply), and
the oracle (restated reference: CSR assembly + scipy qmr, complex128) timed on this
host over a bounded sample.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# scipy qmr iterations the oracle needs on the 128^3 workload to reach rtol 1e-8,
# measured by running the oracle to convergence on a CPU host (scipy 1.15.3: 6447
# iterations, true relative residual 9.7e-9, 7162 s plus 21 s of CSR assembly on one
# core of a 2.7 GHz Xeon)
QMR_ITERS_128 = 6447
# HBM bytes per launch (MB) from rocprofv3 FETCH_SIZE / WRITE_SIZE counters, scaled by the
# factors a 1 GiB copy gives (profiles/*_hbm_traffic.md), keyed by kernel, mode, shape
TRAFFIC_MB = {('update', 'fp64', (128, 128, 128)): 604.4,   # profiles/iteration_kernels_boxes128_hbm_traffic.md
              ('apply', 'fp64', (128, 128, 128)): 231.4}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=5)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--size', type=int, default=128)
    ap.add_argument('--rtol', type=float, default=1e-8)
    ap.add_argument('--apply-reps', type=int, default=200)
    ap.add_argument('--no-fp64', action='store_true')
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--cpu-iters', type=int, default=40)
    ap.add_argument('--apply-sweep', action='store_true',
                    help='only time the operator apply over tile heights and slab depths')
    ap.add_argument('--iter-sweep', action='store_true',
                    help='only time the kernels of the FP64 iteration, in solve order, over tile heights and '
                         'slab depths')
    ap.add_argument('--apply-only', action='store_true',
                    help='only run the operator apply (complex64 and complex128, default tiling) and the '
                         'kernels of one FP64 solver iteration: the program the hardware-counter profiles '
                         'are taken on')
    ap.add_argument('--workload', choices=['single', 'multi_rhs', 'uneven', 'serve', 'fdtd'], default='single',
                    help='single: the 128^3 solve (the metric); multi_rhs: 16 right-hand sides on one '
                         '128x96x96 anisotropic structure; uneven: 32 structures of uneven size with PEC; '
                         'serve: Poisson request stream with continuous batching; fdtd: 256^3 Yee leapfrog '
                         'time stepping (float64 and float32)')
    ap.add_argument('--count', type=int, default=0, help='structures (uneven) or requests (serve)')
    ap.add_argument('--qps', type=float, default=4.0, help='serve: arrival rate')
    ap.add_argument('--form', choices=['auto', 'fused', 'classic'], default='auto',
                    help='COCG iteration form (auto: fused for fp64, classic for mixed)')
    ap.add_argument('--no-configs', action='store_true',
                    help='skip the other benchmark workloads (configs 3-5) and the precision study')
    ap.add_argument('--mode', choices=['mixed', 'fp64'], default='fp64',
                    help='solver of the timed step (the other one is the same-GPU comparison)')
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(w, iters):
    """Oracle on this host: CSR assembly of the scaled operator + `iters` qmr iterations."""
    import scipy.sparse.linalg as spla
    from oracle import fdfd_ref as ref
    t0 = time.perf_counter()
    A0 = ref.e_full(w['omega'], w['dxes'], w['epsilon'])
    Pl, Pr = ref.e_full_preconditioners(w['dxes'])
    A = (Pl @ A0 @ Pr).tocsr()
    b = Pl @ (-1j * w['omega'] * w['J'])
    t1 = time.perf_counter()
    spla.qmr(A, b, rtol=1e-30, atol=0.0, maxiter=iters)
    t2 = time.perf_counter()
    per_iter = (t2 - t1) / iters
    out = dict(assembly_s=t1 - t0, qmr_s_per_iter=per_iter, nnz=int(A.nnz))
    if QMR_ITERS_128 is not None and w['shape'] == (128, 128, 128):
        proj = (t1 - t0) + per_iter * QMR_ITERS_128
        out.update(projected_solve_s=proj, value=1.0 / proj, qmr_iterations_to_1e8=QMR_ITERS_128)
    return out


def time_apply(plan, x, y, reps, stream):
    import torch
    for _ in range(5):
        plan.apply(x, y, stream=stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        plan.apply(x, y, stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) * 1e3 / reps


def apply_sweep(op, dev):
    """Apply time over tile heights x slab depths, complex64 and complex128."""
    import torch
    stream = torch.cuda.current_stream()
    for mode, dt, elem in (('c64', torch.complex64, 8), ('c128', torch.complex128, 16)):
        x = torch.randn(op.n3, dtype=dt, device=dev)
        y = torch.empty_like(x)
        bpc = op.apply_bytes_per_cell(elem)
        for ty in (2, 4, 8):
            for slab in (1, 2, 4, 8, 16, 32, 64):
                plan = op.plan(mode, tile_rows=ty, slab=slab)
                us = time_apply(plan, x, y, 50, stream)
                print(json.dumps(dict(mode=mode, tile_rows=ty, slab=slab, blocks=plan.blocks, us=round(us, 2),
                                      gbs=round(bpc * op.ncell / (us * 1e-6) / 1e9, 1))), flush=True)
        torch.cuda.synchronize()
    # plain copy of the same bytes for reference (HBM roofline as the hardware delivers it)
    nbytes = op.apply_bytes_per_cell(8) * op.ncell // 2
    a = torch.empty(nbytes // 4, dtype=torch.float32, device=dev).fill_(1.0)
    c = torch.empty_like(a)
    for _ in range(5):
        c.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(50):
        c.copy_(a)
    e1.record(stream)
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) * 1e3 / 50
    print(json.dumps(dict(copy_bytes=2 * nbytes, us=round(us, 2), gbs=round(2 * nbytes / (us * 1e-6) / 1e9, 1))))


def fdtd_workload(args):
    """Yee leapfrog steps on a 256^3 grid (larger than the 256 MB Infinity Cache, so the
    rates are HBM rates): eps a full field, mu = 1, stretched widths."""
    import torch
    from meanas_amd import fdtd
    n = args.size if args.size != 128 else 256
    shape = (n, n, n)
    rng = numpy.random.default_rng(9)
    dxes = [[0.9 + 0.2 * rng.random(k) for k in shape] for _ in range(2)]
    stream = torch.cuda.current_stream()
    out = {}
    for name, dt_ in (('f64', torch.float64), ('f32', torch.float32)):
        elem = 8 if dt_ == torch.float64 else 4
        e = torch.zeros((3, *shape), dtype=dt_, device='cuda')
        h = torch.zeros_like(e)
        eps = (1.0 + torch.rand((3, *shape), dtype=dt_, device='cuda'))
        e[1, n // 2, n // 2, n // 2] = 1.0
        me, mh = fdtd.maxwell_e(0.3, dxes), fdtd.maxwell_h(0.3, dxes)
        for _ in range(3):
            mh(e, h)
            me(e, h, eps)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        reps = 20
        ev[0].record(stream)
        for _ in range(reps):
            mh(e, h)
        ev[1].record(stream)
        for _ in range(reps):
            me(e, h, eps)
        ev[2].record(stream)
        fdtd.run(e, h, eps, 0.3, dxes, steps=args.steps * 10)
        ev[3].record(stream)
        torch.cuda.synchronize()
        us_h = ev[0].elapsed_time(ev[1]) * 1e3 / reps
        us_e = ev[1].elapsed_time(ev[2]) * 1e3 / reps
        us_step = ev[2].elapsed_time(ev[3]) * 1e3 / (args.steps * 10)
        cells = n ** 3
        b_h = 3 * elem * 3 * cells                 # e read, h read, h written
        b_e = 4 * 3 * elem * cells                 # h, e, eps read, e written
        # the same step with 10-cell CPML layers on all six faces (psi in float32)
        layers = [[fdtd.cpml_params(a, p, 0.3, thickness=10) for p in (-1, 1)] for a in range(3)]
        uE, uH = fdtd.updates_with_cpml(layers, 0.3, dxes, eps)
        for _ in range(2):
            uH(e, h)
            uE(e, h, eps)
        cev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        csteps = args.steps * 5
        cev[0].record(stream)
        for _ in range(csteps):
            uH(e, h)
            uE(e, h, eps)
        cev[1].record(stream)
        torch.cuda.synchronize()
        us_cpml = cev[0].elapsed_time(cev[1]) * 1e3 / csteps
        # analysis calls on the same fields: the energy at an H step (e0, h1, e2 and eps
        # read, N float64 written), the Poynting vector (e, h read, s written) and one
        # conducting face (E and H); per call, including the upload of the widths
        e2 = e.clone()
        en_, hn_ = fdtd.conducting_boundary(0, -1)
        fdtd.energy_hstep(e, h, e2, epsilon=eps, dxes=dxes)
        fdtd.poynting(e, h, dxes)
        aev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        areps = 10
        aev[0].record(stream)
        for _ in range(areps):
            fdtd.energy_hstep(e, h, e2, epsilon=eps, dxes=dxes)
        aev[1].record(stream)
        for _ in range(areps):
            fdtd.poynting(e, h, dxes)
        aev[2].record(stream)
        for _ in range(areps):
            en_(e)
            hn_(h)
        aev[3].record(stream)
        torch.cuda.synchronize()
        us_en = aev[0].elapsed_time(aev[1]) * 1e3 / areps
        us_s = aev[1].elapsed_time(aev[2]) * 1e3 / areps
        us_cb = aev[2].elapsed_time(aev[3]) * 1e3 / areps
        b_en = 12 * elem * cells + 8 * cells
        b_s = 9 * elem * cells
        del e2
        out[name] = dict(step_us=round(us_step, 2), mcell_updates_per_s=round(cells / us_step, 1),
                         h_update=dict(us=round(us_h, 2), bytes=b_h, gbs=round(b_h / us_h / 1e3, 1)),
                         e_update=dict(us=round(us_e, 2), bytes=b_e, gbs=round(b_e / us_e / 1e3, 1)),
                         cpml_step=dict(us=round(us_cpml, 2), mcell_updates_per_s=round(cells / us_cpml, 1),
                                        layers='10 cells on all six faces, psi float32'),
                         energy_hstep=dict(us=round(us_en, 2), bytes=b_en, gbs=round(b_en / us_en / 1e3, 1)),
                         poynting=dict(us=round(us_s, 2), bytes=b_s, gbs=round(b_s / us_s / 1e3, 1)),
                         conducting_face_e_and_h=dict(us=round(us_cb, 2)))
        log(f'fdtd {name}: {out[name]}')
        del uE, uH
        del e, h, eps
    line = dict(metric='FDTD cell updates/s (Yee leapfrog step: H then E update)', value=out['f64']['mcell_updates_per_s'] * 1e6,
                unit='cell-updates/s', n_gpus=1, dtype='f64 (f32 beside it)', data='synthetic',
                config={'workload': f'fdtd: {n}^3 grid, eps field, mu = 1, stretched widths'}, **out)
    print(json.dumps(line), flush=True)


def batch_workloads(args):
    """Multi-RHS, uneven multi-structure and serving workloads (one JSON line each)."""
    import torch
    from meanas_amd import _lib, workloads
    from meanas_amd.fdfd.device import DeviceOperator, solve_batch, time_kernels
    dev = torch.device('cuda:0')
    mode = _lib.MX_FP64 if args.mode == 'fp64' else _lib.MX_MIXED
    fopt = dict(form=args.form if (mode == _lib.MX_FP64 or args.form != 'fused') else 'auto')
    if args.workload == 'fdtd':
        fdtd_workload(args)
        return
    if args.workload == 'multi_rhs':
        w = workloads.aniso_multi()
        op = DeviceOperator(w['omega'], w['dxes'], w['epsilon'], device=dev)
        bs = [op.map(_lib.MX_MAP_RHS, torch.from_numpy(J).to(dev), coef=-1j * w['omega']) for J in w['Js']]
        xs = [torch.zeros_like(b) for b in bs]
        systems = [dict(op=0, b=b, x=x, rtol=args.rtol, maxiter=200000) for b, x in zip(bs, xs)]
        solve_batch([op], systems[:1], mode=mode, options=fopt)   # warm-up (autotune, code load)
        for x in xs:
            x.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = solve_batch([op], systems, mode=mode, options=fopt)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        line = dict(metric='solves/s, 16 right-hand sides on one structure', value=len(xs) / wall, unit='solves/s',
                    wall_s=wall, converged=sum(s['status'] == 'converged' for s in st),
                    iterations=[s['iterations'] for s in st],
                    config={'workload': 'aniso_multi: 128x96x96, anisotropic box eps, 16 Gaussian sheets', 'rtol': args.rtol,
                            'mode': args.mode, 'form': args.form})
    elif args.workload == 'uneven':
        probs = workloads.uneven_batch(count=args.count or 32)
        ops, systems = [], []
        for k, w in enumerate(probs):
            op = DeviceOperator(w['omega'], w['dxes'], w['epsilon'], pec=w['pec'], device=dev)
            ops.append(op)
            b = op.map(_lib.MX_MAP_RHS, torch.from_numpy(w['J']).to(dev), coef=-1j * w['omega'])
            systems.append(dict(op=k, b=b, x=torch.zeros_like(b), rtol=args.rtol, maxiter=200000))
        for op in ops:                                       # autotune outside the timed region
            from meanas_amd.fdfd.autotune import tuned
            tuned(op, mode)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = solve_batch(ops, systems, mode=mode, options=fopt)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        cells = sum(op.ncell for op in ops)
        line = dict(metric='solves/s, uneven batch of structures in one launch sequence', value=len(ops) / wall,
                    unit='solves/s', wall_s=wall, converged=sum(s['status'] == 'converged' for s in st),
                    total_cells=cells, iterations_max=max(s['iterations'] for s in st),
                    systems=[dict(shape=op.shape, omega=round(w['omega'], 5), status=x['status'], it=x['iterations'],
                                  res=x['rel_residual']) for op, w, x in zip(ops, probs, st)],
                    config={'workload': f'uneven_batch: {len(ops)} structures, edges 48-128, 5% PEC', 'rtol': args.rtol,
                            'mode': args.mode, 'form': args.form})
    else:
        from meanas_amd.serve import Request, Server, summarize
        reqs = []
        for k, r in enumerate(workloads.serving_stream(args.count or 40, args.qps)):
            reqs.append(Request(omega=r['omega'], dxes=r['dxes'], J=r['J'], epsilon=r['epsilon'], adjoint=r['adjoint'],
                                rtol=args.rtol, arrival=r['arrival'], tag=dict(k=k, adjoint=r['adjoint'], shape=r['shape'])))
        srv = Server(capacity=16, mode=mode, chunk=16, options=fopt)
        # warm the autotuner cache for every shape outside the timed run
        from meanas_amd.fdfd.autotune import tuned
        for r in reqs:
            tuned(DeviceOperator(r.omega, r.dxes, r.epsilon, device=dev), mode)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = srv.run(reqs, realtime=True)
        wall = time.perf_counter() - t0
        srv.close()
        summ = summarize(res, wall)
        line = dict(metric='served solves/s with p50/p99 latency (continuous batching)', value=summ['solves_per_s'],
                    unit='solves/s', latency_ms={'p50': summ['p50_ms'], 'p99': summ['p99_ms']}, **summ,
                    config={'workload': f'serving: {len(reqs)} Poisson requests at {args.qps} QPS, edges lognormal '
                                        'in [48,256], 30% adjoint, anisotropic box eps', 'rtol': args.rtol,
                            'mode': args.mode, 'form': args.form, 'capacity': 16})
    line.update(n_gpus=1, dtype='c128' if args.mode == 'fp64' else 'c64 inner + f64 residual', data='synthetic')
    print(json.dumps(line), flush=True)


# ---- the other benchmark workloads at full size (extra keys of the default line) -------
def config3_multi_rhs(args, dev):
    """16 right-hand sides on the 128x96x96 anisotropic structure in one batch, against
    the same structure's right-hand sides solved one at a time (first two)."""
    import torch
    from meanas_amd import _lib, workloads
    from meanas_amd.fdfd.device import DeviceOperator, solve_batch
    w = workloads.aniso_multi()
    op = DeviceOperator(w['omega'], w['dxes'], w['epsilon'], device=dev)
    bs = [op.map(_lib.MX_MAP_RHS, torch.from_numpy(J).to(dev), coef=-1j * w['omega']) for J in w['Js']]
    xs = [torch.zeros_like(b) for b in bs]
    systems = [dict(op=0, b=b, x=x, rtol=args.rtol, maxiter=200000) for b, x in zip(bs, xs)]
    solve_batch([op], [dict(systems[0], x=torch.zeros_like(bs[0]))], mode=_lib.MX_FP64)   # warm-up, autotune
    singles = []
    for k in range(2):
        xk = torch.zeros_like(bs[k])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = solve_batch([op], [dict(systems[k], x=xk)], mode=_lib.MX_FP64)[0]
        torch.cuda.synchronize()
        singles.append((time.perf_counter() - t0, st['iterations']))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = solve_batch([op], systems, mode=_lib.MX_FP64)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    its = [s_['iterations'] for s_ in st]
    cell_its = op.ncell * sum(its)
    single_rate = op.ncell * sum(i for _, i in singles) / sum(t for t, _ in singles)
    out = dict(workload='aniso_multi: 128x96x96, anisotropic box eps, 16 Gaussian sheets, FP64, rtol %g' % args.rtol,
               solves_per_s=len(bs) / wall, wall_s=wall, converged=sum(s_['status'] == 'converged' for s_ in st),
               max_rel_residual=max(s_['rel_residual'] for s_ in st), iterations=its,
               cell_iterations_per_s=cell_its / wall, single_solve_cell_iterations_per_s=single_rate,
               batch_over_single_rate=cell_its / wall / single_rate,
               single_solves=[dict(s=round(t, 3), iterations=i) for t, i in singles])
    log(f'config 3: {out}')
    return out


def config4_uneven(args, dev, form):
    import torch
    from meanas_amd import _lib, workloads
    from meanas_amd.fdfd.autotune import tuned
    from meanas_amd.fdfd.device import DeviceOperator, solve_batch
    probs = workloads.uneven_batch(count=32)
    ops, systems = [], []
    for k, w in enumerate(probs):
        op = DeviceOperator(w['omega'], w['dxes'], w['epsilon'], pec=w['pec'], device=dev)
        ops.append(op)
        b = op.map(_lib.MX_MAP_RHS, torch.from_numpy(w['J']).to(dev), coef=-1j * w['omega'])
        systems.append(dict(op=k, b=b, x=torch.zeros_like(b), rtol=args.rtol, maxiter=400000))
        tuned(op, _lib.MX_FP64)                       # autotune outside the timed region
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = solve_batch(ops, systems, mode=_lib.MX_FP64, options=dict(form=form))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    its = [s_['iterations'] for s_ in st]
    k = int(numpy.argmax(its))
    cell_its = sum(op.ncell * i for op, i in zip(ops, its))
    out = dict(form=form, solves_per_s=len(ops) / wall, wall_s=wall,
               converged=sum(s_['status'] == 'converged' for s_ in st),
               max_rel_residual=max(s_['rel_residual'] for s_ in st), cell_iterations_per_s=cell_its / wall,
               straggler=dict(index=k, shape=ops[k].shape, iterations=its[k]),
               median_iterations=float(numpy.median(its)), iterations=its)
    log(f'config 4 ({form}): {out}')
    del ops, systems
    return out


def config5_serving(args, dev, qps, count):
    from meanas_amd import _lib, workloads
    from meanas_amd.fdfd.autotune import tuned
    from meanas_amd.fdfd.device import DeviceOperator
    from meanas_amd.serve import Request, Server, summarize
    reqs = []
    for k, r in enumerate(workloads.serving_stream(count, qps)):
        reqs.append(Request(omega=r['omega'], dxes=r['dxes'], J=r['J'], epsilon=r['epsilon'], adjoint=r['adjoint'],
                            rtol=args.rtol, arrival=r['arrival'], tag=dict(k=k, adjoint=r['adjoint'], shape=r['shape'])))
        tuned(DeviceOperator(r['omega'], r['dxes'], r['epsilon'], device=dev), _lib.MX_FP64)
    srv = Server(capacity=16, mode=_lib.MX_FP64, chunk=16)
    t0 = time.perf_counter()
    res = srv.run(reqs, realtime=True)
    wall = time.perf_counter() - t0
    srv.close()
    summ = summarize(res, wall)
    out = dict(qps=qps, requests=count, arrivals_s=reqs[-1].arrival, latency_ms={'p50': summ['p50_ms'], 'p99': summ['p99_ms']},
               solves_per_s=summ['solves_per_s'], wall_s=wall, converged=summ['converged'],
               max_rel_residual=max(r.stats['rel_residual'] for r in res), adjoint=summ['adjoint'])
    log(f'config 5 ({qps} QPS): {out}')
    return out


def beyond_cache(args, dev, n=256):
    """The iteration kernels at n^3 (256^3: each complex128 vector 805 MB, far beyond the
    256 MB Infinity Cache), timed in solve order with HIP events: HBM rates without the
    cache's help, on the largest grid of the serving stream."""
    import torch
    from meanas_amd import _lib, workloads
    from meanas_amd.fdfd.device import DeviceOperator, time_kernels
    w = workloads.boxes_128(n)
    op = DeviceOperator(w['omega'], w['dxes'], w['epsilon'], device=dev)
    b = op.map(_lib.MX_MAP_RHS, torch.from_numpy(w['J']).to(dev), coef=-1j * w['omega'])
    kt = time_kernels(op, b, _lib.MX_FP64, reps=20)
    a_bytes = op.apply_bytes_per_cell(16) * op.ncell
    u_bytes = 6 * 16 * op.n3
    out = dict(workload=f'boxes: {n}^3, eps = 2.085 + boxes, FP64 iteration kernels', tiling=[kt['tile_rows'], kt['slab']],
               apply=dict(us=kt['apply'], bytes=a_bytes, gbs=a_bytes / kt['apply'] / 1e3, frac=a_bytes / kt['apply'] / 1e3 / HBM_PEAK_GBS),
               update=dict(us=kt['update'], bytes=u_bytes, gbs=u_bytes / kt['update'] / 1e3, frac=u_bytes / kt['update'] / 1e3 / HBM_PEAK_GBS),
               finishing_us=kt['fin_apply'],
               iteration_us=sum(v for k, v in kt.items() if k not in ('slab', 'tile_rows')))
    log(f'beyond the cache ({n}^3): {out}')
    del op, b
    torch.cuda.empty_cache()
    return out


def cpu_vs_gpu_config1(args, dev):
    """Config 1 (48^3 plain FDFD) solved to rtol 1e-8 by the oracle on this host (restated
    CSR + scipy qmr, complex128, to convergence) and by generic() on the GPU."""
    import torch
    from oracle import fdfd_ref as ref
    from meanas_amd import workloads
    from meanas_amd.fdfd import solvers
    w = workloads.plain_48()
    t0 = time.perf_counter()
    E_cpu = ref.generic(w['omega'], w['dxes'], w['J'], w['epsilon'], matrix_solver_opts=dict(rtol=args.rtol, atol=0.0))
    t_cpu = time.perf_counter() - t0
    its = ref.scipy_qmr.last_iterations
    solvers.generic(w['omega'], w['dxes'], w['J'], w['epsilon'], matrix_solver_opts=dict(rtol=args.rtol))   # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    E_gpu = solvers.generic(w['omega'], w['dxes'], w['J'], w['epsilon'], matrix_solver_opts=dict(rtol=args.rtol))
    t_gpu = time.perf_counter() - t0
    info = dict(solvers.last_info)
    A0 = ref.e_full(w['omega'], w['dxes'], w['epsilon'])
    b0 = -1j * w['omega'] * w['J']
    res = lambda E: float(numpy.linalg.norm(A0 @ E - b0) / numpy.linalg.norm(b0))
    out = dict(workload='plain_48: 48^3, 10-cell SC-PML, Si core along x, Jy dipole, rtol %g' % args.rtol,
               cpu=dict(s=t_cpu, solves_per_s=1 / t_cpu, qmr_iterations=its, physical_residual=res(E_cpu),
                        cores=1, kind='port (oracle: restated CSR e_full + Pl/Pr, scipy qmr)'),
               gpu=dict(s=t_gpu, solves_per_s=1 / t_gpu, iterations=info['iterations'], physical_residual=res(E_gpu),
                        note='generic() end to end: host arrays in and out'),
               field_difference=float(numpy.linalg.norm(E_gpu - E_cpu) / numpy.linalg.norm(E_cpu)),
               gpu_over_cpu=t_cpu / t_gpu)
    log(f'config 1 on the CPU and the GPU: {out}')
    return out


def precision_study(op, b, args):
    """Why complex64 Krylov vectors do not pay on this operator: COCG (torch vector ops
    around the library's complex64 / complex128 apply plans) with complex64 rounding in
    different places, true FP64 residuals reported.  Also the per-iteration kernel times of
    the complex64 and complex128 iterations (the speed-up mixed precision could give at
    equal iteration counts)."""
    import torch
    from meanas_amd import _lib
    from meanas_amd.fdfd.device import time_kernels
    p64, p32 = op.plan('c128'), op.plan('c64')
    c64, c128 = torch.complex64, torch.complex128
    bn = float(torch.linalg.vector_norm(b))

    def dotu(u, v):
        return (u.to(c128) * v.to(c128)).sum()

    def true_rel(x):
        return float(torch.linalg.vector_norm(b - p64.apply(x.to(c128)))) / bn

    def cocg(maxit, store_p=c128, store_x=c128, vec=c128, replace_every=0, report=1000, tol=None):
        """COCG with r in `vec`; p and x_lo stored in the given precisions; FP64 residual
        replacement (x += x_lo, r = b - A x, direction kept) every `replace_every` its."""
        x = torch.zeros_like(b)
        xlo = torch.zeros_like(b, dtype=store_x)
        r = b.to(vec).clone()
        p = r.to(store_p)
        rho = dotu(r, r)
        hist, it = [], 0
        while it < maxit:
            pd = p.to(c64) if vec == c64 else p.to(c128)
            q = (p32 if vec == c64 else p64).apply(pd.contiguous())
            alpha = rho / dotu(pd, q)
            xlo = (xlo.to(c128) + alpha * pd.to(c128)).to(store_x)
            r = r - alpha.to(vec) * q
            rho_new = dotu(r, r)
            it += 1
            if replace_every and it % replace_every == 0:
                x = x + xlo.to(c128)
                xlo.zero_()
                r = (b - p64.apply(x)).to(vec)
                rho_new = dotu(r, r)
            beta = rho_new / rho
            rho = rho_new
            p = (r.to(c128) + beta * pd.to(c128)).to(store_p)
            if it % report == 0 or it == maxit:
                rel = true_rel(x + xlo.to(c128))
                hist.append((it, float('%.3g' % rel)))
                if not numpy.isfinite(rel) or (tol and rel <= tol):
                    break
        return hist
    t0 = time.perf_counter()
    out = {
        'c64_cocg': cocg(6000, store_p=c64, store_x=c64, vec=c64),
        'fp64_cocg_p_stored_c64': cocg(4000, store_p=c64),
        'fp64_cocg_x_lo_c64_replaced_every_500': cocg(9000, store_x=c64, replace_every=500, tol=args.rtol),
        'fp64_cocg': cocg(9000, tol=args.rtol),
    }
    kt = {m: time_kernels(op, b, mode, reps=50) for m, mode in (('c64', _lib.MX_MIXED), ('c128', _lib.MX_FP64))}
    per_it = {m: sum(v for k, v in kt[m].items() if k != 'slab') for m in kt}
    out['iteration_kernels_us'] = {m: {k: round(v, 2) for k, v in kt[m].items()} for m in kt}
    out['iteration_us'] = {m: round(v, 2) for m, v in per_it.items()}
    out['ceiling_at_equal_iterations'] = per_it['c128'] / per_it['c64']
    out['s'] = time.perf_counter() - t0
    log(f'precision study: {out}')
    return out


def main():
    args = parse()
    world = int(os.environ.get('WORLD_SIZE', '1'))
    if args.gpus != 1 or world != 1:
        print(f'bench.py measures one GPU only: --gpus {args.gpus} / WORLD_SIZE {world} is not supported '
              '(there is no multi-GPU path: independent systems would go one per GPU, one process each)')
        sys.exit(2)

    import torch
    from meanas_amd import _lib, workloads
    from meanas_amd.fdfd.device import DeviceOperator, solve_batch, time_kernels

    assert torch.cuda.is_available(), 'bench.py needs a GPU'
    if args.workload != 'single':
        batch_workloads(args)
        return
    dev = torch.device('cuda:0')
    w = workloads.boxes_128(args.size)
    log(f'workload {w["name"]} shape {w["shape"]}')
    op = DeviceOperator(w['omega'], w['dxes'], w['epsilon'], device=dev)
    if args.apply_sweep:
        apply_sweep(op, dev)
        return
    if args.iter_sweep:
        # kernels of the FP64 iteration in solve order (caches as in a solve) over tilings
        J = torch.from_numpy(w['J']).to(dev)
        b = op.map(_lib.MX_MAP_RHS, J, coef=-1j * w['omega'])
        for ty in (2, 4, 8):
            for slab in (4, 8, 16, 32):
                kt = time_kernels(op, b, _lib.MX_FP64, reps=30, slab=slab, tile_rows=ty)
                print(json.dumps(dict(tile_rows=ty, slab=slab, **{k: round(v, 2) for k, v in kt.items() if k != 'slab'})),
                      flush=True)
        return
    if args.apply_only:
        stream = torch.cuda.current_stream()
        for mode, dt, elem in (('c64', torch.complex64, 8), ('c128', torch.complex128, 16)):
            xa = torch.randn(op.n3, dtype=dt, device=dev)
            ya = torch.empty_like(xa)
            us = time_apply(op.plan(mode), xa, ya, args.apply_reps, stream)
            bpc = op.apply_bytes_per_cell(elem)
            print(json.dumps(dict(mode=mode, us=round(us, 2), bytes_per_cell=bpc,
                                  gbs=round(bpc * op.ncell / (us * 1e-6) / 1e9, 1))), flush=True)
        # the kernels of one FP64 COCG iteration, as the solve launches them
        J = torch.from_numpy(w['J']).to(dev)
        b = op.map(_lib.MX_MAP_RHS, J, coef=-1j * w['omega'])
        kt = time_kernels(op, b, _lib.MX_FP64, reps=args.apply_reps)
        print(json.dumps(dict(iteration_kernels_us={k: round(v, 2) for k, v in kt.items()})), flush=True)
        # a plain copy of known size: reference for scaling the FETCH_SIZE / WRITE_SIZE counters
        # (1 GiB each way, far larger than the L2 and the last-level cache, so its reads miss)
        nb = 1 << 30
        a = torch.ones(nb // 4, dtype=torch.float32, device=dev)
        c = torch.empty_like(a)
        for _ in range(3):
            c.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps(dict(copy_bytes_each_way=nb, reps=3)), flush=True)
        return
    J = torch.from_numpy(w['J']).to(dev)
    b = op.map(_lib.MX_MAP_RHS, J, coef=-1j * w['omega'])
    x = torch.zeros_like(b)
    work = {m: torch.empty(op.work_bytes(m), dtype=torch.uint8, device=dev) for m in (_lib.MX_MIXED, _lib.MX_FP64)}
    maxiter = 200000

    def solve(mode, rtol, xout):
        xout.zero_()
        form = args.form if (mode == _lib.MX_FP64 or args.form != 'fused') else 'auto'
        return solve_batch([op], [dict(op=0, b=b, x=xout, rtol=rtol, atol=0.0, maxiter=maxiter, work=work[mode])],
                           mode=mode, options=dict(form=form))[0]

    step_mode = _lib.MX_MIXED if args.mode == 'mixed' else _lib.MX_FP64
    # ---- the measured step: complete solves
    for i in range(args.warmup):
        st = solve(step_mode, args.rtol, x)
        log(f'warmup {i}: {st}')
    torch.cuda.synchronize()
    lat = []
    stats = []
    t_start = time.perf_counter()
    for i in range(args.steps):
        t0 = time.perf_counter()
        st = solve(step_mode, args.rtol, x)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t0)
        stats.append(st)
        log(f'step {i}: {lat[-1] * 1e3:.1f} ms {st}')
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    E_step = op.map(_lib.MX_MAP_E, x)
    ok = all(s['status'] == 'converged' and s['rel_residual'] <= args.rtol for s in stats)

    # ---- same-GPU comparison in the other precision mode, and the field-error reference
    extra = {}
    if not args.no_fp64:
        other = _lib.MX_MIXED if step_mode == _lib.MX_FP64 else _lib.MX_FP64
        x2 = torch.zeros_like(b)
        t0 = time.perf_counter()
        st2 = solve(other, args.rtol, x2)
        torch.cuda.synchronize()
        t2 = time.perf_counter() - t0
        log(f'comparison ({"mixed" if other == _lib.MX_MIXED else "fp64"}): {t2 * 1e3:.1f} ms {st2}')
        st_ref = solve(_lib.MX_FP64, 1e-11, x2)
        E_ref = op.map(_lib.MX_MAP_E, x2)
        ferr = float(torch.linalg.vector_norm(E_step - E_ref) / torch.linalg.vector_norm(E_ref))
        log(f'fp64 reference solve (1e-11): {st_ref}; field error of the measured solve {ferr:.3e}')
        t_step = float(numpy.median(lat))
        t_fp64, t_mixed = (t_step, t2) if step_mode == _lib.MX_FP64 else (t2, t_step)
        from meanas_amd.fdfd.autotune import tuned
        ty2, slab2 = tuned(op, other)     # the tiling that solve ran with (its iteration count depends on it)
        extra.update({'comparison': dict(mode='mixed' if other == _lib.MX_MIXED else 'fp64', solves_per_s=1.0 / t2,
                                         ms=t2 * 1e3, iterations=st2['iterations'], rel_residual=st2['rel_residual'],
                                         status=st2['status'], fp64_fallback=bool(st2.get('fp64_fallback', False)),
                                         mixed_iterations=st2.get('mixed_iterations'),
                                         tiling=dict(tile_rows=ty2, slab=slab2)),
                      'mixed_over_fp64_speedup': t_fp64 / t_mixed,
                      'field_error_vs_fp64_rtol_1e-11': ferr})
        del x2

    # ---- kernel rooflines: each kernel of one solver iteration, timed with HIP events on
    # the stream it is launched on (mx_time_kernels), same operator, right-hand side and
    # tiling as the measured solves
    stream = torch.cuda.current_stream()
    kt = time_kernels(op, b, step_mode, reps=args.apply_reps, stream=stream)
    fp64 = step_mode == _lib.MX_FP64
    elem = 16 if fp64 else 8
    n3 = op.n3
    alg = {   # algorithmic HBM bytes per launch
        'apply': op.apply_bytes_per_cell(elem) * op.ncell,              # p read, q written, materials
        # fused FP64: r, p, q read, r, p written, and x_lo read and written every other
        # launch (the deferred solution update): 6 streams on average; classic complex64:
        # x_lo, r, p, q read, x_lo, r written
        'update': 6 * elem * n3,
        'pupdate': 0 if fp64 else 3 * elem * n3,                        # fused form: no-op launch
    }
    names = {'apply': f'k_apply<{"double" if fp64 else "float"}, COCG epilogue>',
             'update': f'k_update<{"double" if fp64 else "float"}>{" (fused direction update, x_lo every other launch)" if fp64 else ""}',
             'pupdate': f'k_pupdate<{"double" if fp64 else "float"}>'}
    it_us = sum(kt[k] for k in ('apply', 'fin_apply', 'update', 'fin_update', 'pupdate'))
    kern = {k: dict(us=round(kt[k], 2), share=round(kt[k] / it_us, 3)) for k in
            ('apply', 'fin_apply', 'update', 'fin_update', 'pupdate')}
    for k, v in alg.items():
        if v:
            kern[k].update(bytes=v, gbs=round(v / (kt[k] * 1e-6) / 1e9, 1))
    dom = max(('apply', 'update'), key=lambda k: kt[k])
    log(f'iteration kernels (us): {kern}; dominant {dom}')

    def roof(k):
        ach = alg[k] / (kt[k] * 1e-6) / 1e9
        tr = TRAFFIC_MB.get((k, 'fp64' if fp64 else 'mixed', tuple(op.shape)))
        return {'bound': 'hbm', 'achieved': ach, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': ach / HBM_PEAK_GBS,
                'traffic': tr * 1e6 if tr else None, 'kernel': names[k], 'us_per_launch': kt[k],
                'algorithmic_bytes': alg[k], 'timing': 'HIP events on the launch stream (mx_time_kernels)'}
    extra['roofline_apply'] = roof('apply')
    extra['iteration_kernels'] = kern
    extra['tiling'] = dict(tile_rows=kt['tile_rows'], slab=kt['slab'], chosen_by='autotuner (tile height x slab, in-solve apply time)')
    # complex64 plain apply (the mixed solver's inner operator) for reference
    g = torch.Generator(device=dev).manual_seed(0)
    xa = torch.randn(op.n3, dtype=torch.complex64, device=dev, generator=g)
    ya = torch.empty_like(xa)
    from meanas_amd.fdfd.autotune import tuned
    ty64, slab64 = tuned(op, _lib.MX_MIXED)
    us64 = time_apply(op.plan('c64', tile_rows=ty64, slab=slab64), xa, ya, args.apply_reps, stream)
    bpc = op.apply_bytes_per_cell(8)
    extra['apply_c64'] = dict(us=us64, bytes_per_cell=bpc, gbs=bpc * op.ncell / (us64 * 1e-6) / 1e9)

    # ---- fields after the solve: H = e2h(E) and the Poynting product, on the solved field
    from meanas_amd.fdfd import functional
    f_e2h = functional.e2h(w['omega'], w['dxes'])
    f_exh = functional.poynting_e_cross_h(w['dxes'])
    E_t = E_step.contiguous()
    H_t = f_e2h(E_t)
    post = {}
    for name, fn, bpc in (('e2h', lambda: f_e2h(E_t), 96), ('poynting_e_cross_h', lambda: f_exh(E_t, H_t, True), 144)):
        for _ in range(3):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(20):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) * 1e3 / 20
        post[name] = dict(us=round(us, 2), bytes=bpc * op.ncell, gbs=round(bpc * op.ncell / (us * 1e-6) / 1e9, 1))
    log(f'post-processing: {post}')
    extra['post_processing'] = post

    # ---- the drop-in end to end: generic() from host numpy arrays to a host E (operator
    # tables, host-device copies and the solve; not the metric, whose inputs are resident)
    from meanas_amd.fdfd import solvers as gsolvers
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    E_host = gsolvers.generic(w['omega'], w['dxes'], w['J'], w['epsilon'],
                              matrix_solver_opts=dict(rtol=args.rtol))
    t_e2e = time.perf_counter() - t0
    extra['generic_end_to_end_s'] = t_e2e
    log(f'generic() end to end: {t_e2e:.3f} s ({gsolvers.last_info})')
    del E_host

    # ---- CPU baseline (oracle) on a bounded sample, rank 0 only
    cpu = None
    if not args.no_cpu_baseline:
        t0 = time.perf_counter()
        cb = cpu_baseline(w, args.cpu_iters)
        log(f'cpu baseline sample: {cb} ({time.perf_counter() - t0:.1f} s)')
        cores = len(os.sched_getaffinity(0))
        cpu = dict(value=cb.get('value'), unit='solves/s', cores=1, kind='port',
                   sample=(f'oracle (restated reference: CSR e_full + Pl/Pr, scipy {__import__("scipy").__version__} '
                           f'qmr, complex128) on the {w["shape"]} workload: CSR assembly {cb["assembly_s"]:.1f} s + '
                           f'{args.cpu_iters} qmr iterations at {cb["qmr_s_per_iter"]:.3f} s/iteration'
                           + (f', projected to {cb["qmr_iterations_to_1e8"]} iterations (full solve to 1e-8)'
                              if 'value' in cb else ', no projection (iteration count unknown)')
                           + f'; scipy CSR mat-vec is single-threaded, host has {cores} cores'),
                   assembly_s=cb['assembly_s'], s_per_iteration=cb['qmr_s_per_iter'])

    # ---- the other benchmark workloads at full size, and the precision study
    if not args.no_configs:
        extra['precision_study'] = precision_study(op, b, args)
        torch.cuda.empty_cache()
        extra['beyond_cache_256'] = beyond_cache(args, dev)
        cfg = {'config1_cpu_and_gpu': cpu_vs_gpu_config1(args, dev)}
        cfg['config3_multi_rhs'] = config3_multi_rhs(args, dev)
        torch.cuda.empty_cache()
        cfg['config4_uneven'] = {f: config4_uneven(args, dev, f) for f in ('fused', 'classic')}
        torch.cuda.empty_cache()
        cfg['config5_serving'] = {'below_saturation': config5_serving(args, dev, 1.0, 16),
                                  'saturated': config5_serving(args, dev, 4.0, 40)}
        extra['configs'] = cfg

    ms = elapsed / args.steps * 1e3
    line = {
        'metric': 'solves/s to 1e-8 residual @128^3 (+p50/p99 latency; op-apply HBM GB/s)',
        'value': args.steps / elapsed,
        'unit': 'solves/s',
        'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': ms,
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'c128' if step_mode == _lib.MX_FP64 else 'c64 inner + f64 residual/accumulation',
        'method': 'cocg_fp64' if step_mode == _lib.MX_FP64 else 'cocg_mixed',
        'data': 'synthetic (seeded box dielectrics, Gaussian sheet source)',
        'config': {'workload': f'{w["name"]}: {w["shape"][0]}x{w["shape"][1]}x{w["shape"][2]} Yee grid, '
                               '10-cell SC-PML, 40 dielectric boxes, Gaussian Jy sheet, rtol 1e-8',
                   'parallelism': 'single GPU'},
        'latency_ms': {'p50': float(numpy.percentile(lat, 50) * 1e3), 'p99': float(numpy.percentile(lat, 99) * 1e3)},
        'converged': ok,
        'iterations': stats[-1]['iterations'], 'replacements': stats[-1]['replacements'],
        'rel_residual': stats[-1]['rel_residual'],
        'roofline': roof(dom),
        'cpu_baseline': cpu,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
