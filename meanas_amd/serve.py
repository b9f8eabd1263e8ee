"""Continuous-batching FDFD solve server on one GPU.

Requests (the arguments of meanas.fdfd.solvers.generic, forward or adjoint) arrive over
time.  The server keeps a session of slots on the device (`mx_session_*`, C ABI): every
step runs `chunk` COCG iterations on all active slots in one set of launches; finished
slots are read out and retired and waiting requests are admitted as soon as a slot and
enough HBM are free, so new work joins while other solves are in flight.  A request's
result does not depend on what else is in the session: each system's block partition
and reduction order depend on that system alone.

Admission is by HBM footprint: operator tables and materials, b and x (complex128) and
the solver workspace (`mx_work_bytes`).  Host preparation of arrived requests (PML
tables, material compaction, uploads) overlaps with the iterations already queued.
"""
from __future__ import annotations

import collections
import ctypes
import time
from dataclasses import dataclass, field

import numpy
import torch

from . import _lib
from .fdfd.device import DeviceOperator, torch_stream_handle


@dataclass
class Request:
    omega: complex
    dxes: list
    J: numpy.ndarray
    epsilon: numpy.ndarray
    mu: numpy.ndarray | None = None
    pec: numpy.ndarray | None = None
    pmc: numpy.ndarray | None = None
    adjoint: bool = False
    rtol: float = 1e-8
    atol: float = 0.0
    maxiter: int | None = None
    arrival: float = 0.0          # seconds after the start of the run
    tag: object = None


@dataclass
class Result:
    tag: object
    E: numpy.ndarray
    stats: dict
    arrival: float
    admitted: float
    finished: float
    shape: tuple

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


@dataclass
class _Slot:
    req: Request
    op: DeviceOperator
    b: torch.Tensor
    x: torch.Tensor
    work: torch.Tensor
    footprint: int
    admitted: float = 0.0
    extra: dict = field(default_factory=dict)


def footprint_bytes(shape, mode: int) -> int:
    """HBM bytes one admitted request holds (approximate for the operator tables)."""
    n3 = 3 * int(numpy.prod(shape))
    vec = 16 * n3                                    # b and x, complex128
    work = 4 * (8 if mode == _lib.MX_MIXED else 16) * n3 + 16 * n3
    mats = 12 * n3                                   # eps (f32 + f64), up to anisotropic
    return 2 * vec + work + mats


class Server:
    """Continuous-batching solver over one `mx_session`."""

    def __init__(self, capacity: int = 8, mode: int = _lib.MX_FP64, chunk: int = 16,
                 hbm_budget: int | None = None, device: str = 'cuda', options: dict | None = None):
        self.capacity = capacity
        self.mode = mode
        self.chunk = chunk
        self.device = torch.device(device)
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            hbm_budget = int(0.8 * free)
        self.hbm_budget = hbm_budget
        o = dict(tile_rows=0, slab=0, chunk=chunk, check_every=0, use_graph=1, delta=0.0,
                 stall_iters=0, stall_limit=0, form='auto')
        if options:
            o.update(options)
        self.opts = _lib.MxOptions(mode, o['tile_rows'], o['slab'], o['chunk'], o['check_every'], o['use_graph'],
                                   o['delta'], o['stall_iters'], o['stall_limit'], _lib.FORMS[o['form']])
        L = _lib.lib()
        self.handle = L.mx_session_create(capacity, ctypes.byref(self.opts))
        if not self.handle:
            raise _lib.MxError('mx_session_create failed: ' + L.mx_last_error().decode(errors='replace'))
        self.slots: dict[int, _Slot] = {}
        self.used = 0

    def close(self):
        if getattr(self, 'handle', None):
            _lib.lib().mx_session_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()

    # ---- request preparation (host + uploads on torch's stream) ---------------------
    def prepare(self, req: Request) -> _Slot:
        op = DeviceOperator(req.omega, req.dxes, req.epsilon, mu=req.mu, pec=req.pec, pmc=req.pmc,
                            device=self.device)
        J = torch.from_numpy(numpy.ascontiguousarray(req.J, dtype=numpy.complex128)).to(self.device)
        which = _lib.MX_MAP_RHS_ADJ if req.adjoint else _lib.MX_MAP_RHS
        b = op.map(which, J, coef=-1j * complex(req.omega))
        x = torch.zeros_like(b)
        work = torch.empty(op.work_bytes(self.mode), dtype=torch.uint8, device=self.device)
        return _Slot(req, op, b, x, work, footprint_bytes(op.shape, self.mode))

    def fits(self, shape) -> bool:
        return (len(self.slots) < self.capacity and
                self.used + footprint_bytes(shape, self.mode) <= self.hbm_budget)

    def admit(self, slot: _Slot, now: float) -> int:
        req = slot.req
        maxiter = req.maxiter if req.maxiter is not None else 10 * slot.op.n3
        from .fdfd.autotune import tuned_slab
        sysd = _lib.MxSystem(0, int(min(maxiter, 2 ** 31 - 1)), tuned_slab(slot.op, self.mode), 0, req.rtol, req.atol,
                             slot.b.data_ptr(), slot.x.data_ptr(), slot.work.data_ptr())
        idx = _lib.lib().mx_session_admit(self.handle, ctypes.byref(slot.op.desc), ctypes.byref(sysd),
                                          torch_stream_handle())
        if idx < 0:
            raise _lib.MxError('mx_session_admit failed: ' + _lib.lib().mx_last_error().decode(errors='replace'))
        slot.admitted = now
        self.slots[idx] = slot
        self.used += slot.footprint
        return idx

    def step(self) -> None:
        _lib.check(_lib.lib().mx_session_step(self.handle, self.chunk, None), 'mx_session_step')

    def poll(self):
        stats = (_lib.MxStats * self.capacity)()
        done = (ctypes.c_int32 * self.capacity)()
        _lib.check(_lib.lib().mx_session_poll(self.handle, stats, done, torch_stream_handle()), 'mx_session_poll')
        return stats, done

    def finish(self, idx: int, stats: _lib.MxStats, now: float) -> Result:
        slot = self.slots.pop(idx)
        which = _lib.MX_MAP_E_ADJ if slot.req.adjoint else _lib.MX_MAP_E
        E = slot.op.map(which, slot.x).cpu().numpy()
        _lib.check(_lib.lib().mx_session_retire(self.handle, idx), 'mx_session_retire')
        self.used -= slot.footprint
        return Result(slot.req.tag, E, stats.as_dict(), slot.req.arrival, slot.admitted, now, slot.op.shape)

    # ---- the request loop ------------------------------------------------------------
    def run(self, requests, realtime: bool = True) -> list[Result]:
        """Serve `requests` (sorted by arrival).  With realtime, arrivals follow the wall
        clock from the start of the run; otherwise every request is available at once."""
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        ready: collections.deque[_Slot] = collections.deque()
        results: list[Result] = []
        t0 = time.perf_counter()
        clock = (lambda: time.perf_counter() - t0)
        while pending or ready or self.slots:
            now = clock()
            if self.slots:
                self.step()                       # queued: overlaps the host work below
            # prepare arrived requests while the GPU iterates
            while pending and (not realtime or pending[0].arrival <= now) and \
                    len(ready) + len(self.slots) < self.capacity:
                ready.append(self.prepare(pending.popleft()))
            if self.slots:
                stats, done = self.poll()
                now = clock()
                for i in [i for i in list(self.slots) if done[i] == 1]:
                    results.append(self.finish(i, stats[i], now))
            while ready and self.fits(ready[0].op.shape):
                self.admit(ready.popleft(), clock())
            if not self.slots and not ready and pending and realtime:
                wait = pending[0].arrival - clock()
                if wait > 0:
                    time.sleep(min(wait, 0.01))
            if ready and not self.slots and not self.fits(ready[0].op.shape):
                raise _lib.MxError(f'request of shape {ready[0].op.shape} does not fit the HBM budget')
        return results


def summarize(results: list[Result], wall: float) -> dict:
    lat = numpy.array([r.latency for r in results]) if results else numpy.zeros(1)
    return dict(solves=len(results), wall_s=wall, solves_per_s=len(results) / wall if wall > 0 else 0.0,
                p50_ms=float(numpy.percentile(lat, 50) * 1e3), p99_ms=float(numpy.percentile(lat, 99) * 1e3),
                converged=sum(r.stats['status'] == 'converged' for r in results),
                adjoint=sum(1 for r in results if r.tag and isinstance(r.tag, dict) and r.tag.get('adjoint')))
