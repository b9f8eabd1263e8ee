"""Continuous-batching FDFD solve server on one GPU.

Requests (the arguments of meanas.fdfd.solvers.generic, forward or adjoint; reference
meanas/fdfd/solvers.py:66-141) arrive over time.  The server keeps a session of slots on
the device (`mx_session_*`, C ABI): every step runs `chunk` COCG iterations on all
active slots, launch group by launch group, with groups of smaller systems repeating
their chunk (about the same time per group, up to 16 chunks), so a small request is not
held to the pace of a large one; finished slots are read out and retired and
waiting requests are admitted as soon as a slot and enough HBM are free, so new work
joins while other solves are in flight.  A request's result does not depend on what
else is in the session: each system's block partition and reduction order depend on
that system alone.

Admission is by HBM footprint, worked out on the host before anything is allocated:
the operator tables and materials (complex or anisotropic eps, 1/mu, the PEC/PMC mask),
b, x and the source upload, and the solver workspace (`mx_work_bytes`).  A request's
footprint is reserved when its host preparation starts, and that preparation (PML
tables, material compaction, uploads) overlaps with the iterations already queued.

Sessions are kept per lane, a (precision, iteration form) pair.  In mixed precision
(complex64 iterations, FP64 residual replacement) a request whose complex64 iterations
stall or break down continues in an FP64 session from its best FP64 iterate, as
`solve_batch` does; its result carries fp64_fallback and the summed iteration counts.
FP64 sessions run the fused iteration form (three launches per iteration); a request
whose fused recursion loses track finishes there with MX_FLAG_SWITCHED and continues
from its iterate in a classic-form FP64 session, as `mx_solve` does, with the summed
iteration counts.
"""
from __future__ import annotations

import collections
import ctypes
import time
from dataclasses import dataclass, field

import numpy
import torch

from . import _lib
from .fdfd.device import (DeviceOperator, device_bytes, mark_inconsistent, operator_flags, project_masked,
                          torch_stream_handle)


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
    op: DeviceOperator | None
    b: torch.Tensor | None
    x: torch.Tensor | None
    work: torch.Tensor | None
    footprint: int
    mode: int
    form: str = 'fused'            # iteration form of its lane: 'fused' or 'classic'
    admitted: float = 0.0
    prior: dict | None = None      # summed stats of the phases this slot continues
    extra: dict = field(default_factory=dict)

    @property
    def lane(self) -> tuple[int, str]:
        return self.mode, self.form


def request_shape(req: Request) -> tuple[int, int, int]:
    return tuple(int(numpy.size(d)) for d in req.dxes[0])


def hand_over_prior(prior: dict | None, st: dict, lane: tuple[int, str], to_form: str) -> dict:
    """Counts of the phases a continued request has run: `prior` (None before its first
    hand-over) joined with the phase `st` that ran in `lane`, now handed to a lane of
    form `to_form`.  A complex64 phase marks fp64_fallback and mixed_iterations; a fused
    FP64 phase handed to the classic form, fused_iterations."""
    p = prior or dict(iterations=0, replacements=0)
    out = dict(p, iterations=p['iterations'] + st['iterations'], replacements=p['replacements'] + st['replacements'])
    mode, form = lane
    if mode == _lib.MX_MIXED:
        out.update(fp64_fallback=True, mixed_iterations=st['iterations'])
    elif form == 'fused' and to_form == 'classic':
        out.update(fused_iterations=st['iterations'])
    return out


def merge_prior(st: dict, prior: dict | None, lane: tuple[int, str]) -> dict:
    """A request's final statistics: its last phase `st`, run in `lane`, with the counts of
    the phases before it."""
    if lane[0] == _lib.MX_MIXED:   # finished by its complex64 iterations
        st = dict(st, fp64_fallback=False, mixed_iterations=st['iterations'])
    if prior:
        st = dict(st, iterations=st['iterations'] + prior['iterations'],
                  replacements=st['replacements'] + prior['replacements'])
        for k in ('fp64_fallback', 'mixed_iterations', 'fused_iterations'):
            if k in prior:
                st[k] = prior[k]
    return st


def footprint_bytes(req: Request, mode: int) -> int:
    """HBM bytes one admitted request holds, from its shape and material formats (host only)."""
    shape = request_shape(req)
    return device_bytes(shape, operator_flags(shape, req.epsilon, req.mu, req.pec, req.pmc), mode)


class Server:
    """Continuous-batching solver over `mx_session`s (one per lane in use: precision and
    iteration form)."""

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
        self._o = o
        self.lane = (mode, self._form(mode))
        self.sessions: dict[tuple[int, str], int] = {}
        self._session(self.lane)
        self.slots: dict[tuple[tuple[int, str], int], _Slot] = {}
        self.used = 0                   # bytes reserved by prepared and admitted requests

    def _form(self, mode: int) -> str:
        """The iteration form of `mode`'s lane: the options' form where the precision allows
        it (complex64 iterations use the classic form), fused for FP64 by default."""
        form = self._o['form']
        if mode != _lib.MX_FP64:
            return 'classic'
        return 'fused' if form == 'auto' else form

    def _session(self, lane: tuple[int, str]):
        if lane not in self.sessions:
            o = self._o
            mode, form = lane
            opts = _lib.MxOptions(mode, o['tile_rows'], o['slab'], o['chunk'], o['check_every'], o['use_graph'],
                                  o['delta'], o['stall_iters'], o['stall_limit'], _lib.FORMS[form])
            L = _lib.lib()
            h = L.mx_session_create(self.capacity, ctypes.byref(opts))
            if not h:
                raise _lib.MxError('mx_session_create failed: ' + L.mx_last_error().decode(errors='replace'))
            self.sessions[lane] = h
        return self.sessions[lane]

    @property
    def handle(self):
        return self.sessions.get(self.lane)

    def close(self):
        for h in getattr(self, 'sessions', {}).values():
            _lib.lib().mx_session_destroy(h)
        self.sessions = {}

    def __del__(self):
        self.close()

    # ---- admission ---------------------------------------------------------------
    def active(self, lane: tuple[int, str]) -> int:
        return sum(1 for (ln, _) in self.slots if ln == lane)

    def fits(self, nbytes: int) -> bool:
        return self.used + nbytes <= self.hbm_budget

    def prepare(self, req: Request, footprint: int | None = None) -> _Slot:
        """Reserve the request's footprint, then build its operator and b on the device.
        Raises MxError (nothing allocated, nothing reserved) when it does not fit."""
        fp = footprint_bytes(req, self.mode) if footprint is None else footprint
        if not self.fits(fp):
            raise _lib.MxError(f'request of shape {request_shape(req)} needs {fp} bytes of HBM; '
                               f'{self.hbm_budget - self.used} of the {self.hbm_budget}-byte budget are free')
        self.used += fp
        try:
            op = DeviceOperator(req.omega, req.dxes, req.epsilon, mu=req.mu, pec=req.pec, pmc=req.pmc,
                                device=self.device)
            J = torch.from_numpy(numpy.ascontiguousarray(req.J, dtype=numpy.complex128)).to(self.device)
            which = _lib.MX_MAP_RHS_ADJ if req.adjoint else _lib.MX_MAP_RHS
            b = op.map(which, J, coef=-1j * complex(req.omega))
            del J
            b, floor2 = project_masked(op, b)      # a source inside PEC cells: least squares
            x = torch.zeros_like(b)
            work = torch.empty(op.work_bytes(self.mode), dtype=torch.uint8, device=self.device)
        except Exception:
            self.used -= fp
            raise
        return _Slot(req, op, b, x, work, fp, self.mode, self.lane[1], extra=dict(floor2=floor2))

    def admit(self, slot: _Slot, now: float) -> int:
        req = slot.req
        maxiter = req.maxiter if req.maxiter is not None else 10 * slot.op.n3
        if slot.prior:
            maxiter = max(1, maxiter - slot.prior['iterations'])
        from .fdfd.autotune import tuned
        ty, slab = tuned(slot.op, slot.mode)
        sysd = _lib.MxSystem(0, int(min(maxiter, 2 ** 31 - 1)), slab, ty, req.rtol, req.atol,
                             slot.b.data_ptr(), slot.x.data_ptr(), slot.work.data_ptr())
        h = self._session(slot.lane)
        idx = _lib.lib().mx_session_admit(h, ctypes.byref(slot.op.desc), ctypes.byref(sysd), torch_stream_handle())
        if idx < 0:
            raise _lib.MxError('mx_session_admit failed: ' + _lib.lib().mx_last_error().decode(errors='replace'))
        if not slot.prior:
            slot.admitted = now
        self.slots[(slot.lane, idx)] = slot
        return idx

    def step(self) -> None:
        for lane, h in self.sessions.items():
            if self.active(lane):
                _lib.check(_lib.lib().mx_session_step(h, self.chunk, None), 'mx_session_step')

    def poll(self, lane: tuple[int, str]):
        stats = (_lib.MxStats * self.capacity)()
        done = (ctypes.c_int32 * self.capacity)()
        _lib.check(_lib.lib().mx_session_poll(self.sessions[lane], stats, done, torch_stream_handle()),
                   'mx_session_poll')
        return stats, done

    def _hand_over(self, slot: _Slot, st: dict, mode: int, form: str) -> _Slot:
        """The slot, to be continued in lane (mode, form) from its iterate; `st`'s counts
        join those of the phases before it."""
        slot.prior = hand_over_prior(slot.prior, st, slot.lane, form)
        slot.work = None
        self.used -= slot.footprint
        slot.footprint = footprint_bytes(slot.req, mode)
        slot.mode, slot.form = mode, form
        return slot

    def finish(self, key, stats: _lib.MxStats, now: float):
        """Retire a finished slot: a Result, or the slot to continue in another lane from
        its iterate (its workspace traded for one of that lane): complex64 iterations that
        stalled continue in FP64, a fused FP64 recursion that lost track in the classic
        form."""
        lane, idx = key
        slot = self.slots.pop(key)
        _lib.check(_lib.lib().mx_session_retire(self.sessions[lane], idx), 'mx_session_retire')
        st = stats.as_dict()
        if lane[0] == _lib.MX_MIXED and st['status'] in ('stalled', 'breakdown'):
            return self._hand_over(slot, st, _lib.MX_FP64, self._form(_lib.MX_FP64))
        if st['flags'] & _lib.MX_FLAG_SWITCHED:
            return self._hand_over(slot, st, _lib.MX_FP64, 'classic')
        st = merge_prior(st, slot.prior, lane)
        if slot.extra.get('floor2', 0.0) > 0:
            st = mark_inconsistent(st, slot.extra['floor2'])
        which = _lib.MX_MAP_E_ADJ if slot.req.adjoint else _lib.MX_MAP_E
        E = slot.op.map(which, slot.x).cpu().numpy()
        self.used -= slot.footprint
        return Result(slot.req.tag, E, st, slot.req.arrival, slot.admitted, now, slot.op.shape)

    def _continue(self, slot: _Slot, now: float) -> None:
        self.used += slot.footprint
        slot.work = torch.empty(slot.op.work_bytes(slot.mode), dtype=torch.uint8, device=self.device)
        self.admit(slot, now)

    # ---- the request loop ------------------------------------------------------------
    def run(self, requests, realtime: bool = True) -> list[Result]:
        """Serve `requests` (sorted by arrival).  With realtime, arrivals follow the wall
        clock from the start of the run; otherwise every request is available at once."""
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        ready: collections.deque[_Slot] = collections.deque()
        cont: collections.deque[_Slot] = collections.deque()      # continuations in another lane
        results: list[Result] = []
        t0 = time.perf_counter()
        clock = (lambda: time.perf_counter() - t0)
        next_fp = None
        while pending or ready or cont or self.slots:
            now = clock()
            if self.slots:
                self.step()                       # queued: overlaps the host work below
            while cont and self.active(cont[0].lane) < self.capacity and self.fits(cont[0].footprint):
                self._continue(cont.popleft(), clock())
            # prepare arrived requests while the GPU iterates; the footprint is checked
            # and reserved before anything is allocated for the request
            while pending and (not realtime or pending[0].arrival <= now) and \
                    len(ready) + self.active(self.lane) < self.capacity:
                if next_fp is None:
                    next_fp = footprint_bytes(pending[0], self.mode)
                if not self.fits(next_fp):
                    break
                ready.append(self.prepare(pending.popleft(), next_fp))
                next_fp = None
            for lane in list(self.sessions):
                if not self.active(lane):
                    continue
                stats, done = self.poll(lane)
                now = clock()
                for key in [k for k in list(self.slots) if k[0] == lane and done[k[1]] == 1]:
                    out = self.finish(key, stats[key[1]], now)
                    if isinstance(out, Result):
                        results.append(out)
                    else:
                        cont.append(out)
            while ready and self.active(self.lane) < self.capacity:
                self.admit(ready.popleft(), clock())
            if not self.slots and not ready and not cont and pending:
                if next_fp is not None and not self.fits(next_fp) and (not realtime or pending[0].arrival <= clock()):
                    raise _lib.MxError(f'request of shape {request_shape(pending[0])} needs {next_fp} bytes of '
                                       f'HBM, more than the {self.hbm_budget}-byte budget')
                if realtime:
                    wait = pending[0].arrival - clock()
                    if wait > 0:
                        time.sleep(min(wait, 0.01))
            if not self.slots and not ready and cont and not self.fits(cont[0].footprint):
                raise _lib.MxError(f'continuation of shape {cont[0].op.shape} does not fit the HBM budget')
        return results


def summarize(results: list[Result], wall: float) -> dict:
    lat = numpy.array([r.latency for r in results]) if results else numpy.zeros(1)
    return dict(solves=len(results), wall_s=wall, solves_per_s=len(results) / wall if wall > 0 else 0.0,
                p50_ms=float(numpy.percentile(lat, 50) * 1e3), p99_ms=float(numpy.percentile(lat, 99) * 1e3),
                converged=sum(r.stats['status'] == 'converged' for r in results),
                adjoint=sum(1 for r in results if r.tag and isinstance(r.tag, dict) and r.tag.get('adjoint')))
