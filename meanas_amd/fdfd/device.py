"""Device-resident FDFD operators and right-hand sides (torch tensors as HBM buffers).

A `DeviceOperator` holds, in HBM, everything the kernels need to apply the
symmetrised wave operator A = Pl e_full Pr of meanas.fdfd.solvers.generic
(reference fdfd/solvers.py:116-125, fdfd/operators.py:40-123) matrix-free:
  * a per-axis coefficient table: fa, fb (scaled forward differences), se, sh
    (square roots of the E- and H-grid widths, i.e. the separable factors of p);
  * epsilon, stored once if the three components are equal (isotropic) and as a real
    array if it has no imaginary part;
  * 1/mu (optional), with the same compaction;
  * one mask byte per cell for PEC (bits 0-2) and PMC (bits 3-5).
Both complex64 and complex128 copies of the coefficients and materials are kept: the
inner Krylov iteration reads the former, the FP64 residual replacement the latter.
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy
import torch

from .. import _lib


def grid_shape(dxes) -> tuple[int, int, int]:
    shape = tuple(int(numpy.size(d)) for d in dxes[0])
    if len(shape) != 3:
        raise ValueError(f'dxes must describe a 3D grid, got {len(shape)} axes')
    if tuple(int(numpy.size(d)) for d in dxes[1]) != shape:
        raise ValueError('dxes[0] and dxes[1] disagree on the grid shape')
    return shape


def coefficient_table(dxes) -> numpy.ndarray:
    """(nx+ny+nz, 4) complex128 rows (fa, fb, se, sh) for the x, then y, then z axis.

    fa[n] = 1/(se[n] sh[(n+1) mod N]), fb[n] = 1/(se[n] sh[n]) with se = sqrt(dx_e),
    sh = sqrt(dx_h) (principal roots).  These make Pl (ch pm mu^-1 ce) Pr
    (fdfd/operators.py:95,116-122) the product C^T M C of one scaled curl C.
    """
    rows = []
    for a in range(3):
        se = numpy.sqrt(numpy.asarray(dxes[0][a], dtype=complex))
        sh = numpy.sqrt(numpy.asarray(dxes[1][a], dtype=complex))
        fa = 1 / (se * numpy.roll(sh, -1))
        fb = 1 / (se * sh)
        rows.append(numpy.stack([fa, fb, se, sh], axis=1))
    return numpy.ascontiguousarray(numpy.concatenate(rows, axis=0))


def _material(values, n: int, name: str):
    """Compact a per-component material array of 3N entries.

    Returns (array, is_complex, is_aniso) with array of N (isotropic) or 3N entries.
    """
    v = numpy.asarray(values)
    if v.size != 3 * n:
        raise ValueError(f'{name} has {v.size} entries, expected 3*N = {3 * n}')
    v = v.reshape(3, n)
    is_complex = numpy.iscomplexobj(v) and bool(numpy.any(v.imag != 0))
    if not is_complex:
        v = v.real if numpy.iscomplexobj(v) else v
    aniso = not (numpy.array_equal(v[0], v[1]) and numpy.array_equal(v[0], v[2]))
    out = v.reshape(-1) if aniso else v[0]
    return numpy.ascontiguousarray(out), is_complex, aniso


def operator_flags(shape, epsilon, mu=None, pec=None, pmc=None) -> int:
    """The mx_operator flags DeviceOperator would set for these materials (host only)."""
    n = int(numpy.prod(shape))
    _, cplx, aniso = _material(epsilon, n, 'epsilon')
    flags = (_lib.MX_EPS_COMPLEX if cplx else 0) | (_lib.MX_EPS_ANISO if aniso else 0)
    if mu is not None:
        _, cplx_m, aniso_m = _material(1 / numpy.asarray(mu), n, 'mu')
        flags |= _lib.MX_MU | (_lib.MX_MU_COMPLEX if cplx_m else 0) | (_lib.MX_MU_ANISO if aniso_m else 0)
    if any(m is not None and numpy.any(numpy.asarray(m) != 0) for m in (pec, pmc)):
        flags |= _lib.MX_MASK
    return flags


def device_bytes(shape, flags: int, mode: int) -> int:
    """HBM bytes one solve of this operator holds: the per-axis tables and the materials
    in both precisions, the mask, b and x (complex128) plus the source upload, and the
    solver workspace (mx_work_bytes, a host query)."""
    nx, ny, nz = (int(v) for v in shape)
    n = nx * ny * nz

    def mat(cplx, aniso):
        return (3 if aniso else 1) * n * (8 + 4) * (2 if cplx else 1)
    total = (nx + ny + nz) * 4 * (16 + 8)
    total += mat(flags & _lib.MX_EPS_COMPLEX, flags & _lib.MX_EPS_ANISO)
    if flags & _lib.MX_MU:
        total += mat(flags & _lib.MX_MU_COMPLEX, flags & _lib.MX_MU_ANISO)
    if flags & _lib.MX_MASK:
        total += n
    total += 3 * 16 * 3 * n
    desc = _lib.MxOperator(nx, ny, nz, flags, 0.0, 0.0, None, None, None, None, None, None, None)
    return total + int(_lib.lib().mx_work_bytes(ctypes.byref(desc), mode))


def _to_device(a: numpy.ndarray, dtype: torch.dtype, device) -> torch.Tensor:
    return torch.from_numpy(numpy.ascontiguousarray(a)).to(device=device, dtype=dtype)


def torch_stream_handle(stream: torch.cuda.Stream | None = None) -> ctypes.c_void_p:
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


class DeviceOperator:
    """One structure's operator, resident in HBM, described by an `mx_operator`."""

    def __init__(self, omega: complex, dxes, epsilon, mu=None, pec=None, pmc=None,
                 device: str | torch.device = 'cuda'):
        self.shape = grid_shape(dxes)
        nx, ny, nz = self.shape
        self.ncell = nx * ny * nz
        self.n3 = 3 * self.ncell
        self.omega = complex(omega)
        self.device = torch.device(device)
        flags = 0

        coef = coefficient_table(dxes)
        self.coef64 = _to_device(coef.view(numpy.float64), torch.float64, self.device)
        self.coef32 = _to_device(coef.astype(numpy.complex64).view(numpy.float32), torch.float32, self.device)

        eps, cplx, aniso = _material(epsilon, self.ncell, 'epsilon')
        flags |= (_lib.MX_EPS_COMPLEX if cplx else 0) | (_lib.MX_EPS_ANISO if aniso else 0)
        self.eps64, self.eps32 = self._pair(eps, cplx)
        self.eps_format = (cplx, aniso)

        self.inv_mu64 = self.inv_mu32 = None
        if mu is not None:
            imu, cplx_m, aniso_m = _material(1 / numpy.asarray(mu), self.ncell, 'mu')
            flags |= _lib.MX_MU | (_lib.MX_MU_COMPLEX if cplx_m else 0) | (_lib.MX_MU_ANISO if aniso_m else 0)
            self.inv_mu64, self.inv_mu32 = self._pair(imu, cplx_m)

        self.mask = None
        if pec is not None or pmc is not None:
            bits = numpy.zeros(self.ncell, dtype=numpy.uint8)
            for base, m in ((0, pec), (3, pmc)):
                if m is None:
                    continue
                m = numpy.asarray(m).reshape(3, self.ncell) != 0
                for c in range(3):
                    bits |= (m[c].astype(numpy.uint8) << (base + c))
            if bits.any():
                flags |= _lib.MX_MASK
                self.mask = _to_device(bits, torch.uint8, self.device)
        self.flags = flags
        w2 = self.omega ** 2
        ptr = (lambda t: t.data_ptr() if t is not None else None)
        self.desc = _lib.MxOperator(nx, ny, nz, flags, w2.real, w2.imag,
                                    ptr(self.coef32), ptr(self.coef64), ptr(self.eps32), ptr(self.eps64),
                                    ptr(self.inv_mu32), ptr(self.inv_mu64), ptr(self.mask))

    def _pair(self, arr: numpy.ndarray, is_complex: bool):
        if is_complex:
            a = arr.astype(numpy.complex128)
            return (_to_device(a.view(numpy.float64), torch.float64, self.device),
                    _to_device(a.astype(numpy.complex64).view(numpy.float32), torch.float32, self.device))
        a = arr.astype(numpy.float64)
        return _to_device(a, torch.float64, self.device), _to_device(a, torch.float32, self.device)

    # ---- bytes per cell actually streamed by one complex64 apply (x read, y write,
    # materials and masks read once); the per-axis tables are negligible
    def apply_bytes_per_cell(self, elem: int = 8) -> int:
        cplx, aniso = self.eps_format
        eps_b = (2 if cplx else 1) * (elem // 2) * (3 if aniso else 1)
        mu_b = 0
        if self.inv_mu32 is not None:
            f = self.flags
            mu_b = (2 if f & _lib.MX_MU_COMPLEX else 1) * (elem // 2) * (3 if f & _lib.MX_MU_ANISO else 1)
        mask_b = 1 if self.mask is not None else 0
        return 3 * elem * 2 + eps_b + mu_b + mask_b

    # ---- thin wrappers over the C ABI ---------------------------------------------
    def work_bytes(self, mode: int) -> int:
        return int(_lib.lib().mx_work_bytes(ctypes.byref(self.desc), mode))

    def apply(self, x: torch.Tensor, y: torch.Tensor | None = None, mode: int | None = None,
              stream=None) -> torch.Tensor:
        """y = A x (scaled operator); complex64 in -> complex64 out, complex128 -> complex128."""
        if mode is None:
            mode = _lib.MX_MIXED if x.dtype == torch.complex64 else _lib.MX_FP64
        self._check_vec(x, torch.complex64 if mode == _lib.MX_MIXED else torch.complex128)
        if y is None:
            y = torch.empty_like(x)
        _lib.check(_lib.lib().mx_apply(ctypes.byref(self.desc), mode, x.data_ptr(), y.data_ptr(),
                                       torch_stream_handle(stream)), 'mx_apply')
        return y

    def plan(self, mode='c64', tile_rows: int = 0, slab: int = 0) -> 'ApplyPlan':
        """A reusable apply plan (fixed tiling, no per-call table setup)."""
        return ApplyPlan(self, mode, tile_rows, slab)

    def residual(self, b: torch.Tensor, x: torch.Tensor, r: torch.Tensor | None = None, stream=None):
        """r = b - A x in FP64; returns (r, ||r||^2)."""
        self._check_vec(b, torch.complex128)
        self._check_vec(x, torch.complex128)
        if r is None:
            r = torch.empty_like(b)
        out = ctypes.c_double(0.0)
        _lib.check(_lib.lib().mx_residual(ctypes.byref(self.desc), b.data_ptr(), x.data_ptr(), r.data_ptr(),
                                          ctypes.byref(out), torch_stream_handle(stream)), 'mx_residual')
        return r, out.value

    def map(self, which: int, vin: torch.Tensor, coef: complex = 1.0, out: torch.Tensor | None = None,
            stream=None) -> torch.Tensor:
        self._check_vec(vin, torch.complex128)
        if out is None:
            out = torch.empty_like(vin)
        _lib.check(_lib.lib().mx_map(ctypes.byref(self.desc), which, coef.real, coef.imag,
                                     vin.data_ptr(), out.data_ptr(), torch_stream_handle(stream)), 'mx_map')
        return out

    def _check_vec(self, v: torch.Tensor, dtype: torch.dtype) -> None:
        if v.dtype != dtype or not v.is_contiguous() or v.numel() != self.n3 or v.device.type != 'cuda':
            raise ValueError(f'expected a contiguous {dtype} CUDA tensor of {self.n3} entries, '
                             f'got {v.dtype} {tuple(v.shape)} on {v.device}')


def _mode(mode) -> int:
    if mode in ('c64', _lib.MX_MIXED):
        return _lib.MX_MIXED
    if mode in ('c128', _lib.MX_FP64):
        return _lib.MX_FP64
    raise ValueError(f'unknown mode {mode!r}')


class ApplyPlan:
    """mx_plan wrapper: y = A x with a fixed tile height and slab depth."""

    def __init__(self, op: DeviceOperator, mode='c64', tile_rows: int = 0, slab: int = 0):
        self.op = op
        self.mode = _mode(mode)
        self.dtype = torch.complex64 if self.mode == _lib.MX_MIXED else torch.complex128
        L = _lib.lib()
        self.handle = L.mx_plan_create(ctypes.byref(op.desc), self.mode, tile_rows, slab)
        if not self.handle:
            raise _lib.MxError('mx_plan_create failed: ' + L.mx_last_error().decode(errors='replace'))
        self.blocks = int(L.mx_plan_blocks(self.handle))

    def apply(self, x: torch.Tensor, y: torch.Tensor | None = None, stream=None) -> torch.Tensor:
        self.op._check_vec(x, self.dtype)
        if y is None:
            y = torch.empty_like(x)
        else:
            self.op._check_vec(y, self.dtype)
        _lib.check(_lib.lib().mx_plan_apply(self.handle, x.data_ptr(), y.data_ptr(), torch_stream_handle(stream)),
                   'mx_plan_apply')
        return y

    def __del__(self):
        h = getattr(self, 'handle', None)
        if h:
            _lib.lib().mx_plan_destroy(h)
            self.handle = None


def project_masked(op: DeviceOperator, b: torch.Tensor) -> tuple[torch.Tensor, float]:
    """Split off the part of b on PEC entries.

    e_full has zero rows (and columns) on PEC entries (fdfd/operators.py:84-95), so b's
    entries there can never be matched: A x = b is inconsistent, and COCG on it would mix
    those constant entries into r^T r.  Returns (b with them zeroed, ||b on them||^2);
    the solve of the projected system is the least-squares solution of the original,
    whose residual norm is exactly that floor.  b itself is returned when nothing is
    masked or b vanishes there.
    """
    if op.mask is None:
        return b, 0.0
    bits = op.mask.to(torch.int32)
    pec = torch.cat([((bits >> c) & 1).bool() for c in range(3)])
    floor2 = float(torch.sum(torch.abs(b[pec]) ** 2).item())
    if floor2 == 0.0:
        return b, 0.0
    bp = b.clone()
    bp[pec] = 0
    return bp, floor2


def mark_inconsistent(st: dict, floor2: float) -> dict:
    """Stats of a solve whose right-hand side had a part on PEC entries (project_masked):
    status 'inconsistent'; rel_residual becomes the returned iterate's residual relative to
    the full ||b|| (the floor included); residual_floor is the floor's share of it."""
    bn2 = st['bnorm'] ** 2 + floor2
    free2 = (st['rel_residual'] * st['bnorm']) ** 2
    out = dict(st, bnorm=bn2 ** 0.5, rel_residual=((free2 + floor2) / bn2) ** 0.5,
               residual_floor=(floor2 / bn2) ** 0.5, solved_status=st['status'])
    out['status'] = 'inconsistent'
    return out


def solve_batch(ops: Sequence[DeviceOperator], systems: Sequence[dict], mode: int = _lib.MX_MIXED,
                options: dict | None = None, stream=None, fp64_fallback: bool = True) -> list[dict]:
    """Solve several systems in one batched launch sequence.

    systems: dicts with keys op (index into ops), b (complex128 tensor), x (complex128
    tensor, in: guess, out: solution), rtol, atol, maxiter; optional work (uint8 tensor).
    With fp64_fallback, mixed-precision systems that stall (the complex64 inner solve
    cannot resolve the operator, e.g. near-static omega with kappa ~ 1/eps32) continue
    in FP64 COCG from their FP64 iterate; their stats then carry fp64_fallback=True and
    the summed iteration counts.
    A right-hand side with a part on PEC entries is solved without it (project_masked)
    and reported with status 'inconsistent' (mark_inconsistent).
    """
    floors = []
    systems = list(systems)
    for i, s in enumerate(systems):
        bp, f2 = project_masked(ops[s['op']], s['b'])
        floors.append(f2)
        if f2 > 0:
            systems[i] = dict(s, b=bp)
    stats = _solve_once(ops, systems, mode, options, stream)
    if mode == _lib.MX_MIXED:
        # every mixed-precision result says whether complex64 finished it on its own
        stats = [dict(st, fp64_fallback=False, mixed_iterations=st['iterations']) for st in stats]
    if mode == _lib.MX_MIXED and fp64_fallback:
        redo = [i for i, st in enumerate(stats) if st['status'] in ('stalled', 'breakdown')]
        if redo:
            sub = []
            for i in redo:
                s = dict(systems[i])
                s.pop('work', None)
                s['maxiter'] = max(1, int(s.get('maxiter', 10 * ops[s['op']].n3)) - stats[i]['iterations'])
                sub.append(s)
            more = _solve_once(ops, sub, _lib.MX_FP64, options, stream)
            for i, st in zip(redo, more):
                st = dict(st, fp64_fallback=True, iterations=st['iterations'] + stats[i]['iterations'],
                          replacements=st['replacements'] + stats[i]['replacements'],
                          mixed_iterations=stats[i]['iterations'])
                stats[i] = st
    return [mark_inconsistent(st, f2) if f2 > 0 else st for st, f2 in zip(stats, floors)]


def _solve_once(ops, systems, mode, options, stream) -> list[dict]:
    o = dict(tile_rows=0, slab=0, chunk=0, check_every=0, use_graph=1, delta=0.0, stall_iters=0, stall_limit=0,
             autotune=True, form='auto')
    if options:
        o.update(options)
    descs = (_lib.MxOperator * len(ops))(*[op.desc for op in ops])
    sysarr = (_lib.MxSystem * len(systems))()
    keep = []
    for i, s in enumerate(systems):
        op = ops[s['op']]
        op._check_vec(s['b'], torch.complex128)
        op._check_vec(s['x'], torch.complex128)
        work = s.get('work')
        if work is None:
            work = torch.empty(op.work_bytes(mode), dtype=torch.uint8, device=op.device)
        keep.append(work)
        slab, ty = int(s.get('slab', 0)), int(s.get('tile_rows', 0))
        if slab == 0 and o['slab'] == 0 and o['tile_rows'] == 0 and ty == 0 and o.get('autotune', True):
            from .autotune import tuned
            ty, slab = tuned(op, mode)
        sysarr[i] = _lib.MxSystem(s['op'], int(s.get('maxiter', 10 * op.n3)), slab, ty, float(s.get('rtol', 1e-5)),
                                  float(s.get('atol', 0.0)), s['b'].data_ptr(), s['x'].data_ptr(), work.data_ptr())
    opts = _lib.MxOptions(mode, o['tile_rows'], o['slab'], o['chunk'], o['check_every'], o['use_graph'], o['delta'],
                          o['stall_iters'], o['stall_limit'], _lib.FORMS[o['form']])
    stats = (_lib.MxStats * len(systems))()
    _lib.check(_lib.lib().mx_solve(descs, len(ops), sysarr, len(systems), ctypes.byref(opts), stats,
                                   torch_stream_handle(stream)), 'mx_solve')
    return [st.as_dict() for st in stats]


KERNEL_NAMES = ('apply', 'fin_apply', 'update', 'fin_update', 'pupdate')


def time_kernels(op: DeviceOperator, b: torch.Tensor, mode: int = _lib.MX_FP64, reps: int = 50,
                 slab: int | None = None, stream=None, tile_rows: int = 0, form: str = 'auto') -> dict:
    """Average microseconds per launch of each kernel of one solver iteration
    (`mx_time_kernels`), timed with HIP events on `stream` on the system A x = b.
    The tiling is the one a solve of this operator uses (the autotuned tile height and slab
    depth unless given)."""
    op._check_vec(b, torch.complex128)
    x = torch.zeros_like(b)
    work = torch.empty(op.work_bytes(mode), dtype=torch.uint8, device=op.device)
    if slab is None:
        from .autotune import tuned
        ty, slab = tuned(op, mode)
        tile_rows = tile_rows or ty
    sysd = _lib.MxSystem(0, 1, int(slab), int(tile_rows), 0.0, 0.0, b.data_ptr(), x.data_ptr(), work.data_ptr())
    opts = _lib.MxOptions(mode, int(tile_rows), 0, 0, 0, 1, 0.0, 0, 0, _lib.FORMS[form])
    us = (ctypes.c_double * len(KERNEL_NAMES))()
    _lib.check(_lib.lib().mx_time_kernels(ctypes.byref(op.desc), ctypes.byref(sysd), ctypes.byref(opts), int(reps),
                                          us, len(KERNEL_NAMES), torch_stream_handle(stream)), 'mx_time_kernels')
    return dict(zip(KERNEL_NAMES, list(us)), slab=int(slab), tile_rows=int(tile_rows))
