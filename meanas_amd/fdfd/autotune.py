"""Autotuner for the apply kernel's tile shape: tile height (y rows per block) and slab
depth (x-planes per block).

Each block of the apply kernel owns a TY x 64 (y, z) tile and marches over `slab`
x-planes.  A deeper slab amortises the two extra planes a block reads at its start; a
shallower one, or a lower tile, gives more blocks to fill the 256 CUs; the register
budget of the kernel decides how many of them are resident.  The best shape depends on
the grid, the precision and the material kinds, so it is measured: every candidate
(tile height, slab) is timed with `mx_time_kernels`, i.e. the solver's own iteration
launched in solve order on the real operator (so the caches hold what they hold in a
solve), and the fastest apply is kept per (device, shape, precision, material flags), in
memory and in a JSON cache file.  Shapes within 3% of the fastest (timing noise) are
ranked by a fixed preference (4-row tiles, then the deepest slab), so that near-ties
resolve the same way on every box.

A system's block partition -- and so the order of its reductions -- depends on that
system alone (its shape and materials through this table), never on the rest of a
batch, so a system gives the same bits alone and in any batch.
Set MEANAS_AMD_AUTOTUNE=0 to use the library's built-in rule (4-row tiles, automatic
slab) instead.
"""
from __future__ import annotations

import json
import os
import threading

import torch

from .. import _lib

TILE_ROWS = (2, 4, 8)
SLABS = (2, 4, 6, 8, 11, 12, 16, 24, 32)
MAX_PARTIALS = 4096       # per-system reduction limit of the finishing kernel
# Shapes within this fraction of the fastest count as tied (timing noise) and are ranked
# by a fixed preference instead: near-ties would otherwise resolve differently on
# different boxes, and a system's bits -- even its Krylov iteration count, through the
# order of its reductions -- depend on its block partition.
MARGIN = 0.03
CACHE_PATH = os.path.join(os.path.expanduser('~'), '.cache', 'meanas_amd', 'autotune3.json')

_lock = threading.Lock()
_memo: dict[str, tuple[int, int]] = {}
_loaded = False


def blocks(shape, tile_rows: int, slab: int) -> int:
    nx, ny, nz = shape
    return ((ny + tile_rows - 1) // tile_rows) * ((nz + 63) // 64) * ((nx + slab - 1) // slab)


def choose(times: dict[tuple[int, int], float]) -> tuple[int, int]:
    """Among the shapes within MARGIN of the fastest, the one a fixed preference ranks
    first: 4-row tiles, then the deepest slab, then the lower tile."""
    best = min(times.values())
    near = [c for c, t in times.items() if t <= (1.0 + MARGIN) * best]
    return min(near, key=lambda c: (c[0] != 4, -c[1], c[0]))


def candidates(shape) -> list[tuple[int, int]]:
    """Admissible (tile height, slab) pairs: slab no deeper than the grid (the shallowest
    candidate always), and at most MAX_PARTIALS blocks."""
    out = []
    for ty in TILE_ROWS:
        for slab in SLABS:
            if slab > shape[0] and slab != SLABS[0]:
                continue
            if blocks(shape, ty, slab) > MAX_PARTIALS:
                continue
            out.append((ty, slab))
    return out


def _key(op, mode: int) -> str:
    name = torch.cuda.get_device_name(op.device)
    return f'{name}|{op.shape[0]}x{op.shape[1]}x{op.shape[2]}|{"c64" if mode == _lib.MX_MIXED else "c128"}|f{op.flags}'


def _load():
    global _loaded
    if _loaded:
        return
    _loaded = True
    try:
        with open(CACHE_PATH) as f:
            _memo.update({k: (int(v[0]), int(v[1])) for k, v in json.load(f).items()})
    except (OSError, ValueError, TypeError, IndexError):
        pass


def _save():
    try:
        os.makedirs(os.path.dirname(CACHE_PATH), exist_ok=True)
        tmp = CACHE_PATH + '.tmp'
        with open(tmp, 'w') as f:
            json.dump({k: list(v) for k, v in _memo.items()}, f, indent=0, sort_keys=True)
        os.replace(tmp, CACHE_PATH)
    except OSError:
        pass


def measure(op, mode: int, reps: int = 8) -> dict[tuple[int, int], float]:
    """Microseconds per in-solve apply for every admissible (tile height, slab)."""
    from .device import time_kernels
    g = torch.Generator(device=op.device).manual_seed(0)
    b = torch.randn(op.n3, dtype=torch.complex128, device=op.device, generator=g)
    out = {}
    for ty, slab in candidates(op.shape):
        kt = time_kernels(op, b, mode, reps=reps, slab=slab, tile_rows=ty)
        out[(ty, slab)] = kt['apply']
    return out


def tuned(op, mode: int) -> tuple[int, int]:
    """Best measured (tile height, slab) for this operator's shape, precision and
    material kinds; (0, 0) means the library's built-in rule."""
    if os.environ.get('MEANAS_AMD_AUTOTUNE', '1') == '0':
        return 0, 0
    key = _key(op, mode)
    with _lock:
        _load()
        if key in _memo:
            return _memo[key]
        times = measure(op, mode)
        if not times:
            return 0, 0
        best = choose(times)
        _memo[key] = best
        _save()
        return best


def tuned_slab(op, mode: int) -> int:
    """The slab depth of `tuned` (its tile height goes with it: use `tuned` for both)."""
    return tuned(op, mode)[1]
