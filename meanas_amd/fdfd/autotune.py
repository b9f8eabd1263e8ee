"""Autotuner for the apply kernel's plane-block depth (x-planes per block).

Each block of the apply kernel owns a 4 x 64 (y, z) tile and marches over `slab`
x-planes; a deeper slab amortises the two extra planes a block reads at its start,
a shallower one gives more blocks to fill the 256 CUs.  The best depth depends on the
grid shape and the precision, so it is measured: every candidate depth is timed with
an apply plan on the real operator (HIP events on one stream) and the fastest is kept,
per (device, shape, precision), in memory and in a JSON cache file.

The tile height stays fixed (4 rows): a system's block partition -- and so the order
of its reductions -- must depend on that system alone, never on the rest of a batch.
Set MEANAS_AMD_AUTOTUNE=0 to use the library's built-in depth rule instead.
"""
from __future__ import annotations

import json
import os
import threading

import torch

from .. import _lib

CANDIDATES = (2, 4, 8, 16, 32)
TILE_ROWS = 4
MAX_PARTIALS = 4096       # per-system reduction limit of the finishing kernel
CACHE_PATH = os.path.join(os.path.expanduser('~'), '.cache', 'meanas_amd', 'autotune.json')

_lock = threading.Lock()
_memo: dict[str, int] = {}
_loaded = False


def _blocks(shape, slab: int) -> int:
    nx, ny, nz = shape
    return ((ny + TILE_ROWS - 1) // TILE_ROWS) * ((nz + 63) // 64) * ((nx + slab - 1) // slab)


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
            _memo.update({k: int(v) for k, v in json.load(f).items()})
    except (OSError, ValueError):
        pass


def _save():
    try:
        os.makedirs(os.path.dirname(CACHE_PATH), exist_ok=True)
        tmp = CACHE_PATH + '.tmp'
        with open(tmp, 'w') as f:
            json.dump(_memo, f, indent=0, sort_keys=True)
        os.replace(tmp, CACHE_PATH)
    except OSError:
        pass


def measure(op, mode: int, reps: int = 10) -> dict[int, float]:
    """Microseconds per apply for every admissible candidate depth."""
    dtype = torch.complex64 if mode == _lib.MX_MIXED else torch.complex128
    x = torch.randn(op.n3, dtype=dtype, device=op.device)
    y = torch.empty_like(x)
    stream = torch.cuda.current_stream(op.device)
    out = {}
    for slab in CANDIDATES:
        if slab > op.shape[0] and slab != CANDIDATES[0]:
            continue
        if _blocks(op.shape, slab) > MAX_PARTIALS:
            continue
        plan = op.plan('c64' if mode == _lib.MX_MIXED else 'c128', tile_rows=TILE_ROWS, slab=slab)
        for _ in range(2):
            plan.apply(x, y, stream=stream)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            plan.apply(x, y, stream=stream)
        e1.record(stream)
        e1.synchronize()
        out[slab] = e0.elapsed_time(e1) * 1e3 / reps
    return out


def tuned_slab(op, mode: int) -> int:
    """Best measured slab depth for this operator's shape and precision (0 = library rule)."""
    if os.environ.get('MEANAS_AMD_AUTOTUNE', '1') == '0':
        return 0
    key = _key(op, mode)
    with _lock:
        _load()
        if key in _memo:
            return _memo[key]
        times = measure(op, mode)
        if not times:
            return 0
        best = min(times, key=times.get)
        _memo[key] = best
        _save()
        return best
