"""Build recipe for the HIP library libmeanas_fdfd.so (gfx950 only).

Compiles every `csrc/*.hip` translation unit with hipcc for gfx950 and links them into
`meanas_amd/lib/libmeanas_fdfd.so`, in-tree, so the built library travels with the
source tree.  Packed-FP32 instruction selection is disabled for device code.
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
OUT_DIR = os.path.join(HERE, 'lib')
LIB_NAME = 'libmeanas_fdfd.so'
ARCH = 'gfx950'

HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
FLAGS = [
    '-O3', '-std=c++17', '-fPIC', f'--offload-arch={ARCH}',
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wno-unused-command-line-argument',
]


def sources() -> list[str]:
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith('.hip'))


def _compile(src: str, obj: str) -> None:
    cmd = [HIPCC, *FLAGS, '-c', src, '-o', obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f'hipcc failed for {src}:\n{res.stderr}')


def build(verbose: bool = True) -> str:
    """Compile and link the library; returns the path of the built .so."""
    os.makedirs(OUT_DIR, exist_ok=True)
    objs = []
    jobs = []
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        for src in sources():
            obj = os.path.join(OUT_DIR, os.path.basename(src)[:-4] + '.o')
            objs.append(obj)
            jobs.append(ex.submit(_compile, src, obj))
        for j in jobs:
            j.result()
    out = os.path.join(OUT_DIR, LIB_NAME)
    tmp = out + '.tmp'
    cmd = [HIPCC, *FLAGS, '-shared', *objs, '-o', tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f'link failed:\n{res.stderr}')
    os.replace(tmp, out)
    for o in objs:
        os.remove(o)
    if verbose:
        print(f'built {out}', file=sys.stderr)
    return out


if __name__ == '__main__':
    build()
