"""Writes the golden fixtures under tests/golden/ with the FP64 CPU oracle.

plain_48.npz: the oracle's solution (scipy qmr on the restated CSR operator,
rtol 1e-12) of the 48^3 plain FDFD workload (meanas_amd.workloads.plain_48), stored
as complex64 (enough to check a relative field error of 1e-6) plus its FP64 norm and
the oracle's achieved residuals.

Run:  python tests/golden/make_golden.py
"""
import os
import sys
import time

import numpy

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import fdfd_ref as ref           # noqa: E402
from meanas_amd import workloads             # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def plain_48(rtol=1e-12):
    w = workloads.plain_48()
    t0 = time.time()
    opts = dict(rtol=rtol, atol=0.0)
    E = ref.generic(w['omega'], w['dxes'], w['J'], w['epsilon'], matrix_solver_opts=opts)
    secs = time.time() - t0
    A0 = ref.e_full(w['omega'], w['dxes'], w['epsilon'])
    b0 = -1j * w['omega'] * w['J']
    phys = numpy.linalg.norm(A0 @ E - b0) / numpy.linalg.norm(b0)
    Pl, Pr = ref.e_full_preconditioners(w['dxes'])
    A = Pl @ A0 @ Pr
    b = Pl @ b0
    scaled = numpy.linalg.norm(A @ (Pl @ E) - b) / numpy.linalg.norm(b)
    numpy.savez(os.path.join(HERE, 'plain_48.npz'), E=E.astype(numpy.complex64),
                E_norm=numpy.linalg.norm(E), physical_residual=phys, scaled_residual=scaled,
                qmr_iterations=ref.scipy_qmr.last_iterations, qmr_info=ref.scipy_qmr.last_info,
                seconds=secs, rtol=rtol)
    print(f'plain_48: qmr {ref.scipy_qmr.last_iterations} its, info {ref.scipy_qmr.last_info}, '
          f'{secs:.1f} s, scaled residual {scaled:.3e}, physical {phys:.3e}')


if __name__ == '__main__':
    plain_48()
