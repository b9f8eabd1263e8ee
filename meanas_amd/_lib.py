"""ctypes binding of libmeanas_fdfd.so (C ABI declared in include/meanas_fdfd.h).

The library is the product path: there is no CPU fallback.  Importing the binding
raises if the built library is missing (build it with `python -m meanas_amd.build`).
"""
from __future__ import annotations

import ctypes
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'lib', 'libmeanas_fdfd.so')

MX_MIXED = 0
MX_FP64 = 1

MX_EPS_COMPLEX = 1
MX_EPS_ANISO = 2
MX_MU = 4
MX_MU_COMPLEX = 8
MX_MU_ANISO = 16
MX_MASK = 32

MX_CONVERGED = 0
MX_MAXITER = 1
MX_BREAKDOWN = 2
MX_NONFINITE = 3
MX_STALLED = 4
MX_FORM_AUTO, MX_FORM_FUSED, MX_FORM_CLASSIC = 0, 1, 2
MX_FLAG_SWITCHED = 8192   # mx_stats.flags: a fused session system that lost track, to continue classic
MX_REAL32, MX_REAL64 = 0, 1
MX_FIELD_E, MX_FIELD_H = 0, 1
MX_ENERGY_DXMUL, MX_ENERGY_HSTEP, MX_ENERGY_ESTEP, MX_ENERGY_H2E, MX_ENERGY_E2H, MX_ENERGY_J = range(6)
FORMS = {'auto': MX_FORM_AUTO, 'fused': MX_FORM_FUSED, 'classic': MX_FORM_CLASSIC}
STATUS_NAMES = {0: 'converged', 1: 'maxiter', 2: 'breakdown', 3: 'nonfinite', 4: 'stalled'}

MX_MAP_RHS = 0
MX_MAP_RHS_ADJ = 1
MX_MAP_E = 2
MX_MAP_E_ADJ = 3
MX_MAP_SIGN = 4

c_void_p = ctypes.c_void_p


class MxOperator(ctypes.Structure):
    _fields_ = [
        ('nx', ctypes.c_int32), ('ny', ctypes.c_int32), ('nz', ctypes.c_int32),
        ('flags', ctypes.c_int32),
        ('omega2_re', ctypes.c_double), ('omega2_im', ctypes.c_double),
        ('coef32', c_void_p), ('coef64', c_void_p),
        ('eps32', c_void_p), ('eps64', c_void_p),
        ('inv_mu32', c_void_p), ('inv_mu64', c_void_p),
        ('mask', c_void_p),
    ]


class MxSystem(ctypes.Structure):
    _fields_ = [
        ('op', ctypes.c_int32), ('maxiter', ctypes.c_int32),
        ('slab', ctypes.c_int32), ('tile_rows', ctypes.c_int32),
        ('rtol', ctypes.c_double), ('atol', ctypes.c_double),
        ('b', c_void_p), ('x', c_void_p), ('work', c_void_p),
    ]


class MxOptions(ctypes.Structure):
    _fields_ = [
        ('mode', ctypes.c_int32), ('tile_rows', ctypes.c_int32), ('slab', ctypes.c_int32),
        ('chunk', ctypes.c_int32), ('check_every', ctypes.c_int32), ('use_graph', ctypes.c_int32),
        ('delta', ctypes.c_double),
        ('stall_iters', ctypes.c_int32), ('stall_limit', ctypes.c_int32),
        ('form', ctypes.c_int32), ('reserved', ctypes.c_int32),
    ]


class MxGrid(ctypes.Structure):
    """mx_grid: cell widths dxes[0] (E grid) and dxes[1] (H grid) as device complex128 arrays."""
    _fields_ = [('nx', ctypes.c_int32), ('ny', ctypes.c_int32), ('nz', ctypes.c_int32), ('pad', ctypes.c_int32),
                ('dxe', c_void_p * 3), ('dxh', c_void_p * 3)]


class MxFdtdGrid(ctypes.Structure):
    """mx_fdtd_grid: real cell widths (float64 device arrays) or NULL for unit widths."""
    _fields_ = [('nx', ctypes.c_int32), ('ny', ctypes.c_int32), ('nz', ctypes.c_int32), ('pad', ctypes.c_int32),
                ('dxe', c_void_p * 3), ('dxh', c_void_p * 3)]


class MxCpmlLayer(ctypes.Structure):
    """mx_cpml_layer: one CPML face (profiles and psi state on the device)."""
    _fields_ = [('axis', ctypes.c_int32), ('polarity', ctypes.c_int32), ('thickness', ctypes.c_int32),
                ('reserved', ctypes.c_int32), ('p_e', c_void_p), ('p_h', c_void_p), ('psi_e', c_void_p),
                ('psi_h', c_void_p)]


MX_CPML_MAX_LAYERS = 6


class MxStats(ctypes.Structure):
    _fields_ = [
        ('status', ctypes.c_int32), ('iterations', ctypes.c_int32),
        ('replacements', ctypes.c_int32), ('flags', ctypes.c_int32),
        ('bnorm', ctypes.c_double), ('rel_residual', ctypes.c_double),
    ]

    def as_dict(self) -> dict:
        return dict(status=STATUS_NAMES.get(self.status, str(self.status)), iterations=self.iterations,
                    replacements=self.replacements, flags=self.flags, bnorm=self.bnorm,
                    rel_residual=self.rel_residual)


# exported symbols and their prototypes: (restype, argtypes)
PROTOTYPES = {
    'mx_abi_version': (ctypes.c_int, []),
    'mx_last_error': (ctypes.c_char_p, []),
    'mx_work_bytes': (ctypes.c_size_t, [ctypes.POINTER(MxOperator), ctypes.c_int32]),
    'mx_apply': (ctypes.c_int, [ctypes.POINTER(MxOperator), ctypes.c_int32, c_void_p, c_void_p, c_void_p]),
    'mx_plan_create': (c_void_p, [ctypes.POINTER(MxOperator), ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    'mx_plan_apply': (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    'mx_plan_blocks': (ctypes.c_int, [c_void_p]),
    'mx_plan_destroy': (None, [c_void_p]),
    'mx_residual': (ctypes.c_int, [ctypes.POINTER(MxOperator), c_void_p, c_void_p, c_void_p,
                                   ctypes.POINTER(ctypes.c_double), c_void_p]),
    'mx_residual_noalloc': (ctypes.c_int, [ctypes.POINTER(MxOperator), c_void_p, c_void_p, c_void_p,
                                           ctypes.POINTER(ctypes.c_double), c_void_p]),
    'mx_map': (ctypes.c_int, [ctypes.POINTER(MxOperator), ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                              c_void_p, c_void_p, c_void_p]),
    'mx_solve': (ctypes.c_int, [ctypes.POINTER(MxOperator), ctypes.c_int32, ctypes.POINTER(MxSystem),
                                ctypes.c_int32, ctypes.POINTER(MxOptions), ctypes.POINTER(MxStats), c_void_p]),
    'mx_solve_plan_reuses': (ctypes.c_int64, []),
    'mx_solve_plan_release': (ctypes.c_int, []),
    'mx_m2j': (ctypes.c_int, [ctypes.POINTER(MxGrid), ctypes.c_double, ctypes.c_double, c_void_p, c_void_p, c_void_p,
                              c_void_p]),
    'mx_e2h': (ctypes.c_int, [ctypes.POINTER(MxGrid), ctypes.c_double, ctypes.c_double, c_void_p, c_void_p,
                              c_void_p, c_void_p, c_void_p]),
    'mx_poynting_h_cross_e': (ctypes.c_int, [ctypes.POINTER(MxGrid), c_void_p, c_void_p, ctypes.c_int32, c_void_p,
                                             c_void_p]),
    'mx_poynting_e_cross_h': (ctypes.c_int, [ctypes.POINTER(MxGrid), c_void_p, c_void_p, ctypes.c_int32, c_void_p,
                                             c_void_p]),
    'mx_fdtd_h': (ctypes.c_int, [ctypes.POINTER(MxFdtdGrid), ctypes.c_int32, ctypes.c_double, c_void_p, c_void_p,
                                 c_void_p, ctypes.c_double, c_void_p]),
    'mx_fdtd_e': (ctypes.c_int, [ctypes.POINTER(MxFdtdGrid), ctypes.c_int32, ctypes.c_double, c_void_p, c_void_p,
                                 c_void_p, ctypes.c_double, c_void_p]),
    'mx_fdtd_run': (ctypes.c_int, [ctypes.POINTER(MxFdtdGrid), ctypes.c_int32, ctypes.c_double, c_void_p, c_void_p,
                                   c_void_p, ctypes.c_double, c_void_p, ctypes.c_double, ctypes.c_int32, c_void_p]),
    'mx_fdtd_cpml_e': (ctypes.c_int, [ctypes.POINTER(MxFdtdGrid), ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                      c_void_p, c_void_p, c_void_p, ctypes.c_double, ctypes.POINTER(MxCpmlLayer),
                                      ctypes.c_int32, c_void_p]),
    'mx_fdtd_cpml_h': (ctypes.c_int, [ctypes.POINTER(MxFdtdGrid), ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                      c_void_p, c_void_p, c_void_p, ctypes.c_double, ctypes.POINTER(MxCpmlLayer),
                                      ctypes.c_int32, c_void_p]),
    'mx_fdtd_conducting_boundary': (ctypes.c_int, [ctypes.POINTER(MxFdtdGrid), ctypes.c_int32, ctypes.c_int32,
                                                   ctypes.c_int32, ctypes.c_int32, c_void_p, c_void_p]),
    'mx_fdtd_energy': (ctypes.c_int, [ctypes.POINTER(MxFdtdGrid), ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                      ctypes.POINTER(c_void_p), c_void_p, ctypes.c_double, c_void_p, ctypes.c_double,
                                      c_void_p, c_void_p]),
    'mx_fdtd_poynting': (ctypes.c_int, [ctypes.POINTER(MxFdtdGrid), ctypes.c_int32, c_void_p, c_void_p, c_void_p,
                                        c_void_p]),
    'mx_fdtd_poynting_divergence': (ctypes.c_int, [ctypes.POINTER(MxFdtdGrid), ctypes.c_int32, c_void_p, c_void_p,
                                                   c_void_p]),
    'mx_time_kernels': (ctypes.c_int, [ctypes.POINTER(MxOperator), ctypes.POINTER(MxSystem), ctypes.POINTER(MxOptions),
                                       ctypes.c_int32, ctypes.POINTER(ctypes.c_double), ctypes.c_int32, c_void_p]),
    'mx_session_create': (c_void_p, [ctypes.c_int32, ctypes.POINTER(MxOptions)]),
    'mx_session_admit': (ctypes.c_int, [c_void_p, ctypes.POINTER(MxOperator), ctypes.POINTER(MxSystem), c_void_p]),
    'mx_session_step': (ctypes.c_int, [c_void_p, ctypes.c_int32, c_void_p]),
    'mx_session_poll': (ctypes.c_int, [c_void_p, ctypes.POINTER(MxStats), ctypes.POINTER(ctypes.c_int32), c_void_p]),
    'mx_session_retire': (ctypes.c_int, [c_void_p, ctypes.c_int32]),
    'mx_session_destroy': (None, [c_void_p]),
}

_lib = None


class MxError(RuntimeError):
    pass


def lib() -> ctypes.CDLL:
    """Load libmeanas_fdfd.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MxError(f'{LIB_PATH} is missing: the HIP library must be built '
                          '(python -m meanas_amd.build); there is no CPU fallback')
        h = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().mx_last_error().decode(errors='replace')
        raise MxError(f'{what} failed: {msg}')
