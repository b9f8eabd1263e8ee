"""Time-domain (FDTD) Yee updates on the GPU: the meanas.fdtd update callables and the
convolutional PML (meanas/fdtd/base.py, meanas/fdtd/pml.py)."""
from .base import maxwell_e, maxwell_h, run
from .pml import cpml_params, updates_with_cpml

__all__ = ['maxwell_e', 'maxwell_h', 'run', 'cpml_params', 'updates_with_cpml']
