"""Time-domain (FDTD) Yee updates on the GPU: the meanas.fdtd update callables, the
convolutional PML and the conducting boundary (meanas/fdtd/base.py, meanas/fdtd/pml.py,
meanas/fdtd/boundaries.py)."""
from .base import maxwell_e, maxwell_h, run
from .boundaries import conducting_boundary
from .pml import cpml_params, updates_with_cpml

__all__ = ['maxwell_e', 'maxwell_h', 'run', 'cpml_params', 'updates_with_cpml', 'conducting_boundary']
