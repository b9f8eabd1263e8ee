"""Time-domain (FDTD) Yee updates on the GPU: the meanas.fdtd update callables, the
convolutional PML, the conducting boundary and the energy functions
(meanas/fdtd/base.py, pml.py, boundaries.py, energy.py)."""
from .base import maxwell_e, maxwell_h, run
from .boundaries import conducting_boundary
from .energy import (delta_energy_e2h, delta_energy_h2e, delta_energy_j, dxmul, energy_estep, energy_hstep, poynting,
                     poynting_divergence)
from .pml import cpml_params, updates_with_cpml

__all__ = ['maxwell_e', 'maxwell_h', 'run', 'cpml_params', 'updates_with_cpml', 'conducting_boundary', 'poynting',
           'poynting_divergence', 'energy_hstep', 'energy_estep', 'delta_energy_h2e', 'delta_energy_e2h',
           'delta_energy_j', 'dxmul']
