"""Time-domain (FDTD) Yee updates on the GPU: the meanas.fdtd update callables."""
from .base import maxwell_e, maxwell_h, run

__all__ = ['maxwell_e', 'maxwell_h', 'run']
