"""MI355X (gfx950) implementation of meanas' 3D FDFD solve path.

`meanas_amd.fdfd.solvers.generic` is a drop-in for meanas.fdfd.solvers.generic; it
runs hand-written HIP kernels (libmeanas_fdfd.so, C ABI in include/meanas_fdfd.h)
through ctypes, with torch tensors as device buffers.
"""
__version__ = '0.1.0'
