"""FDFD solve path: host-side PML preparation, device operators and the generic() drop-in."""
from . import scpml  # noqa: F401
