"""
FP64 CPU oracle for the 3D FDFD solve path -- TEST INFRASTRUCTURE ONLY.

This package restates, in numpy/scipy, the algorithm of meanas' FDFD solve
path so that the HIP implementation in `meanas_amd` can be checked against it.
Only `tests/`, `__graft_entry__.smoke()` and the `cpu_baseline` leg of
`bench.py` may import it, and there only as the checker / the timed CPU
baseline.  The product path (`meanas_amd`) never imports it and has no CPU
fallback.

Every function cites the reference file:line it restates.  How each function is
pinned (golden vectors, closed forms, invariants, brute force) is listed in
`oracle/fdfd_ref.py`'s module docstring.
"""
