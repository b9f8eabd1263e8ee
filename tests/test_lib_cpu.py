"""CPU checks of the product boundary: the C-ABI library loads and exports exactly the
entry points include/meanas_fdfd.h declares, the ctypes prototypes cover them, the
struct layouts agree with the header, and the host-side PML / coefficient preparation
matches the oracle.  No compute call is made (no GPU here)."""
import ctypes
import os
import re

import numpy
import pytest
from scipy import sparse

from oracle import fdfd_ref as ref

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'meanas_fdfd.h')


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(mx_[a-z_0-9]+)\s*\(', text)))


def test_library_exports_every_declared_symbol():
    from meanas_amd import _lib
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.PROTOTYPES), set(names) ^ set(_lib.PROTOTYPES)
    assert L.mx_abi_version() == 1


def test_struct_sizes_match_header_layout():
    from meanas_amd import _lib
    # mx_operator: 4 int32 + 2 double + 7 pointers ; mx_system: 4 int32 + 2 double + 3 pointers
    assert ctypes.sizeof(_lib.MxOperator) == 16 + 16 + 7 * 8
    assert ctypes.sizeof(_lib.MxSystem) == 16 + 16 + 3 * 8
    assert ctypes.sizeof(_lib.MxOptions) == 6 * 4 + 8 + 4 * 4
    assert ctypes.sizeof(_lib.MxStats) == 16 + 16
    assert ctypes.sizeof(_lib.MxGrid) == 16 + 6 * 8
    assert ctypes.sizeof(_lib.MxFdtdGrid) == 16 + 6 * 8
    assert ctypes.sizeof(_lib.MxCpmlLayer) == 16 + 4 * 8


def test_work_bytes_is_a_pure_host_query():
    from meanas_amd import _lib
    op = _lib.MxOperator(32, 16, 70, 0, 1.0, 0.0, 1, 1, 1, 1, None, None, None)
    n3 = 3 * 32 * 16 * 70
    mixed = _lib.lib().mx_work_bytes(ctypes.byref(op), _lib.MX_MIXED)
    fp64 = _lib.lib().mx_work_bytes(ctypes.byref(op), _lib.MX_FP64)
    assert mixed >= 4 * 8 * n3 and fp64 >= 4 * 16 * n3
    assert _lib.lib().mx_work_bytes(ctypes.byref(op), 7) == 0


@pytest.mark.parametrize('shape', [(1, 2048, 2048), (64, 1024, 1280)])
def test_oversized_cross_section_is_refused_on_the_host(shape):
    """A (y, z) cross-section with more apply tiles than one finishing block reduces is
    refused with a clear error by every entry point whose kernels leave partial sums,
    before any device call (the pointers below are never dereferenced)."""
    from meanas_amd import _lib
    L = _lib.lib()
    fake = 0x1000
    op = _lib.MxOperator(*shape, 0, 1.0, 0.0, fake, fake, fake, fake, None, None, None)
    assert L.mx_work_bytes(ctypes.byref(op), _lib.MX_FP64) > 0
    norm = ctypes.c_double(0.0)
    assert L.mx_residual(ctypes.byref(op), fake, fake, fake, ctypes.byref(norm), None) != 0
    assert b'apply tiles' in L.mx_last_error(), L.mx_last_error()
    assert L.mx_residual_noalloc(ctypes.byref(op), fake, fake, fake, ctypes.byref(norm), None) != 0
    assert b'apply tiles' in L.mx_last_error(), L.mx_last_error()
    sysd = _lib.MxSystem(0, 10, 0, 0, 1e-8, 0.0, fake, fake, fake)
    opts = _lib.MxOptions(_lib.MX_FP64, 0, 0, 0, 0, 1, 0.0, 0, 0, 0, 0)
    stats = (_lib.MxStats * 1)()
    assert L.mx_solve(ctypes.byref(op), 1, ctypes.byref(sysd), 1, ctypes.byref(opts), stats, None) != 0
    assert b'apply tiles' in L.mx_last_error(), L.mx_last_error()


def test_oversized_grid_is_refused_on_the_host():
    """A grid whose complex128 field reaches 4 GiB (448^3 cells) is refused before any
    device call by every entry point that runs the operator apply: the kernels address a
    field component with 32-bit byte offsets, and their bounded output stores assume a
    component below 2 GiB (the pointers below are never dereferenced)."""
    from meanas_amd import _lib
    L = _lib.lib()
    fake = 0x1000
    op = _lib.MxOperator(448, 448, 448, 0, 1.0, 0.0, fake, fake, fake, fake, None, None, None)
    norm = ctypes.c_double(0.0)
    assert L.mx_residual(ctypes.byref(op), fake, fake, fake, ctypes.byref(norm), None) != 0
    assert b'too large' in L.mx_last_error(), L.mx_last_error()
    assert L.mx_residual_noalloc(ctypes.byref(op), fake, fake, fake, ctypes.byref(norm), None) != 0
    assert b'too large' in L.mx_last_error(), L.mx_last_error()
    assert not L.mx_plan_create(ctypes.byref(op), _lib.MX_FP64, 4, 16)
    assert b'too large' in L.mx_last_error(), L.mx_last_error()
    sysd = _lib.MxSystem(0, 10, 0, 0, 1e-8, 0.0, fake, fake, fake)
    opts = _lib.MxOptions(_lib.MX_FP64, 0, 0, 0, 0, 1, 0.0, 0, 0, 0, 0)
    stats = (_lib.MxStats * 1)()
    assert L.mx_solve(ctypes.byref(op), 1, ctypes.byref(sysd), 1, ctypes.byref(opts), stats, None) != 0
    assert b'too large' in L.mx_last_error(), L.mx_last_error()


def test_conducting_boundary_checks_arguments_on_the_host():
    """mx_fdtd_conducting_boundary refuses bad arguments before any device call (each
    call below fails a check; the buffer is never handed to a kernel)."""
    from meanas_amd import _lib
    L = _lib.lib()
    buf = numpy.zeros(3 * 4 * 4 * 2)
    ptr = buf.ctypes.data
    grid = _lib.MxFdtdGrid(4, 4, 2, 0, (ctypes.c_void_p * 3)(), (ctypes.c_void_p * 3)())
    cases = [((_lib.MX_REAL64, 0, -1, _lib.MX_FIELD_E, None), b'null field'),
             ((5, 0, -1, _lib.MX_FIELD_E, ptr), b'prec'),
             ((_lib.MX_REAL64, 3, -1, _lib.MX_FIELD_E, ptr), b'direction'),
             ((_lib.MX_REAL64, 0, 0, _lib.MX_FIELD_E, ptr), b'polarity'),
             ((_lib.MX_REAL64, 0, 1, 2, ptr), b'field'),
             ((_lib.MX_REAL64, 2, 1, _lib.MX_FIELD_H, ptr), b'at least')]   # nz = 2 < 3
    for args, msg in cases:
        assert L.mx_fdtd_conducting_boundary(ctypes.byref(grid), *args, None) != 0
        assert msg in L.mx_last_error(), (args, L.mx_last_error())


def test_energy_entry_points_check_arguments_on_the_host():
    """mx_fdtd_energy and the Poynting entry points refuse bad arguments before any
    device call (each call below fails a check; no buffer reaches a kernel)."""
    from meanas_amd import _lib
    L = _lib.lib()
    buf = numpy.zeros(3 * 4 * 4 * 2)
    out = numpy.zeros(4 * 4 * 2)
    grid = _lib.MxFdtdGrid(4, 4, 2, 0, (ctypes.c_void_p * 3)(), (ctypes.c_void_p * 3)())
    p = buf.ctypes.data
    table = (ctypes.c_void_p * 4)(p, p, None, None)

    def energy(form, tab=table, dt=0.5, o=out.ctypes.data):
        return L.mx_fdtd_energy(ctypes.byref(grid), _lib.MX_REAL64, form, dt, tab, None, 1.0, None, 1.0, o, None)
    for rc_args, msg in [((9,), b'unknown form'), ((_lib.MX_ENERGY_HSTEP,), b'null field for this form'),
                         ((_lib.MX_ENERGY_DXMUL, None), b'null field table'),
                         ((_lib.MX_ENERGY_DXMUL, table, 0.5, None), b'null field')]:
        assert energy(*rc_args) != 0
        assert msg in L.mx_last_error(), (rc_args, L.mx_last_error())
    full = (ctypes.c_void_p * 4)(p, p, p, p)
    assert energy(_lib.MX_ENERGY_H2E, full, 0.0) != 0 and b'dt' in L.mx_last_error()
    assert L.mx_fdtd_poynting(ctypes.byref(grid), _lib.MX_REAL64, p, p, None, None) != 0
    assert b'null output' in L.mx_last_error()
    assert L.mx_fdtd_poynting_divergence(ctypes.byref(grid), 7, p, p, None) != 0
    assert b'prec' in L.mx_last_error()


def test_product_scpml_matches_oracle():
    from meanas_amd.fdfd import scpml
    shape = (30, 1, 41)
    omega = 2 * numpy.pi / 38.75
    mine = [[numpy.linspace(1, 1.2, n) for n in shape] for _ in range(2)]
    theirs = [[d.copy() for d in row] for row in mine]
    for a in (0, 2):
        for p in (-1, 1):
            scpml.stretch_with_scpml(mine, axis=a, polarity=p, omega=omega, epsilon_effective=2.0, thickness=7)
            ref.stretch_with_scpml(theirs, axis=a, polarity=p, omega=omega, epsilon_effective=2.0, thickness=7)
    for eh in (0, 1):
        for a in range(3):
            numpy.testing.assert_allclose(mine[eh][a], theirs[eh][a], rtol=1e-15)


def test_coefficient_table_reproduces_scaled_operator():
    """The per-axis table (fa, fb) builds C^T M C - w^2 eps equal to Pl e_full Pr (oracle)."""
    from scipy import sparse
    from meanas_amd.fdfd.device import coefficient_table
    g = numpy.random.default_rng(0)
    shape = (4, 5, 3)
    dxes = [[g.uniform(0.7, 1.4, n) + 1j * g.uniform(-0.3, 0.3, n) for n in shape] for _ in range(2)]
    tab = coefficient_table(dxes)
    n3 = 3 * int(numpy.prod(shape))
    eps = g.uniform(1, 3, n3) + 1j * g.uniform(-0.2, 0.2, n3)
    mu = g.uniform(0.8, 1.5, n3)
    w = 0.9 - 0.05j
    offs = numpy.cumsum([0, *shape])
    D = []
    for a in range(3):
        fa, fb = tab[offs[a]:offs[a + 1], 0], tab[offs[a]:offs[a + 1], 1]
        S = ref.shift_circ(a, shape, 1)
        grid = lambda v: numpy.meshgrid(*[v if b == a else numpy.ones(n) for b, n in enumerate(shape)],
                                        indexing='ij')[a].ravel()
        D.append(sparse.diags_array(grid(fa)) @ S - sparse.diags_array(grid(fb)))
    C = ref.cross(D)
    A = C.T @ sparse.diags_array(1 / mu) @ C - w ** 2 * sparse.diags_array(eps)
    Pl, Pr = ref.e_full_preconditioners(dxes)
    want = Pl @ ref.e_full(w, dxes, eps, mu=mu) @ Pr
    assert abs(A - want).max() <= 1e-14 * abs(want).max()


def test_generic_rejects_callable_matrix_solver():
    from meanas_amd.fdfd import solvers
    dxes = [[numpy.ones(3)] * 3] * 2
    with pytest.raises(TypeError):
        solvers.generic(1.0, dxes, numpy.zeros(81), numpy.ones(81), matrix_solver=lambda A, b: b)


@pytest.mark.parametrize('periodic', [False, True])
@pytest.mark.parametrize('weighted', [False, True])
def test_boundary_mask_matches_oracle(periodic, weighted):
    """Host edge mask of e_boundary_source (fdfd/operators.py:420-445) against the oracle's
    sparse construction: diag(mask) @ e_tfsf_source equals e_boundary_source."""
    from meanas_amd.fdfd.sources import boundary_mask
    rng = numpy.random.default_rng(3)
    shape = (6, 5, 1)
    dxes = [[1.0 + 0.2 * rng.random(n) for n in shape] for _ in range(2)]
    eps = 1.0 + rng.random(3 * 30)
    m = (rng.random(3 * 30) < 0.4).astype(float)
    if weighted:   # a fractional mask: edges are where the shifted value differs
        m = m * rng.choice([0.5, 1.0], size=m.size)
    keep = boundary_mask(m, shape, periodic)
    full = ref.e_tfsf_source(m, 0.3, dxes, eps)
    want = ref.e_boundary_source(m, 0.3, dxes, eps, periodic_mask_edges=periodic)
    got = sparse.diags_array(keep) @ full
    assert abs(got - want).max() == 0


def test_autotuner_candidates_respect_the_partial_limit():
    """The tuner's (tile height, slab) candidates stay within the per-system partial-sum
    limit and the grid depth; a cross-section beyond the limit has none."""
    from meanas_amd.fdfd import autotune
    c = autotune.candidates((128, 128, 128))
    assert (4, 16) in c and (4, 11) in c and (2, 32) in c
    assert all(autotune.blocks((128, 128, 128), ty, s) <= autotune.MAX_PARTIALS for ty, s in c)
    assert all(s <= 8 or s == autotune.SLABS[0] for _, s in autotune.candidates((8, 48, 48)))
    assert autotune.candidates((1, 2048, 2048)) == []


def test_autotuner_resolves_near_ties_by_a_fixed_preference():
    """Shapes within 3% of the fastest are tied; the tie goes to 4-row tiles and the
    deepest slab, so timing noise does not change a system's partition between boxes."""
    from meanas_amd.fdfd import autotune
    t = {(4, 16): 42.8, (4, 11): 42.3, (2, 16): 43.9, (4, 8): 48.0}
    assert autotune.choose(t) == (4, 16)
    t[(4, 11)] = 40.0
    assert autotune.choose(t) == (4, 11)
    assert autotune.choose({(2, 8): 10.0, (8, 32): 10.1}) == (8, 32)


def test_admission_footprint_counts_materials_and_mask():
    """The server's host-side footprint (computed before anything is allocated) counts the
    workspace (mx_work_bytes), b, x and the source upload, and grows with complex and
    anisotropic eps, 1/mu and the PEC/PMC mask, in both precisions."""
    from meanas_amd import _lib
    from meanas_amd.fdfd.device import device_bytes, operator_flags
    shape = (6, 5, 4)
    n = 6 * 5 * 4
    iso = numpy.full(3 * n, 2.0)
    aniso = numpy.concatenate([numpy.full(n, 2.0), numpy.full(n, 3.0), numpy.full(n, 4.0)])
    assert operator_flags(shape, iso) == 0
    assert operator_flags(shape, aniso) == _lib.MX_EPS_ANISO
    assert operator_flags(shape, iso + 0.1j) == _lib.MX_EPS_COMPLEX
    pec = numpy.zeros(3 * n)
    assert operator_flags(shape, iso, pec=pec) == 0          # an all-zero mask is no mask
    pec[7] = 1
    assert operator_flags(shape, iso, mu=iso, pec=pec) == _lib.MX_MU | _lib.MX_MASK
    for mode, real in ((_lib.MX_FP64, 8), (_lib.MX_MIXED, 4)):
        base = device_bytes(shape, 0, mode)
        desc = _lib.MxOperator(*shape, 0, 0.0, 0.0, None, None, None, None, None, None, None)
        work = _lib.lib().mx_work_bytes(ctypes.byref(desc), mode)
        assert base == sum(shape) * 4 * 24 + n * 12 + 3 * 16 * 3 * n + work
        assert device_bytes(shape, _lib.MX_EPS_ANISO, mode) - base == 2 * n * 12
        assert device_bytes(shape, _lib.MX_EPS_COMPLEX, mode) - base == n * 12
        assert device_bytes(shape, _lib.MX_MU, mode) - base == n * 12
        assert device_bytes(shape, _lib.MX_MASK, mode) - base == n
        assert work >= 4 * real * 2 * 3 * n


def test_server_phase_counts_across_hand_overs():
    """A request the server continues in another lane reports the sum of its phases'
    iterations and replacements, and which phases ran: complex64 -> fused FP64 (switched)
    -> classic FP64 gives fp64_fallback, mixed_iterations and fused_iterations; a request
    that stays in one lane reports its own counts unchanged."""
    from meanas_amd import _lib
    from meanas_amd.serve import hand_over_prior, merge_prior
    mixed, fused, classic = (_lib.MX_MIXED, 'fused'), (_lib.MX_FP64, 'fused'), (_lib.MX_FP64, 'classic')
    ph = [dict(iterations=70, replacements=2, flags=0), dict(iterations=300, replacements=4, flags=0),
          dict(iterations=55, replacements=1, flags=0)]
    p = hand_over_prior(None, ph[0], mixed, 'fused')
    assert p == dict(iterations=70, replacements=2, fp64_fallback=True, mixed_iterations=70)
    p = hand_over_prior(p, ph[1], fused, 'classic')
    assert p == dict(iterations=370, replacements=6, fp64_fallback=True, mixed_iterations=70,
                     fused_iterations=300)
    st = merge_prior(dict(ph[2]), p, classic)
    assert (st['iterations'], st['replacements']) == (425, 7)
    assert (st['fp64_fallback'], st['mixed_iterations'], st['fused_iterations']) == (True, 70, 300)
    # complex64 to the end, and FP64 fused to the end
    st = merge_prior(dict(ph[0]), None, mixed)
    assert st == dict(ph[0], fp64_fallback=False, mixed_iterations=70)
    assert merge_prior(dict(ph[1]), None, fused) == ph[1]
    # FP64 fused switched to classic without a complex64 phase
    st = merge_prior(dict(ph[2]), hand_over_prior(None, ph[1], fused, 'classic'), classic)
    assert st == dict(ph[2], iterations=355, replacements=5, fused_iterations=300)
    # a classic FP64 continuation of a complex64 phase (fused form disabled)
    st = merge_prior(dict(ph[2]), hand_over_prior(None, ph[0], mixed, 'classic'), classic)
    assert st == dict(ph[2], iterations=125, replacements=3, fp64_fallback=True, mixed_iterations=70)


def test_bench_refuses_several_gpus():
    """bench.py measures one GPU: `--gpus N` with N > 1 (or a WORLD_SIZE above 1) exits
    with status 2 and says so, before touching a device."""
    import subprocess
    import sys
    env = dict(os.environ)
    for argv, extra in ((['--gpus', '2'], {}), ([], {'WORLD_SIZE': '4'})):
        env2 = dict(env, **extra)
        res = subprocess.run([sys.executable, os.path.join(ROOT, 'bench.py'), *argv], capture_output=True, text=True,
                             env=env2, timeout=120)
        assert res.returncode == 2, res.stderr
        assert 'one GPU only' in res.stdout
