"""GPU parity of the non-periodic-y transforms (FC-Gram continuation along y
as well as z; fftp.fpp:388-425 with Cy > 0, fftp.fpp:527-716 and :922-1056)
against the float64 oracle, through the C ABI (specter_amd.FCPlan with Cy > 0).

Random fields continue into values up to ~1e7 times the input in the (y, z)
corner, so errors are relative to each result's maximum (rel_err), where
float64 rounding of the continued field sits near 1e-16 * max.
"""
import numpy as np
import pytest

from conftest import TABLES, rel_err

pytestmark = pytest.mark.gpu

# (nx, npy, Cy, oy, npz, Cz, oz): ny = npy + Cy must be even
SHAPES = [
    (16, 39, 25, 5, 39, 25, 5),     # ny = 64 (radix 2), nz = 64
    (12, 41, 25, 5, 32, 25, 5),     # ny = 66 = 2*3*11, nz = 57 = 3*19
    (8, 37, 27, 8, 16, 25, 5),      # oy = 8: ny = 64, nz = 41 (Bluestein)
    (8, 487, 25, 5, 16, 25, 5),     # ny = 512 (the 512-point wave plan)
]
TOL = 1e-13


def _ids(s):
    return "nx%d_ny%d+%d_o%d_nz%d+%d_o%d" % s


def _setup(shape):
    import torch
    from oracle import specter_ref as O
    from specter_amd import FCPlan
    nx, npy, Cy, oy, npz, Cz, oz = shape
    ny, nz = npy + Cy, npz + Cz
    g = O.make_grid(nx, ny, nz, Cz, oz, TABLES, Cy=Cy, oy=oy)
    p = FCPlan(nx, ny, nz, Cz, oz, Cy=Cy, oy=oy)
    return torch, O, g, p


def _cu(torch, a):
    return torch.as_tensor(np.ascontiguousarray(a)).cuda()


def _rand_real(g, seed):
    return np.random.default_rng(seed).standard_normal(g.real_shape)


def _rand_spec(g, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(g.spec_shape) + 1j * rng.standard_normal(g.spec_shape)


@pytest.mark.parametrize("shape", SHAPES, ids=_ids)
def test_r2c_3d(shape):
    torch, O, g, p = _setup(shape)
    R = _rand_real(g, 1)
    ref = O.fftp3d_real_to_complex(g, R.copy())
    out = p.empty_spectral()
    p.fftp3d_real_to_complex(_cu(torch, R), out)
    assert rel_err(out.cpu().numpy(), ref) < TOL


@pytest.mark.parametrize("shape", SHAPES, ids=_ids)
def test_c2r_3d(shape):
    torch, O, g, p = _setup(shape)
    S = _rand_spec(g, 2)
    ref = O.fftp3d_complex_to_real(g, S.copy())
    out = p.empty_real()
    p.fftp3d_complex_to_real(_cu(torch, S), out)
    assert rel_err(out.cpu().numpy(), ref) < TOL


@pytest.mark.parametrize("shape", SHAPES[:3], ids=_ids)
def test_pieces(shape):
    """fftp1d_real_to_complex_x, fftp2d_real_to_complex_yz,
    fftp2d_complex_to_real_yz, fftp1d_complex_to_real_x against the oracle's."""
    torch, O, g, p = _setup(shape)
    R = _rand_real(g, 3)
    X = O.fftp1d_real_to_complex_x(g, R.copy())
    out = p.empty_spectral()
    p.fftp1d_real_to_complex_x(_cu(torch, R), out)
    assert rel_err(out.cpu().numpy(), X) < TOL
    O.fftp2d_real_to_complex_yz(g, X)
    p.fftp2d_real_to_complex_yz(out)
    assert rel_err(out.cpu().numpy(), X) < TOL
    S = _rand_spec(g, 4)
    ref = S.copy()
    O.fftp2d_complex_to_real_yz(g, ref)
    t = _cu(torch, S)
    p.fftp2d_complex_to_real_yz(t)
    assert rel_err(t.cpu().numpy(), ref) < TOL
    rr = O.fftp1d_complex_to_real_x(g, S.copy())
    o = p.empty_real()
    p.fftp1d_complex_to_real_x(_cu(torch, S), o)
    assert rel_err(o.cpu().numpy(), rr) < TOL


@pytest.mark.parametrize("shape", SHAPES[:2], ids=_ids)
def test_derivatives_kat(shape):
    """Spectral derivatives of sin(4x) cos(8y) sin(6z) on the physical points
    (the reference's fc_dirichlet check, tests/fc_dirichlet.f90, with y
    non-periodic): x exact to rounding, y and z to the continuation's accuracy
    at about 40 physical points (the oracle's own errors, times 2)."""
    torch, O, g, p = _setup(shape)
    z, y, x = np.meshgrid(g.z, g.y, g.x, indexing="ij")
    f = np.sin(4 * x) * np.cos(8 * y) * np.sin(6 * z)
    exact = {1: 4 * np.cos(4 * x) * np.cos(8 * y) * np.sin(6 * z),
             2: -8 * np.sin(4 * x) * np.sin(8 * y) * np.sin(6 * z),
             3: 6 * np.sin(4 * x) * np.cos(8 * y) * np.cos(6 * z)}
    S = p.empty_spectral()
    p.fftp3d_real_to_complex(_cu(torch, f), S)
    Sref = O.fftp3d_real_to_complex(g, f.copy())
    N = g.nx * g.ny * g.nz
    sl = (slice(0, g.np_), slice(0, g.npy), slice(None))
    for dirn, ex in exact.items():
        d = p.empty_spectral()
        p.derivk(S, d, dirn)
        r = p.empty_real()
        p.fftp3d_complex_to_real(d, r)
        got = r.cpu().numpy()[sl] / N
        oref = O.fftp3d_complex_to_real(g, O.derivk(g, Sref, dirn))[sl] / N
        e_gpu = np.abs(got - ex[sl]).max() / np.abs(ex[sl]).max()
        e_ora = np.abs(oref - ex[sl]).max() / np.abs(ex[sl]).max()
        assert e_gpu <= 2 * e_ora + 1e-13, (dirn, e_gpu, e_ora)
        if dirn == 1:
            assert e_gpu < 1e-12
        else:
            assert e_gpu < 1e-3


@pytest.mark.parametrize("shape", SHAPES[:3], ids=_ids)
def test_gradre(shape):
    """(A.grad)A with y continued in the three forward transforms."""
    torch, O, g, p = _setup(shape)
    a, b, c = (_rand_spec(g, s) for s in (5, 6, 7))
    ref = O.gradre(g, a, b, c)
    outs = [p.empty_spectral() for _ in range(3)]
    p.gradre(_cu(torch, a), _cu(torch, b), _cu(torch, c), *outs)
    for o, r in zip(outs, ref):
        assert rel_err(o.cpu().numpy(), r) < 1e-12


def test_wall_solvers_refused():
    """No boundary conditions in y (boundary_mod.fpp:104-105): the BC calls
    and the solvers refuse plans with Cy > 0."""
    import ctypes
    from specter_amd import lib as L
    torch, O, g, p = _setup(SHAPES[0])
    v = [p.empty_spectral() for _ in range(4)]
    lib = L.load_library()
    rc = lib.spx_sol_project(p.handle, *(ctypes.c_void_p(t.data_ptr()) for t in v), 1, 0, 0)
    assert rc == L.SPX_ERR_UNSUPPORTED
    rc = lib.spx_s_imposebc(p.handle, ctypes.c_void_p(v[0].data_ptr()))
    assert rc == L.SPX_ERR_UNSUPPORTED
    prm = L.SpxParams()
    prm.solver, prm.ord, prm.dt, prm.nu = L.SPX_SOLVER_HD, 2, 1e-3, 1e-2
    s = ctypes.c_void_p()
    assert lib.spx_state_create(p.handle, ctypes.byref(prm), ctypes.byref(s)) == \
        L.SPX_ERR_UNSUPPORTED
    assert not s.value
