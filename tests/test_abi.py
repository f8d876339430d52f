"""C ABI checks that need no GPU: the library loads, exports every entry point
declared in include/spx.h, and rejects invalid arguments before touching the
device.  Also host-side logic (tables, initial conditions)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT


@pytest.fixture(scope="module")
def lib():
    from specter_amd import lib as L
    if not os.path.exists(L.LIB_PATH):
        from specter_amd.build import build
        build()
    return L.load_library()


def test_every_header_symbol_is_exported(lib):
    from specter_amd.lib import header_symbols, _SIGS
    syms = header_symbols()
    assert len(syms) >= 30
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(ROOT, "specter_amd",
                                                                      "libspx.so")],
                         capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    for s in syms:
        assert s in exported, s
        assert hasattr(lib, s)
        assert s in _SIGS, f"{s} has no ctypes signature"


def _err(lib):
    buf = ctypes.create_string_buffer(512)
    lib.spx_last_error(buf, 512)
    return buf.value.decode()


@pytest.mark.parametrize("grid,code,msg", [
    ((15, 16, 41, 25, 5, 1.0, 1.0, 1.0, 0), 1, "nx must be even"),
    ((16, 16, 41, 25, 5, 1.0, -1.0, 1.0, 0), 1, "box lengths"),
    ((16, 16, 41, 25, 0, 1.0, 1.0, 1.0, 0), 1, "matching order"),
    ((16, 16, 30, 25, 5, 1.0, 1.0, 1.0, 0), 1, "too few physical"),
    ((16, 16, 41, 25, 5, 1.0, 1.0, 1.0, 7), 1, "precision must be"),
])
def test_plan_create_validates_before_device(lib, grid, code, msg):
    from specter_amd.lib import SpxGrid
    from specter_amd.tables import load_dirichlet_tables
    A, Q = load_dirichlet_tables(25, 5)
    g = SpxGrid(*grid)
    p = ctypes.c_void_p()
    rc = lib.spx_plan_create(ctypes.byref(g), A.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                             Q.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), None,
                             ctypes.byref(p))
    assert rc == code
    assert msg in _err(lib)
    assert not p.value


@pytest.mark.parametrize("Cy,oy,prec,with_y,code,msg", [
    (25, 5, 1, True, 2, "non-periodic y needs SPX_FP64"),
    (25, 5, 0, False, 1, "tables Ay and Qy"),
    (25, 0, 0, True, 1, "matching order oy"),
    (0, 5, 0, False, 1, "Cy == 0 requires oy == 0"),
    (60, 5, 0, True, 1, "too few physical y"),
])
def test_plan_create_fcy_validates_before_device(lib, Cy, oy, prec, with_y, code, msg):
    """spx_plan_create_fcy checks its y arguments before any device call."""
    from specter_amd.lib import SpxGrid
    from specter_amd.tables import load_dirichlet_tables
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
    A, Q = load_dirichlet_tables(25, 5)
    g = SpxGrid(16, 64, 41, 25, 5, 1.0, 1.0, 1.0, prec)
    p = ctypes.c_void_p()
    yA, yQ = (dp(A), dp(Q)) if with_y else (None, None)
    rc = lib.spx_plan_create_fcy(ctypes.byref(g), dp(A), dp(Q), Cy, oy, yA, yQ, 1, None,
                                 ctypes.byref(p))
    assert rc == code
    assert msg in _err(lib)
    assert not p.value


def test_missing_tables_rejected(lib):
    from specter_amd.lib import SpxGrid
    g = SpxGrid(16, 16, 41, 25, 5, 1.0, 1.0, 1.0, 0)
    p = ctypes.c_void_p()
    assert lib.spx_plan_create(ctypes.byref(g), None, None, None, ctypes.byref(p)) == 1
    assert "tables" in _err(lib)


def test_null_plan_rejected(lib):
    assert lib.spx_fftp3d_real_to_complex(None, None, None) == 1
    assert lib.spx_substep(None, None, 1) == 1
    assert lib.spx_derivk(None, None, None, 1) == 1


def test_table_loader():
    from specter_amd.tables import load_dirichlet_tables
    A, Q = load_dirichlet_tables(27, 8)
    assert A.shape == (27, 8) and Q.shape == (8, 8)
    with pytest.raises(FileNotFoundError):
        load_dirichlet_tables(26, 5)


def test_initial_condition_properties():
    from specter_amd import ic
    nx, ny, nz, C = 16, 16, 41, 25
    u = ic.spectrum_peaked_velocity(nx, ny, nz, C, seed=7)
    npz = nz - C
    e = sum(np.mean(c[:npz] ** 2) for c in u)
    assert abs(e - 1.0) < 1e-12
    for c in u:
        assert np.all(c[0] == 0) and np.all(c[npz - 1] == 0) and np.all(c[npz:] == 0)
    u2 = ic.spectrum_peaked_velocity(nx, ny, nz, C, seed=7)
    assert all(np.array_equal(a, b) for a, b in zip(u, u2))       # counter-based RNG
    u3 = ic.spectrum_peaked_velocity(nx, ny, nz, C, seed=8)
    assert not np.array_equal(u[0], u3[0])
    th = ic.thermal_perturbation(nx, ny, nz, C, seed=1)
    assert np.all(th[0] == 0) and np.all(th[npz - 1] == 0) and abs(th[1:npz - 1].std() - 1e-3) < 2e-4
    fx, fy, fz = ic.large_scale_forcing(nx, ny, nz, C)
    x, y, _ = ic.physical_grid(nx, ny, nz, C)
    env = ic.wall_envelope(nz, C)
    assert np.allclose(fx[:npz], 0.5 * np.sin(2 * y)[None, :, None] * env[:, None, None])
    assert np.all(fz == 0)


def test_ensemble_stream_properties():
    """The ensemble workload generator: Poisson arrivals, grid / order /
    viscosity ranges, seeds 1..n (no GPU needed)."""
    from specter_amd.ensemble import member_bytes, poisson_stream
    reqs = poisson_stream()
    assert len(reqs) == 96
    t = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(t) > 0)
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert abs(gaps.mean() - 0.25) < 0.08          # rate 4/s
    assert {r.n for r in reqs} <= {32, 64, 128, 192}
    assert {r.d for r in reqs} <= {5, 8}
    assert all(5e-4 <= r.nu <= 5e-3 for r in reqs)
    assert [r.seed for r in reqs] == list(range(1, 97))
    assert all(r.nsteps == 300 and r.ord == 2 for r in reqs)
    sizes = sorted({r.n for r in reqs})
    mb = [member_bytes(next(r for r in reqs if r.n == n)) for n in sizes]
    assert mb == sorted(mb)


def test_solver_params_layout():
    """SolverParams maps the four solver kinds, the rotation vector and the MHD
    parameters onto
    spx_params (the C struct's field order: include/spx.h), and rejects
    unknown kinds before any device call."""
    import ctypes
    from specter_amd import SolverParams
    from specter_amd import lib as L
    names = [f[0] for f in L.SpxParams._fields_]
    assert names == ["solver", "ord", "dt", "nu", "kappa", "xmom", "xtemp", "v_zsta", "v_zend",
                     "omega", "mu", "b0", "bzsta", "bzend"]
    dsz = ctypes.sizeof(ctypes.c_double)
    assert L.SpxParams.omega.offset == L.SpxParams.v_zend.offset + 2 * dsz
    assert L.SpxParams.mu.offset == L.SpxParams.omega.offset + 3 * dsz
    assert L.SpxParams.bzsta.offset == L.SpxParams.b0.offset + 3 * dsz
    for kind, code in (("HD", 0), ("bouss", 1), ("ROTBOUSS", 2), ("MHD", 3)):
        assert SolverParams(solver=kind).to_c().solver == code
    c = SolverParams(solver="ROTBOUSS", omega=(0.5, -1.0, 2.0)).to_c()
    assert list(c.omega) == [0.5, -1.0, 2.0]
    c = SolverParams(solver="MHD", mu=0.1, b0=(1.0, 0.0, -2.0), bzsta=0, bzend=1).to_c()
    assert (c.mu, list(c.b0), c.bzsta, c.bzend) == (0.1, [1.0, 0.0, -2.0], 0, 1)
    with pytest.raises(ValueError):
        SolverParams(solver="MHDB").to_c()
