"""GPU parity of the libspx transforms and pointwise operators against the
float64 oracle, through the C ABI (specter_amd.FCPlan)."""
import numpy as np
import pytest

from conftest import TABLES, rel_err

pytestmark = pytest.mark.gpu

# (nx, ny, nz, Cz, d): nz prime (Bluestein), even smooth, 3*19, 7*13 (direct radix 7/13)
SHAPES = [
    (16, 16, 41, 25, 5),
    (16, 16, 64, 25, 5),
    (32, 24, 57, 25, 5),
    (12, 20, 91, 27, 8),
    (64, 64, 89, 25, 5),
]
TOL = 1e-13


def _proj_tol(shape):
    """The d=8 continuation amplifies roundoff of white-noise input: a 1-ulp
    input perturbation moves the oracle's own sol_project output by 1.2e-11
    (d=5: 5e-15)."""
    return 1e-12 if shape[4] <= 5 else 1e-10


def _setup(shape):
    import torch
    from oracle import specter_ref as O
    from specter_amd import FCPlan
    nx, ny, nz, C, d = shape
    g = O.make_grid(nx, ny, nz, C, d, TABLES)
    p = FCPlan(nx, ny, nz, C, d)
    return torch, O, g, p


def _rand_spec(g, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(g.spec_shape) + 1j * rng.standard_normal(g.spec_shape)


def _rand_real(g, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(g.real_shape)


def _cu(torch, a):
    return torch.as_tensor(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_r2c_3d(shape):
    torch, O, g, p = _setup(shape)
    R = _rand_real(g, 1)
    ref = O.fftp3d_real_to_complex(g, R.copy())
    out = p.empty_spectral()
    p.fftp3d_real_to_complex(_cu(torch, R), out)
    assert rel_err(out.cpu().numpy(), ref) < TOL


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_c2r_3d(shape):
    torch, O, g, p = _setup(shape)
    S = _rand_spec(g, 2)
    ref = O.fftp3d_complex_to_real(g, S.copy())
    out = p.empty_real()
    p.fftp3d_complex_to_real(_cu(torch, S), out)
    assert rel_err(out.cpu().numpy(), ref) < TOL


@pytest.mark.parametrize("shape", SHAPES[:3], ids=lambda s: "x".join(map(str, s)))
def test_xy_2d(shape):
    torch, O, g, p = _setup(shape)
    R = _rand_real(g, 3)
    ref = O.fftp2d_real_to_complex_xy(g, R.copy())
    out = p.empty_spectral()
    p.fftp2d_real_to_complex_xy(_cu(torch, R), out)
    assert rel_err(out.cpu().numpy(), ref) < TOL
    S = _rand_spec(g, 4)
    ref2 = O.fftp2d_complex_to_real_xy(g, S.copy())
    out2 = p.empty_real()
    p.fftp2d_complex_to_real_xy(_cu(torch, S), out2)
    assert rel_err(out2.cpu().numpy(), ref2) < TOL


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_z_transforms_and_continuation(shape):
    torch, O, g, p = _setup(shape)
    S = _rand_spec(g, 5)
    ref = S.copy()
    O.fc_continue(g, ref)
    t = _cu(torch, S)
    p.fc_continue(t)
    # continuation alone is the same float64 operation sequence: bit-exact
    assert np.array_equal(t.cpu().numpy(), ref)
    ref = S.copy()
    O.fftp1d_real_to_complex_z(g, ref)
    t = _cu(torch, S)
    p.fftp1d_real_to_complex_z(t)
    assert rel_err(t.cpu().numpy(), ref) < TOL
    ref = S.copy()
    O.fftp1d_complex_to_real_z(g, ref)
    t = _cu(torch, S)
    p.fftp1d_complex_to_real_z(t)
    assert rel_err(t.cpu().numpy(), ref) < TOL


@pytest.mark.parametrize("shape", SHAPES[:2], ids=lambda s: "x".join(map(str, s)))
def test_pointwise_operators(shape):
    torch, O, g, p = _setup(shape)
    S = _rand_spec(g, 6)
    for dirn in (1, 2, 3):
        out = p.empty_spectral()
        p.derivk(_cu(torch, S), out, dirn)
        assert np.array_equal(out.cpu().numpy(), O.derivk(g, S, dirn))
    out = p.empty_spectral()
    p.laplak(_cu(torch, S), out)
    assert np.array_equal(out.cpu().numpy(), O.laplak(g, S))
    t = _cu(torch, S)
    p.fc_filter(t)
    ref = S.copy()
    O.fc_filter(g, ref)
    assert rel_err(t.cpu().numpy(), ref) < 1e-15


@pytest.mark.parametrize("shape", SHAPES[:4], ids=lambda s: "x".join(map(str, s)))
def test_gradre(shape):
    torch, O, g, p = _setup(shape)
    a, b, c = (_rand_spec(g, s) for s in (7, 8, 9))
    ref = O.gradre(g, a, b, c)
    outs = [p.empty_spectral() for _ in range(3)]
    p.gradre(_cu(torch, a), _cu(torch, b), _cu(torch, c), *outs)
    for o, r in zip(outs, ref):
        assert rel_err(o.cpu().numpy(), r) < 1e-12


@pytest.mark.parametrize("shape", SHAPES[:4], ids=lambda s: "x".join(map(str, s)))
def test_advect(shape):
    torch, O, g, p = _setup(shape)
    a, b, c, d = (_rand_spec(g, s) for s in (10, 11, 12, 13))
    ref = O.advect(g, a, b, c, d)
    out = p.empty_spectral()
    p.advect(_cu(torch, a), _cu(torch, b), _cu(torch, c), _cu(torch, d), out)
    assert rel_err(out.cpu().numpy(), ref) < 1e-12


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_sol_project(shape):
    torch, O, g, p = _setup(shape)
    a, b, c, d = (_rand_spec(g, s) for s in (14, 15, 16, 17))
    ra, rb, rc, rd = a.copy(), b.copy(), c.copy(), d.copy()
    O.sol_project(g, ra, rb, rc, rd, 1, 0, 0)
    ta, tb, tc, td = (_cu(torch, x) for x in (a, b, c, d))
    p.sol_project(ta, tb, tc, td, 1, 0, 0)
    for t, r in ((ta, ra), (tb, rb), (tc, rc), (td, rd)):
        assert rel_err(t.cpu().numpy(), r) < _proj_tol(shape)


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("o,ordr", [(2, 2), (1, 2), (3, 4)])
def test_v_imposebc_and_project(shape, o, ordr):
    torch, O, g, p = _setup(shape)
    vx, vy, vz, pr = (_rand_spec(g, s) for s in (18, 19, 20, 21))
    r = [x.copy() for x in (vx, vy, vz, pr)]
    O.v_imposebc_and_project(g, *r, o, ordr, (0.0, 0.0), (0.0, 0.0))
    ts = [_cu(torch, x) for x in (vx, vy, vz, pr)]
    p.v_imposebc_and_project(*ts, o, ordr)
    for t, rr in zip(ts, r):
        assert rel_err(t.cpu().numpy(), rr) < _proj_tol(shape)


@pytest.mark.parametrize("shape", SHAPES[:3], ids=lambda s: "x".join(map(str, s)))
def test_sol_project_without_basis_table(shape, monkeypatch):
    """SPX_HARMONIC_TABLE=0: the harmonic correction transforms its profiles
    per line; same results as the table path and the oracle."""
    torch, O, g, p = _setup(shape)
    from specter_amd import FCPlan
    monkeypatch.setenv("SPX_HARMONIC_TABLE", "0")
    q = FCPlan(*shape)
    vx, vy, vz, pr = (_rand_spec(g, s) for s in (18, 19, 20, 21))
    r = [x.copy() for x in (vx, vy, vz, pr)]
    O.v_imposebc_and_project(g, *r, 2, 2, (0.0, 0.0), (0.0, 0.0))
    ta = [_cu(torch, x) for x in (vx, vy, vz, pr)]
    tb = [_cu(torch, x) for x in (vx, vy, vz, pr)]
    p.v_imposebc_and_project(*ta, 2, 2)
    q.v_imposebc_and_project(*tb, 2, 2)
    for a, b, rr in zip(ta, tb, r):
        assert rel_err(b.cpu().numpy(), rr) < _proj_tol(shape)
        assert rel_err(b.cpu().numpy(), a.cpu().numpy()) < 1e-12


@pytest.mark.parametrize("shape", SHAPES[:3], ids=lambda s: "x".join(map(str, s)))
def test_s_imposebc(shape):
    torch, O, g, p = _setup(shape)
    th = _rand_spec(g, 22)
    r = th.copy()
    O.s_imposebc(g, r)
    t = _cu(torch, th)
    p.s_imposebc(t)
    assert rel_err(t.cpu().numpy(), r) < 1e-12


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_diagnostics(shape):
    torch, O, g, p = _setup(shape)
    a, b, c, d, e, f = (_rand_spec(g, s) for s in range(30, 36))
    ta, tb, tc, td, te, tf = (_cu(torch, x) for x in (a, b, c, d, e, f))
    for kin in (0, 1):
        assert rel_err(p.energy(ta, tb, tc, kin), O.energy(g, a, b, c, kin)) < 1e-12
    assert rel_err(p.cross(ta, tb, tc, td, te, tf), O.cross(g, a, b, c, d, e, f)) < 1e-12
    assert rel_err(p.divergence(ta, tb, tc), O.divergence(g, a, b, c)) < 1e-12
    assert rel_err(p.helicity(ta, tb, tc), O.helicity(g, a, b, c)) < 1e-12
    bt = p.bouncheck_z(ta, tb)
    rb = O.bouncheck_z(g, a, b)
    assert rel_err(np.array(bt), np.array(rb)) < 1e-12
    bt = p.bouncheck_z(tc)
    rb = O.bouncheck_z(g, c)
    assert rel_err(np.array(bt), np.array(rb)) < 1e-12


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("o,ordr", [(2, 2), (3, 4)])
def test_wall_update_mfma_matches_transform_path(shape, o, ordr, monkeypatch):
    """The wall update as a rank-(2+C) correction on the matrix cores (the
    default) against the same update through the z transforms
    (SPX_WALL_MFMA=0 at plan creation), and both against the oracle; also the
    scalar wall condition."""
    torch, O, g, p = _setup(shape)
    from specter_amd import FCPlan
    monkeypatch.setenv("SPX_WALL_MFMA", "0")
    q = FCPlan(*shape)
    vx, vy, vz, pr = (_rand_spec(g, s) for s in (40, 41, 42, 43))
    r = [x.copy() for x in (vx, vy, vz, pr)]
    O.v_imposebc_and_project(g, *r, o, ordr, (0.0, 0.0), (0.0, 0.0))
    ta = [_cu(torch, x) for x in (vx, vy, vz, pr)]
    tb = [_cu(torch, x) for x in (vx, vy, vz, pr)]
    p.v_imposebc_and_project(*ta, o, ordr)
    q.v_imposebc_and_project(*tb, o, ordr)
    for a, b, rr in zip(ta, tb, r):
        assert rel_err(a.cpu().numpy(), rr) < _proj_tol(shape)
        assert rel_err(a.cpu().numpy(), b.cpu().numpy()) < _proj_tol(shape)
    th = _rand_spec(g, 44)
    rs = th.copy()
    O.s_imposebc(g, rs)
    t1, t2 = _cu(torch, th), _cu(torch, th)
    p.s_imposebc(t1)
    q.s_imposebc(t2)
    assert rel_err(t1.cpu().numpy(), rs) < 1e-12
    assert rel_err(t1.cpu().numpy(), t2.cpu().numpy()) < 1e-12


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_wall_update_fragment_tables_match_roots(shape, monkeypatch):
    """The two forms of the wall correction's DFT operands -- formed from the
    roots of unity staged in LDS (the default) and read from per-plan tables
    in the operands' lane order (SPX_WALL_TAB=0 at plan creation) -- give the
    same wall update to rounding, and both match the oracle."""
    torch, O, g, p = _setup(shape)
    from specter_amd import FCPlan
    monkeypatch.setenv("SPX_WALL_TAB", "0")
    q = FCPlan(*shape)
    vx, vy, vz, pr = (_rand_spec(g, s) for s in (50, 51, 52, 53))
    r = [x.copy() for x in (vx, vy, vz, pr)]
    O.v_imposebc_and_project(g, *r, 2, 2, (0.0, 0.0), (0.0, 0.0))
    ta = [_cu(torch, x) for x in (vx, vy, vz, pr)]
    tb = [_cu(torch, x) for x in (vx, vy, vz, pr)]
    p.v_imposebc_and_project(*ta, 2, 2)
    q.v_imposebc_and_project(*tb, 2, 2)
    for a, b, rr in zip(ta, tb, r):
        assert rel_err(b.cpu().numpy(), rr) < _proj_tol(shape)
        assert rel_err(a.cpu().numpy(), b.cpu().numpy()) < _proj_tol(shape)


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_fused_projection_matches_two_kernel_path(precision, monkeypatch):
    """sol_project in one pass per line (the default) gives the same bits as
    the Poisson-projection kernel followed by the harmonic-correction kernel
    (SPX_FUSED_PROJECT=0 at plan creation): two full RK4 steps, HD and
    Boussinesq."""
    from specter_amd import FCPlan, Solver, SolverParams, ic
    nx, ny, nz, C, d = 32, 24, 57, 25, 5
    u = ic.spectrum_peaked_velocity(nx, ny, nz, C, 21)
    th = ic.thermal_perturbation(nx, ny, nz, C, 22)
    for prm in (dict(solver="HD", ord=4, dt=1e-3, nu=2e-3),
                dict(solver="BOUSS", ord=4, dt=1e-3, nu=2e-3, kappa=2e-3, xmom=1.0, xtemp=1.0)):
        out = []
        for fused in ("1", "0"):
            monkeypatch.setenv("SPX_FUSED_PROJECT", fused)
            sol = Solver(FCPlan(nx, ny, nz, C, d, precision=precision), SolverParams(**prm))
            for name, arr in zip(("vx", "vy", "vz"), u):
                sol.set_real_field(name, arr)
            if prm["solver"] == "BOUSS":
                sol.set_real_field("th", th)
            sol.step(2)
            out.append({n: sol.spectral(n) for n in ("vx", "vy", "vz", "pr")})
        for n in ("vx", "vy", "vz", "pr"):
            assert np.array_equal(out[0][n], out[1][n]), (prm["solver"], n)


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_fused_x_stage_matches_six_launches(precision, monkeypatch):
    """gradre's x stage in one launch (the default: the real velocity kept in
    registers, stored only when advect reuses it) agrees with the three xc2r
    + three xgrad_prod launches (SPX_XFUSED=0 at plan creation) to rounding
    (the same operations; the compiler contracts the FFT's products into
    fused multiply-adds differently in the two kernels): gradre alone, and
    two RK4 steps of HD and Boussinesq (whose advect reads the real velocity
    the fused stage stores)."""
    import torch
    from specter_amd import FCPlan, Solver, SolverParams, ic
    nx, ny, nz, C, d = 32, 24, 57, 25, 5
    u = ic.spectrum_peaked_velocity(nx, ny, nz, C, 31)
    th = ic.thermal_perturbation(nx, ny, nz, C, 32)
    rng = np.random.default_rng(33)
    shape = (nx // 2 + 1, ny, nz)
    abc = [torch.as_tensor(rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).cuda()
           for _ in range(3)]
    res = []
    for fused in ("1", "0"):
        monkeypatch.setenv("SPX_XFUSED", fused)
        p = FCPlan(nx, ny, nz, C, d, precision=precision)
        outs = [p.empty_nonlinear() for _ in range(3)]
        p.gradre(*abc, *outs)
        r = {"gradre": [o.cpu().numpy() for o in outs]}
        for prm in (dict(solver="HD", ord=4, dt=1e-3, nu=2e-3),
                    dict(solver="BOUSS", ord=4, dt=1e-3, nu=2e-3, kappa=2e-3, xmom=1.0,
                         xtemp=1.0)):
            sol = Solver(FCPlan(nx, ny, nz, C, d, precision=precision), SolverParams(**prm))
            for name, arr in zip(("vx", "vy", "vz"), u):
                sol.set_real_field(name, arr)
            if prm["solver"] == "BOUSS":
                sol.set_real_field("th", th)
            sol.step(2)
            names = ("vx", "vy", "vz", "th") if prm["solver"] == "BOUSS" else ("vx", "vy", "vz")
            r[prm["solver"]] = [sol.spectral(n) for n in names]
        res.append(r)
    # gradre itself to rounding; the steps' differences grow through the
    # pressure assembly (measured 1e-12 after two float64 RK4 steps), far
    # below what a wrong real velocity for advect would give (O(1) in th)
    tols = {"gradre": 1e-14 if precision == 0 else 1e-6,
            "HD": 1e-10 if precision == 0 else 1e-4, "BOUSS": 1e-10 if precision == 0 else 1e-4}
    for key in res[0]:
        tol = tols[key]
        for a, b in zip(res[0][key], res[1][key]):
            e = rel_err(a, b)
            print(f"{key}: fused vs six launches {e:.2e}")
            assert e < tol, key
