"""The Fortran driver (specter_amd/fortran/spx_hd.f90) calls libspx through
the iso_c_binding module with SPECTER's parameter.inp surface; its balance.txt
series and output fields must match the same run driven from Python and the
float64 oracle."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, TABLES, rel_l2

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "specter_amd", "fortran", "spx_hd")

PARAM = """&grid
nx = {nx}, ny = {ny}, nz = {nz}, Cz = {C}, oz = {d}, ord = {ord}, solver = "{solver}"
/
&boussinesq
gama = 0.5, xmom = 2.0, xtemp = 1.0
/
&scalar
c0 = 0.0, s0 = 0.0, skdn = 1.0, skup = 2.0, kappa = {kappa}
/
&status
idir = "in", odir = "out", tdir = "{tdir}", stat = 1, mult = 1, bench = 0, outs = 1
/
&boxparams
Lx = 1.0, Ly = 1.0, Lz = 1.0
/
&parameter
dt = {dt}, step = {step}, tstep = {tstep}, cstep = {cstep}, seed = 1000
/
&velocity
f0 = 0.0, u0 = 1.0, kdn = 2.0, kup = 4.0, nu = {nu}
/
&velbound
vbczsta = "noslip", vbczend = "noslip", vxzsta = 0.0, vxzend = 0.0, vyzsta = 0.0, vyzend = 0.0
/
&rotation
omegax = {ox}, omegay = {oy}, omegaz = {oz}
/
"""


def test_fortran_driver_matches_python_and_oracle(tmp_path):
    from oracle import specter_ref as O
    from specter_amd import FCPlan, Solver, SolverParams, ic
    if not os.path.exists(BIN):
        pytest.fail("Fortran driver not built (specter_amd.build.build_fortran)")
    nx, ny, nz, C, d, ordr, dt, nu, step, cstep = 32, 32, 57, 25, 5, 2, 1e-3, 2e-3, 12, 4
    tstep = 8   # one output (index 0002) at the start of step 9: the fields after 8 steps
    npz = nz - C
    (tmp_path / "in").mkdir()
    (tmp_path / "out").mkdir()
    (tmp_path / "parameter.inp").write_text(PARAM.format(
        nx=nx, ny=ny, nz=nz, C=C, d=d, ord=ordr, tdir=TABLES, dt=dt, step=step,
        tstep=tstep, cstep=cstep, nu=nu, solver="HD", kappa=1e-3, ox=0.0, oy=0.0, oz=0.0))
    u = ic.spectrum_peaked_velocity(nx, ny, nz, C, seed=1000)
    fr = ic.large_scale_forcing(nx, ny, nz, C)
    for name, arr in zip(("vx", "vy", "vz", "fx", "fy", "fz"), (*u, *fr)):
        # raw float64, Fortran order (nx, ny, nz-Cz) == C order (nz-Cz, ny, nx)
        np.ascontiguousarray(arr[:npz]).tofile(tmp_path / "in" / f"{name}.0001.out")
    res = subprocess.run([BIN], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    bal = np.loadtxt(tmp_path / "balance.txt", ndmin=2)
    nsd = np.loadtxt(tmp_path / "noslip_diagnostic.txt", ndmin=2)
    assert bal.shape[0] == (step - 1) // cstep

    # the same run from Python through the same library
    plan = FCPlan(nx, ny, nz, C, d)
    sol = Solver(plan, SolverParams(solver="HD", ord=ordr, dt=dt, nu=nu))
    for name, arr in zip(("vx", "vy", "vz", "fx", "fy", "fz"), (*u, *fr)):
        sol.set_real_field(name, arr)
    # and the oracle
    g = O.make_grid(nx, ny, nz, C, d, TABLES)
    st = O.HDState(*(O.fftp3d_real_to_complex(g, a.copy()) for a in u),
                   np.zeros(g.spec_shape, complex),
                   *(O.fftp3d_real_to_complex(g, a.copy()) for a in fr))
    rows_py, rows_or = [], []
    timec = 0
    snap = None
    for t in range(1, step + 1):
        if t == tstep + 1:
            snap = {n: sol.real_field(n) for n in ("vx", "vy", "vz")}
            snap["pr"] = plan.fftp2d_complex_to_real_xy(sol.field("pr").clone(),
                                                        plan.empty_real()).cpu().numpy()
            snap_or = {n: getattr(st, n).copy() for n in ("vx", "vy", "vz")}
        if timec == cstep:
            timec = 0
            rows_py.append(((t - 1) * dt, *sol.hdcheck()))
            rows_or.append(((t - 1) * dt, *O.hdcheck(g, st), *O.vdiagnostic(g, st)))
        sol.step(1)
        O.hd_step(g, st, ordr, dt, nu)
        timec += 1
    rows_py = np.array(rows_py)
    rows_or = np.array(rows_or)
    # balance.txt is written with 16 decimal digits (E23.16)
    assert np.allclose(bal[:, 1:3], rows_py[:, 1:3], rtol=1e-14, atol=0)
    assert np.allclose(bal[:, 1:3], rows_or[:, 1:3], rtol=1e-11, atol=0)
    # <div^2>, wall <v_t^2>, <v_n^2> (vboundary.f90:214-269) on the energy scale
    assert np.max(np.abs(nsd[:, 1:] - rows_or[:, 4:])) < 1e-11 * np.max(rows_or[:, 1])
    # output fields after tstep steps (physical space, v/(nx ny nz))
    for name in ("vx", "vy", "vz"):
        out = np.fromfile(tmp_path / "out" / f"{name}.0002.out").reshape(npz, ny, nx)
        py = snap[name]
        diff = np.abs(out - py)
        print(f"{name}: driver vs Python max |diff| {diff.max():.3e} at "
              f"{np.unravel_index(diff.argmax(), diff.shape)}, rel L2 {rel_l2(out, py):.3e}")
        # same library trajectory; the driver's final v/(nx ny nz) is Fortran
        # arithmetic, which the compiler may evaluate as a reciprocal product
        assert diff.max() <= 4 * np.finfo(float).eps * np.abs(py).max(), name
        ref = O.fftp3d_complex_to_real(g, snap_or[name].copy())[:npz] / (nx * ny * nz)
        assert rel_l2(out, ref) < 1e-11, name
    # pressure p'/(nx ny dt) from the mixed domain (specter.fpp:1041-1053)
    out = np.fromfile(tmp_path / "out" / "pr.0002.out").reshape(npz, ny, nx)
    py = snap["pr"][:npz] / (nx * ny * dt)
    assert np.abs(out - py).max() <= 4 * np.finfo(float).eps * np.abs(py).max()
    # vorticity (outs = 1): wx = dy vz - dz vy etc. against the oracle
    for name, (b, db, a, da) in {"wx": ("vz", 2, "vy", 3), "wy": ("vx", 3, "vz", 1),
                                 "wz": ("vy", 1, "vx", 2)}.items():
        out = np.fromfile(tmp_path / "out" / f"{name}.0002.out").reshape(npz, ny, nx)
        w = O.derivk(g, snap_or[b], db) - O.derivk(g, snap_or[a], da)
        ref = O.fftp3d_complex_to_real(g, w)[:npz] / (nx * ny * nz)
        assert rel_l2(out, ref) < 1e-10, name
    # no second output within 12 steps; the grid files
    assert not (tmp_path / "out" / "vx.0003.out").exists()
    x, y, z = ic.physical_grid(nx, ny, nz, C)
    assert np.allclose(np.loadtxt(tmp_path / "x.txt"), x, rtol=1e-14, atol=1e-15)
    assert np.allclose(np.loadtxt(tmp_path / "z.txt"), z[:npz], rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("solver", ["BOUSS", "ROTBOUSS"])
def test_fortran_driver_boussinesq(tmp_path, solver):
    """solver = BOUSS / ROTBOUSS (&rotation read for the latter): the scalar
    series (scalar.txt, wall <th^2>, Nusselt) and the th output against the
    Python-driven run and the oracle."""
    from oracle import specter_ref as O
    from specter_amd import FCPlan, Solver, SolverParams, ic
    if not os.path.exists(BIN):
        pytest.fail("Fortran driver not built (specter_amd.build.build_fortran)")
    nx, ny, nz, C, d, ordr, dt, nu, kappa, step, cstep = 16, 16, 41, 25, 5, 4, 1e-3, 2e-3, 3e-3, 7, 2
    xmom, xtemp = 0.5 * 2.0, 0.5 * 1.0          # gama * xmom, gama * xtemp
    omega = (0.25, -0.5, 2.0) if solver == "ROTBOUSS" else (0.0, 0.0, 0.0)
    tstep = 5
    npz = nz - C
    (tmp_path / "in").mkdir()
    (tmp_path / "out").mkdir()
    (tmp_path / "parameter.inp").write_text(PARAM.format(
        nx=nx, ny=ny, nz=nz, C=C, d=d, ord=ordr, tdir=TABLES, dt=dt, step=step,
        tstep=tstep, cstep=cstep, nu=nu, solver=solver, kappa=kappa, ox=omega[0],
        oy=omega[1], oz=omega[2]))
    u = ic.spectrum_peaked_velocity(nx, ny, nz, C, seed=1000)
    th = 1e-2 * ic.spectrum_peaked_velocity(nx, ny, nz, C, seed=1001)[0]
    fs = 0.1 * ic.spectrum_peaked_velocity(nx, ny, nz, C, seed=1002)[1]
    for name, arr in zip(("vx", "vy", "vz", "th", "fs"), (*u, th, fs)):
        np.ascontiguousarray(arr[:npz]).tofile(tmp_path / "in" / f"{name}.0001.out")
    res = subprocess.run([BIN], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    sca = np.loadtxt(tmp_path / "scalar.txt", ndmin=2)
    scd = np.loadtxt(tmp_path / "scalar_constant_diagnostic.txt", ndmin=2)
    nus = np.loadtxt(tmp_path / "nusselt.txt", ndmin=2)
    assert sca.shape[0] == scd.shape[0] == nus.shape[0] == (step - 1) // cstep

    prm = SolverParams(solver=solver, ord=ordr, dt=dt, nu=nu, kappa=kappa, xmom=xmom,
                       xtemp=xtemp, omega=omega)
    plan = FCPlan(nx, ny, nz, C, d)
    sol = Solver(plan, prm)
    for name, arr in zip(("vx", "vy", "vz", "th", "fs"), (*u, th, fs)):
        sol.set_real_field(name, arr)
    g = O.make_grid(nx, ny, nz, C, d, TABLES)
    zero = np.zeros(g.spec_shape, complex)
    st = O.HDState(*(O.fftp3d_real_to_complex(g, a.copy()) for a in u), zero.copy(),
                   zero.copy(), zero.copy(), zero.copy())
    st.th = O.fftp3d_real_to_complex(g, th.copy())
    st.fs = O.fftp3d_real_to_complex(g, fs.copy())
    rows_py, rows_or, snap = [], [], None
    timec = 0
    for t in range(1, step + 1):
        if t == tstep + 1:
            snap = sol.real_field("th")
            snap_or = st.th.copy()
        if timec == cstep:
            timec = 0
            rows_py.append(((t - 1) * dt, *sol.pscheck()))
            rows_or.append(((t - 1) * dt, *O.pscheck(g, st.th, st.fs),
                            *O.bouncheck_z(g, st.th), *O.nusselt(g, st.th)))
        sol.step(1)
        if solver == "ROTBOUSS":
            O.rotbouss_step(g, st, ordr, dt, nu, kappa, xmom, xtemp, omega)
        else:
            O.bouss_step(g, st, ordr, dt, nu, kappa, xmom, xtemp)
        timec += 1
    rows_py = np.array(rows_py)
    rows_or = np.array(rows_or)
    # scalar.txt carries 14 decimal digits (E22.14)
    assert np.allclose(sca[:, 1:4], rows_py[:, 1:4], rtol=1e-12, atol=0)
    assert np.allclose(sca[:, 1:3], rows_or[:, 1:3], rtol=1e-9, atol=0)
    scale = np.max(rows_or[:, 1])
    assert np.max(np.abs(scd[:, 1:] - rows_or[:, 4:6])) < 1e-9 * scale
    assert np.max(np.abs(nus[:, 1:] - rows_or[:, 6:8])) < 1e-9
    out = np.fromfile(tmp_path / "out" / "th.0002.out").reshape(npz, ny, nx)
    assert np.abs(out - snap).max() <= 4 * np.finfo(float).eps * np.abs(snap).max()
    ref = O.fftp3d_complex_to_real(g, snap_or)[:npz] / (nx * ny * nz)
    assert rel_l2(out, ref) < 1e-9


MHD_PARAM = """&magfield
m0 = 0.0, a0 = 1.0, mkdn = 2.0, mkup = 4.0, mu = {mu}, mparam0 = 0.0, aparam9 = 0.0
/
&magbound
bbcxsta = "periodic", bbcxend = "periodic", bbcysta = "periodic", bbcyend = "periodic",
bbczsta = "{bsta}", bbczend = "{bend}"
/
&uniformb
bx0 = {bx0}, by0 = {by0}, bz0 = {bz0}
/
"""


def test_fortran_driver_mhd(tmp_path):
    """solver = MHD with &magfield / &magbound / &uniformb: the mhdcheck files
    (balance, energy, helicity, cross) and the ax, bx, ph outputs against the
    Python-driven run of the same library and the oracle."""
    from oracle import specter_ref as O
    from specter_amd import FCPlan, Solver, SolverParams, ic
    if not os.path.exists(BIN):
        pytest.fail("Fortran driver not built (specter_amd.build.build_fortran)")
    nx, ny, nz, C, d, ordr, dt, nu, mu, step, cstep = 16, 16, 41, 25, 5, 2, 1e-3, 5e-2, 4e-2, 7, 2
    b0 = (0.2, 0.0, 0.3)
    tstep = 5
    npz = nz - C
    (tmp_path / "in").mkdir()
    (tmp_path / "out").mkdir()
    (tmp_path / "parameter.inp").write_text(PARAM.format(
        nx=nx, ny=ny, nz=nz, C=C, d=d, ord=ordr, tdir=TABLES, dt=dt, step=step,
        tstep=tstep, cstep=cstep, nu=nu, solver="MHD", kappa=1e-3, ox=0.0, oy=0.0, oz=0.0)
        + MHD_PARAM.format(mu=mu, bsta="conducting", bend="vacuum", bx0=b0[0], by0=b0[1],
                           bz0=b0[2]))
    u = ic.spectrum_peaked_velocity(nx, ny, nz, C, seed=1000)
    x, y, z = ic.physical_grid(nx, ny, nz, C)
    zz, yy, xx = np.meshgrid(z, y, x, indexing="ij")
    s = np.sin(np.pi * zz)
    ar = [0.3 * s * np.cos(yy), 0.3 * s * np.sin(xx + yy), 0.3 * np.cos(np.pi * zz) * np.cos(xx)]
    for name, arr in zip(("vx", "vy", "vz", "ax", "ay", "az"), (*u, *ar)):
        np.ascontiguousarray(arr[:npz]).tofile(tmp_path / "in" / f"{name}.0001.out")
    res = subprocess.run([BIN], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    files = {n: np.loadtxt(tmp_path / f"{n}.txt", ndmin=2)
             for n in ("balance", "energy", "helicity", "cross")}
    for f in files.values():
        assert f.shape[0] == (step - 1) // cstep

    prm = SolverParams(solver="MHD", ord=ordr, dt=dt, nu=nu, mu=mu, b0=b0, bzsta=0, bzend=1)
    plan = FCPlan(nx, ny, nz, C, d)
    sol = Solver(plan, prm)
    for name, arr in zip(("vx", "vy", "vz", "ax", "ay", "az"), (*u, *ar)):
        sol.set_real_field(name, arr)
    g = O.make_grid(nx, ny, nz, C, d, TABLES)
    neu, neu2 = O.neumann_vector(g, TABLES, 1), O.neumann_vector(g, TABLES, 2)
    zero = np.zeros(g.spec_shape, complex)
    sp_ = lambda r: O.fftp3d_real_to_complex(g, r.copy())  # noqa: E731
    st = O.MHDState(*(sp_(r) for r in u), zero.copy(), zero.copy(), zero.copy(), zero.copy(),
                    *(sp_(r) for r in ar), zero.copy(), zero.copy(), zero.copy(), zero.copy())
    rows_py, rows_or, snap = [], [], None
    timec = 0
    for t in range(1, step + 1):
        if t == tstep + 1:
            snap = {n: sol.real_field(n) for n in ("ax", "ay", "az")}
            snap["ph"] = plan.fftp2d_complex_to_real_xy(sol.field("ph").clone(),
                                                        plan.empty_real()).cpu().numpy()
            snap_or = {n: getattr(st, n).copy() for n in ("ax", "ay", "az")}
        if timec == cstep:
            timec = 0
            rows_py.append(((t - 1) * dt, *sol.mhdcheck()))
            rows_or.append(((t - 1) * dt, *O.mhdcheck(g, st)))
        sol.step(1)
        O.mhd_step(g, st, ordr, dt, nu, mu, b0, neu, neu2, 0, 1)
        timec += 1
    rows_py, rows_or = np.array(rows_py), np.array(rows_or)
    # columns of mhdcheck: eng ens cur engk engm helk helm crh asq
    cols = {"balance": [1, 2, 3], "energy": [4, 5], "helicity": [6, 7], "cross": [8, 9]}
    for n, c in cols.items():
        assert np.allclose(files[n][:, 1:], rows_py[:, c], rtol=1e-14, atol=1e-15), n
        scale = np.abs(rows_or[:, c]).max()
        assert np.abs(files[n][:, 1:] - rows_or[:, c]).max() < 1e-9 * scale, n
    for name in ("ax", "ay", "az"):
        out = np.fromfile(tmp_path / "out" / f"{name}.0002.out").reshape(npz, ny, nx)
        py = snap[name]
        assert np.abs(out - py).max() <= 4 * np.finfo(float).eps * np.abs(py).max(), name
        ref = O.fftp3d_complex_to_real(g, snap_or[name].copy())[:npz] / (nx * ny * nz)
        assert rel_l2(out, ref) < 1e-10, name
    out = np.fromfile(tmp_path / "out" / "ph.0002.out").reshape(npz, ny, nx)
    py = snap["ph"][:npz] / (nx * ny * dt)
    assert np.abs(out - py).max() <= 4 * np.finfo(float).eps * np.abs(py).max()
    # b = curl a (outs = 1)
    out = np.fromfile(tmp_path / "out" / "bx.0002.out").reshape(npz, ny, nx)
    w = O.derivk(g, snap_or["az"], 2) - O.derivk(g, snap_or["ay"], 3)
    ref = O.fftp3d_complex_to_real(g, w)[:npz] / (nx * ny * nz)
    assert rel_l2(out, ref) < 1e-10
