"""Pinning of the oracle's RK step compositions and of the pressure that
sol_project returns (CPU only).

The operator-level checks in test_oracle.py pin each routine; these pin how
they are composed into RK stages and steps (include/hd/hd_rkstep2.f90:3-36,
include/bouss/bouss_rkstep2.f90:3-59, specter.fpp:1142-1161) and the pressure
output d of sol_project (boundary_mod.fpp:354-381), with facts the
mathematics fixes:

* linear decay: a field with kx = 0 and only an x component has no nonlinear
  term and no pressure, so a step is a polynomial in the wall-projected
  Laplacian.  For an eigenvector of that operator (built here from the DFT
  matrix, the FC-Gram continuation matrix and the wall rows) one step must
  multiply it by sum_{j<=ORD} (nu dt lambda)^j / j! (the SPECTER scheme
  v = v0 + dt/o RHS(v), o = ORD..1).  Its slowest eigenvalue is the
  continuous -(ky^2 + pi^2/Lz^2) to FC accuracy.
* no-slip: the pressure time-extrapolation factor of noslip_z
  (vboundary.f90:195-196) makes the tangential wall velocity after a stage
  O(dt^2); any other factor leaves O(dt).
* pressure identity: what the projection removes from (vx, vy) is
  (i kx, i ky) d and from vz the z derivative of d.
* Boussinesq rest state: theta(z) alone must leave v = 0, decay as the
  linear mode, and give the hydrostatic pressure d pr/dz = dt xmom theta;
  a vertical velocity with theta = 0 must produce theta = dt xtemp vz.

Each check is shown to fail on a planted mistake of the kind it guards
against (stage order, stage factor 1/o, stage start value, wall enforcement,
the extrapolation factor, the pressure's sign, the buoyancy and heat-current
signs).
"""
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN, TABLES
from oracle import specter_ref as O
from specter_amd import ic


def grid(nx, ny, nz, C, d):
    return O.make_grid(nx, ny, nz, C, d, TABLES)


def zero_like(g):
    return np.zeros(g.spec_shape, complex)


def taylor(h, ordr):
    return sum(h ** j / math.factorial(j) for j in range(ordr + 1))


def to_mixed(g, a):
    """IFFT_z(a)/nz: the mixed (z, ky, kx) domain of boundary_mod.fpp:72-150."""
    b = a.copy()
    O.fftp1d_complex_to_real_z(g, b)
    return b / g.nz


# ---------------------------------------------------------------------------
# the per-line operator of one stage on a kx = 0 line, from definitions
# ---------------------------------------------------------------------------
def wall_projected_laplacian(g, ky):
    """K = P diag(-(ky^2 + kz^2)) on one z line, where P maps spectral
    coefficients to physical values (backward DFT / nz), zeroes the two wall
    values, re-applies the FC continuation (the matrix of fftp.fpp:755-773
    built from A*Q^T) and transforms back (forward DFT)."""
    nz, npz, C, d = g.nz, g.np_, g.Cz, g.d
    m = np.arange(nz)
    F = np.exp(-2j * np.pi * np.outer(m, m) / nz)
    A, Q = O.load_tables(TABLES, C, d)
    D = A @ Q.T
    cont = np.zeros((nz, npz))
    cont[:npz] = np.eye(npz)
    for ii in range(C):
        for jj in range(d):
            cont[npz + ii, npz - d + jj] += D[ii, jj]
            cont[npz + ii, d - 1 - jj] += D[C - 1 - ii, jj]
    walls = np.eye(npz)
    walls[0, 0] = walls[-1, -1] = 0.0
    P = F @ cont @ walls @ (np.conj(F)[:npz] / nz)
    K = P @ np.diag(-(ky ** 2 + g.kz ** 2))
    w, V = np.linalg.eig(K)
    keep = np.abs(w) > 1e-6 * np.abs(w).max()
    return w[keep].real, V[:, keep], P


def hd_step_mut(g, st, ordr, dt, nu, order="ord..1", rmp_one=False, chain=False, walls=True):
    """A copy of the HD step with one planted mistake: the stage order, the
    stage factor, the stage start value or the wall enforcement."""
    c0 = (st.vx.copy(), st.vy.copy(), st.vz.copy())
    stages = range(ordr, 0, -1) if order == "ord..1" else range(1, ordr + 1)
    for o in stages:
        rmp = 1.0 if rmp_one else 1.0 / o
        C4, C5, C6 = O.gradre(g, st.vx, st.vy, st.vz)
        for a in (C4, C5, C6):
            O.fc_filter(g, a)
        start = (st.vx.copy(), st.vy.copy(), st.vz.copy()) if chain else c0
        for name, s0, nl, f in zip(("vx", "vy", "vz"), start, (C4, C5, C6),
                                   (st.fx, st.fy, st.fz)):
            v = getattr(st, name)
            v[...] = s0 + dt * (nu * O.laplak(g, v) - nl + f) * rmp
        if walls:
            O.v_imposebc_and_project(g, st.vx, st.vy, st.vz, st.pr, o, ordr)


def _eigen_state(g, lam_target):
    """Oracle state whose vx holds, on the (kx, ky) = (0, 2) line, the
    eigenvector of the stage operator nearest lam_target."""
    w, V, P = wall_projected_laplacian(g, 2.0)
    k = int(np.argmin(np.abs(w - lam_target)))
    vx = zero_like(g)
    vx[0, 2] = V[:, k] / np.linalg.norm(V[:, k])
    z = zero_like(g)
    return O.HDState(vx, z.copy(), z.copy(), z.copy(), z.copy(), z.copy(), z.copy()), w[k]


def test_linear_mode_eigenvalue_is_the_continuous_one():
    """The three slowest decay rates of the discrete wall-projected Laplacian
    on the ky = 2 line: -(ky^2 + (m pi / Lz)^2), m = 1, 2, 3, to FC accuracy,
    converging at high order in the number of physical points (measured
    relative errors 8.4e-7, 1.6e-5, 8.0e-5 at 32 points; 2.4e-8, 4.9e-7,
    2.6e-6 at 64; 7.4e-10, 1.5e-8, 8.0e-8 at 128)."""
    tol = {57: (2e-6, 4e-5, 2e-4), 89: (1e-7, 1.5e-6, 8e-6), 153: (3e-9, 5e-8, 2.5e-7)}
    errs = {}
    for nz in (57, 89, 153):
        g = grid(8, 8, nz, 25, 5)
        w, _, _ = wall_projected_laplacian(g, 2.0)
        w = np.sort(w)[::-1]
        for m in (1, 2, 3):
            ref = -(4.0 + (m * np.pi / g.Lz) ** 2)
            errs[nz, m] = abs(w[m - 1] - ref) / abs(ref)
            assert errs[nz, m] < tol[nz][m - 1], (nz, m, w[m - 1])
    for m in (1, 2, 3):
        assert errs[57, m] > 20 * errs[89, m] and errs[89, m] > 20 * errs[153, m]


def test_linear_mode_decays_as_the_analytic_field():
    """The real field sin(2y) sin(pi z/Lz) e_x, not an exact discrete
    eigenvector, decays at the continuous rate to FC accuracy."""
    g = grid(8, 8, 89, 25, 5)
    R = np.zeros(g.real_shape)
    R[: g.np_] = (np.sin(2 * g.y)[None, :, None] * np.sin(np.pi * g.z[: g.np_] / g.Lz)[:, None, None]
                  * np.ones(g.nx)[None, None, :])
    vx = O.fftp3d_real_to_complex(g, R.copy())
    z = zero_like(g)
    st = O.HDState(vx.copy(), z.copy(), z.copy(), z.copy(), z.copy(), z.copy(), z.copy())
    nu, dt, n = 1.0, 4e-5, 50
    for _ in range(n):
        O.hd_step(g, st, 4, dt, nu)
    fac = taylor(-nu * (4 + np.pi ** 2) * dt, 4) ** n
    assert np.linalg.norm(st.vx - fac * vx) / np.linalg.norm(vx) < 5e-5
    assert np.abs(st.vy).max() == 0 and np.abs(st.vz).max() == 0 and np.abs(st.pr).max() == 0


@pytest.mark.parametrize("ordr", [1, 2, 3, 4])
def test_step_is_the_taylor_polynomial_on_an_eigenmode(ordr):
    """One step multiplies an eigenvector of the stage operator by
    sum_{j<=ORD} (h lambda)^j/j!, h = nu dt; h lambda = -0.5 here, so every
    term of the polynomial is far above roundoff."""
    g = grid(8, 8, 57, 25, 5)
    nu, dt = 1.0, 2e-4
    st, lam = _eigen_state(g, -0.5 / (nu * dt))
    v0 = st.vx.copy()
    nsteps = 6
    for _ in range(nsteps):
        O.hd_step(g, st, ordr, dt, nu)
    fac = taylor(nu * dt * lam, ordr) ** nsteps
    assert np.linalg.norm(st.vx - fac * v0) < 1e-10
    # one term more or less is far outside the tolerance
    assert abs(taylor(nu * dt * lam, ordr + 1) ** nsteps - fac) > 1e-6


@pytest.mark.parametrize("mistake", ["order", "rmp", "chain", "walls"])
def test_planted_step_composition_mistakes_fail(mistake):
    g = grid(8, 8, 57, 25, 5)
    nu, dt, ordr = 1.0, 2e-4, 4
    st, lam = _eigen_state(g, -0.5 / (nu * dt))
    v0 = st.vx.copy()
    # the mutation harness itself reproduces the oracle when nothing is planted
    ok = st.copy()
    hd_step_mut(g, ok, ordr, dt, nu)
    ref = st.copy()
    O.hd_step(g, ref, ordr, dt, nu)
    assert np.array_equal(ok.vx, ref.vx)
    kw = {"order": dict(order="1..ord"), "rmp": dict(rmp_one=True), "chain": dict(chain=True),
          "walls": dict(walls=False)}[mistake]
    hd_step_mut(g, st, ordr, dt, nu, **kw)
    fac = taylor(nu * dt * lam, ordr)
    assert np.linalg.norm(st.vx - fac * v0) > 1e-6


# ---------------------------------------------------------------------------
# no-slip to second order in dt: the pressure extrapolation factor tau
# ---------------------------------------------------------------------------
def _wall_slip(g, u, ordr, dt, t_end, noslip=None, monkeypatch=None):
    if noslip is not None:
        monkeypatch.setattr(O, "noslip_z", noslip)
    z = zero_like(g)
    st = O.HDState(*(O.fftp3d_real_to_complex(g, a.copy()) for a in u), z.copy(), z.copy(),
                   z.copy(), z.copy())
    for _ in range(int(round(t_end / dt))):
        O.hd_step(g, st, ordr, dt, 2e-3)
    if noslip is not None:
        monkeypatch.undo()
    bot, top = O.bouncheck_z(g, st.vx, st.vy)
    return math.sqrt((bot + top) / O.energy(g, st.vx, st.vy, st.vz, 1))


def _noslip_tau(tau_of):
    """noslip_z with the extrapolation factor tau_of(o, ord) (vboundary.f90:154-211)."""
    def f(g, o, ordr, vx, vy, pr, vbound, pos):
        ind = 0 if pos == 0 else g.np_ - 1
        tmp = tau_of(o, ordr)
        p = pr[:, :, ind]
        vx[:, :, ind] = O._ik(g.kx[:, None], p) * tmp
        vy[:, :, ind] = O._ik(g.ky[None, :], p) * tmp
        vx[0, 0, ind] = float(g.nx * g.ny) * vbound[0]
        vy[0, 0, ind] = float(g.nx * g.ny) * vbound[1]
    return f


@pytest.mark.parametrize("ordr", [2, 4])
def test_wall_slip_is_second_order(ordr):
    g = grid(16, 16, 89, 25, 5)
    u = ic.spectrum_peaked_velocity(16, 16, 89, 25, 1000)
    s1 = _wall_slip(g, u, ordr, 2e-3, 0.02)
    s2 = _wall_slip(g, u, ordr, 1e-3, 0.02)
    assert s1 < 2e-5
    assert 3.6 < s1 / s2 < 4.4, (s1, s2)


@pytest.mark.parametrize("tau", ["always_1/o", "branches_swapped"])
def test_planted_tau_mistakes_leave_first_order_slip(tau, monkeypatch):
    g = grid(16, 16, 89, 25, 5)
    u = ic.spectrum_peaked_velocity(16, 16, 89, 25, 1000)
    # the harness reproduces the oracle's factor (1/o at o = ord, (o+1)/o otherwise)
    good = _noslip_tau(lambda o, ordr: 1.0 / o if o == ordr else (o + 1.0) / o)
    assert _wall_slip(g, u, 2, 2e-3, 0.02, good, monkeypatch) == _wall_slip(g, u, 2, 2e-3, 0.02)
    bad = {"always_1/o": lambda o, ordr: 1.0 / o,
           "branches_swapped": lambda o, ordr: (o + 1.0) / o if o == ordr else 1.0 / o}[tau]
    s1 = _wall_slip(g, u, 2, 2e-3, 0.02, _noslip_tau(bad), monkeypatch)
    s2 = _wall_slip(g, u, 2, 1e-3, 0.02, _noslip_tau(bad), monkeypatch)
    assert s1 > 1e-4 and s1 / s2 < 2.5


# ---------------------------------------------------------------------------
# the pressure returned by sol_project
# ---------------------------------------------------------------------------
def _smooth_nonsolenoidal(g):
    u = ic.spectrum_peaked_velocity(g.nx, g.ny, g.nz, g.Cz, 1000)
    Z, Y, X = np.meshgrid(g.z, g.y, g.x, indexing="ij")
    ph = np.arange(g.nz)[:, None, None] < g.np_
    pert = (0.3 * np.sin(X) * np.cos(2 * Y) * np.cos(3 * Z), 0.2 * np.cos(Y + X) * Z ** 2,
            0.4 * np.cos(2 * X) * np.exp(Z) + 0.1 * Z)
    return [O.fftp3d_real_to_complex(g, np.where(ph, a + p, 0.0)) for a, p in zip(u, pert)]


def _fc_dz(g, a):
    """d/dz of the physical values of a mixed-domain line set, via the FC
    continuation and the spectral derivative."""
    b = a.copy()
    b[:, :, g.np_:] = 0
    O.fftp1d_real_to_complex_z(g, b)
    return to_mixed(g, O._ik(g.kz[None, None, :], b))[:, :, : g.np_]


@pytest.mark.parametrize("nz,ztol", [(57, 1e-5), (89, 5e-7), (153, 3e-8)])
def test_sol_project_pressure_identity(nz, ztol):
    g = grid(16, 16, nz, 25, 5)
    a, b, c = _smooth_nonsolenoidal(g)
    a0, b0, c0 = a.copy(), b.copy(), c.copy()
    d = zero_like(g)
    O.sol_project(g, a, b, c, d, 1, 0, 0)
    p = g.np_
    ra = to_mixed(g, a0 - a)[:, :, :p]
    rb = to_mixed(g, b0 - b)[:, :, :p]
    rc = to_mixed(g, c0 - c)[:, :, :p]
    # horizontal parts: exact (the z transforms commute with i kx, i ky)
    gx = O._ik(g.kx[:, None, None], d)[:, :, :p]
    gy = O._ik(g.ky[None, :, None], d)[:, :, :p]
    assert np.abs(ra - gx).max() < 1e-12 * np.abs(ra).max()
    assert np.abs(rb - gy).max() < 1e-12 * np.abs(rb).max()
    # vertical part: the FC derivative of the returned pressure
    dz = _fc_dz(g, d)
    assert np.linalg.norm(rc - dz) < ztol * np.linalg.norm(rc)
    # a pressure of the wrong sign, or without the harmonic part, fails
    assert np.abs(ra + gx).max() > 1.0 * np.abs(ra).max()
    dh = to_mixed(g, d.copy())   # not the pressure: an extra z transform
    assert np.abs(ra - O._ik(g.kx[:, None, None], dh)[:, :, :p]).max() > 1e-3 * np.abs(ra).max()


def test_sol_project_pressure_without_harmonic_part_fails():
    g = grid(16, 16, 89, 25, 5)
    a, b, c = _smooth_nonsolenoidal(g)
    a0 = a.copy()
    d = zero_like(g)
    # the potential alone (IFFT_z(d)/nz without laplace_z's a, :370-381)
    dpot = O.poisson_inhomogeneous(g, a, b, c)
    O.sol_project(g, a, b, c, d, 1, 0, 0)
    p = g.np_
    ra = to_mixed(g, a0 - a)[:, :, :p]
    gx = O._ik(g.kx[:, None, None], to_mixed(g, dpot))[:, :, :p]
    assert np.abs(ra - gx).max() > 1e-3 * np.abs(ra).max()


# ---------------------------------------------------------------------------
# Boussinesq: rest state with a vertical temperature profile, heat current
# ---------------------------------------------------------------------------
def bouss_substep_mut(g, st, C1, C2, C3, C7, o, ordr, dt, nu, kappa, xmom, xtemp,
                      buoy=-1.0, heat=-1.0):
    """bouss_substep with the signs of the buoyancy (C6 - xmom th) and heat
    current (C8 - xtemp vz) terms as parameters (-1: the reference's)."""
    rmp = 1.0 / float(o)
    C4, C5, C6 = O.gradre(g, st.vx, st.vy, st.vz)
    C8 = O.advect(g, st.vx, st.vy, st.vz, st.th)
    C6 = C6 + buoy * xmom * st.th
    C8 = C8 + heat * xtemp * st.vz
    for arr in (C4, C5, C6, C8):
        O.fc_filter(g, arr)
    lx, ly, lz, lt = (O.laplak(g, st.vx), O.laplak(g, st.vy), O.laplak(g, st.vz),
                      O.laplak(g, st.th))
    st.vx[...] = C1 + dt * (nu * lx - C4 + st.fx) * rmp
    st.vy[...] = C2 + dt * (nu * ly - C5 + st.fy) * rmp
    st.vz[...] = C3 + dt * (nu * lz - C6 + st.fz) * rmp
    st.th[...] = C7 + dt * (kappa * lt - C8 + st.fs) * rmp
    O.v_imposebc_and_project(g, st.vx, st.vy, st.vz, st.pr, o, ordr)
    O.s_imposebc(g, st.th)
    O.fc_filter(g, st.th)
    R1 = O.fftp3d_complex_to_real(g, st.th.copy())
    R1 = R1 / g.nx / g.ny / g.nz
    st.th[...] = O.fftp3d_real_to_complex(g, R1)


def _rest_state(buoy, ordr=4, nsteps=10):
    g = grid(8, 8, 89, 25, 5)
    R = np.zeros(g.real_shape)
    R[: g.np_] = 1e-2 * np.sin(np.pi * g.z[: g.np_] / g.Lz)[:, None, None]
    th0 = O.fftp3d_real_to_complex(g, R.copy())
    z = zero_like(g)
    st = O.HDState(z.copy(), z.copy(), z.copy(), z.copy(), z.copy(), z.copy(), z.copy(),
                   th0.copy(), z.copy())
    kappa, dt, xmom, xtemp = 1.0, 2e-5, 3.0, 2.0
    for step in range(nsteps):
        C1, C2, C3, C7 = st.vx.copy(), st.vy.copy(), st.vz.copy(), st.th.copy()
        for o in range(ordr, 0, -1):
            th_in = st.th.copy()
            if buoy is None:
                O.bouss_substep(g, st, C1, C2, C3, C7, o, ordr, dt, 1.0, kappa, xmom, xtemp)
            else:
                bouss_substep_mut(g, st, C1, C2, C3, C7, o, ordr, dt, 1.0, kappa, xmom, xtemp,
                                  buoy=buoy)
    return g, st, th0, th_in, dict(kappa=kappa, dt=dt, xmom=xmom, nsteps=nsteps, ordr=ordr)


def _hydrostatic_error(g, st, th_in, q):
    """|d pr/dz - dt xmom th_in| / |dt xmom th_in| on the (0, 0) line after a
    step (the last stage, o = 1, projects (dt/1) xmom th_in out of vz)."""
    dpr = _fc_dz(g, st.pr)[0, 0]
    ref = q["dt"] * q["xmom"] * to_mixed(g, th_in)[0, 0, : g.np_]
    return np.abs(dpr - ref).max() / np.abs(ref).max()


def test_bouss_rest_state():
    g, st, th0, th_in, q = _rest_state(None)
    scale = q["dt"] * q["xmom"] * np.abs(to_mixed(g, th0)).max()
    for name in ("vx", "vy", "vz"):
        assert np.abs(to_mixed(g, getattr(st, name))[:, :, : g.np_]).max() < 1e-12 * scale, name
    fac = taylor(-q["kappa"] * np.pi ** 2 * q["dt"], q["ordr"]) ** q["nsteps"]
    assert np.linalg.norm(st.th - fac * th0) / np.linalg.norm(th0) < 5e-5
    assert _hydrostatic_error(g, st, th_in, q) < 1e-6
    # only the horizontal mean line carries pressure
    assert np.abs(st.pr[1:]).max() == 0 and np.abs(st.pr[0, 1:]).max() == 0


def test_bouss_planted_buoyancy_sign_fails():
    # the harness with the reference's sign reproduces the oracle
    g, st, _, th_in, q = _rest_state(-1.0, nsteps=2)
    _, st_ref, _, _, _ = _rest_state(None, nsteps=2)
    assert np.array_equal(st.pr, st_ref.pr)
    g, st, _, th_in, q = _rest_state(+1.0, nsteps=2)
    assert _hydrostatic_error(g, st, th_in, q) > 1.0


def _heat_current(heat):
    g = grid(16, 16, 89, 25, 5)
    u = ic.spectrum_peaked_velocity(16, 16, 89, 25, 7)
    z = zero_like(g)
    st = O.HDState(*(O.fftp3d_real_to_complex(g, a.copy()) for a in u), z.copy(), z.copy(),
                   z.copy(), z.copy(), z.copy(), z.copy())
    vz0 = st.vz.copy()
    dt, xtemp = 1e-4, 2.0
    C = (st.vx.copy(), st.vy.copy(), st.vz.copy(), st.th.copy())
    if heat is None:
        O.bouss_substep(g, st, *C, 1, 1, dt, 1e-3, 1e-3, 0.0, xtemp)
    else:
        bouss_substep_mut(g, st, *C, 1, 1, dt, 1e-3, 1e-3, 0.0, xtemp, heat=heat)
    # theta = dt xtemp vz0 with the Nyquist modes removed by the filter; the
    # filter runs on C8 before the update and on theta after the wall
    # condition (bouss_rkstep2.f90:54), which the single filter here matches
    # to ~1e-6 of the norm
    ref = dt * xtemp * vz0
    O.fc_filter(g, ref)
    a = to_mixed(g, st.th)[:, :, : g.np_]
    b = to_mixed(g, ref)[:, :, : g.np_]
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_bouss_heat_current():
    assert _heat_current(None) < 1e-5


def test_bouss_planted_heat_current_sign_fails():
    assert _heat_current(-1.0) == _heat_current(None)
    assert _heat_current(+1.0) > 1.0


# ---------------------------------------------------------------------------
# rotating Boussinesq (include/rotbouss/rotbouss_rkstep2.f90)
# ---------------------------------------------------------------------------
def _inertial_state(g):
    """Horizontally uniform flow (U(z), 0, 0), U = sin(pi z/Lz): no nonlinear
    term (kx = ky = 0, vz = 0), no pressure, zero wall values."""
    R = np.zeros(g.real_shape)
    R[: g.np_] = np.sin(np.pi * g.z[: g.np_] / g.Lz)[:, None, None]
    u0 = O.fftp3d_real_to_complex(g, R.copy())
    z = zero_like(g)
    return O.HDState(u0.copy(), z.copy(), z.copy(), z.copy(), z.copy(), z.copy(), z.copy(),
                     z.copy(), z.copy()), u0


def _rot_steps(g, st, ordr, dt, omega, nsteps, sign=1.0):
    """rotbouss steps with nu = kappa = xmom = xtemp = 0; sign = -1 plants a
    Coriolis term of the wrong sign."""
    for _ in range(nsteps):
        if sign == 1.0:
            O.rotbouss_step(g, st, ordr, dt, 0.0, 0.0, 0.0, 0.0, omega)
        else:
            O.rotbouss_step(g, st, ordr, dt, 0.0, 0.0, 0.0, 0.0, tuple(-w for w in omega))


def _inertial_line_model(g, u0, ordr, dt, omz, nsteps, filtered=True):
    """The (kx, ky) = (0, 0) line under rotation about z, from definitions:
    per stage U <- P (U0 + dt/o 2 Omega_z f V), V <- P (V0 - dt/o 2 Omega_z f U),
    with f the z filter factors and P the wall map (backward DFT, zero wall
    values, FC-Gram continuation, forward DFT) of wall_projected_laplacian."""
    _, _, P = wall_projected_laplacian(g, 0.0)
    f = O.filter_factors(g)[2] if filtered else np.ones(g.nz)
    U, V = u0[0, 0].copy(), np.zeros(g.nz, complex)
    for _ in range(nsteps):
        C1, C2 = U.copy(), V.copy()
        for o in range(ordr, 0, -1):
            U, V = (P @ (C1 + dt / o * 2 * omz * f * V), P @ (C2 - dt / o * 2 * omz * f * U))
    return U, V


@pytest.mark.parametrize("ordr", [2, 3, 4])
def test_rotbouss_inertial_oscillation(ordr):
    """With Omega = Omega_z e_z and no dissipation, (U, V) obeys
    U' = 2 Omega_z V, V' = -2 Omega_z U.  Continuously, one RK step multiplies
    it by P(h J) = c I + s J, J = [[0, 1], [-1, 0]], c + i s =
    sum_{j<=ORD} (i h)^j/j!, h = 2 Omega_z dt, so after n steps
    (U, V) = (C U0, -S U0) with C + i S = (c + i s)^n (here to the accuracy
    with which the filtered FC line represents sin(pi z/Lz): 1e-5); the
    discrete line follows the matrix model of _inertial_line_model to 1e-9
    (the roundoff of the continuation, amplified as in the HD eigenmode
    test), and a model without the filter of the Coriolis term is 1e-6 off.
    h = 0.5, so the neighbouring orders' polynomials are far off."""
    g = grid(8, 8, 89, 25, 5)
    dt, h, n = 1e-3, 0.5, 4
    st, u0 = _inertial_state(g)
    omz = h / (2 * dt)
    _rot_steps(g, st, ordr, dt, (0.0, 0.0, omz), n)
    nrm = np.linalg.norm(u0)
    U, V = _inertial_line_model(g, u0, ordr, dt, omz, n)
    eu = np.linalg.norm(st.vx[0, 0] - U) / nrm
    ev = np.linalg.norm(st.vy[0, 0] - V) / nrm
    assert eu < 1e-9 and ev < 1e-9, (eu, ev)
    U1, V1 = _inertial_line_model(g, u0, ordr, dt, omz, n, filtered=False)
    assert np.linalg.norm(st.vx[0, 0] - U1) / nrm > 1e-8
    assert np.abs(st.vx[1:]).max() == 0 and np.abs(st.vx[0, 1:]).max() == 0
    assert np.abs(st.vz).max() < 1e-13 * nrm and np.abs(st.th).max() == 0
    z = taylor(1j * h, ordr) ** n
    assert np.linalg.norm(st.vx - z.real * u0) / nrm < 1e-5
    assert np.linalg.norm(st.vy + z.imag * u0) / nrm < 1e-5
    for other in (ordr - 1, ordr + 1):
        assert abs(taylor(1j * h, other) ** n - z) > 1e-3


def test_rotbouss_planted_coriolis_sign_fails():
    g = grid(8, 8, 89, 25, 5)
    dt, h, n = 1e-3, 0.5, 4
    st, u0 = _inertial_state(g)
    _rot_steps(g, st, 4, dt, (0.0, 0.0, h / (2 * dt)), n, sign=-1.0)
    z = taylor(1j * h, 4) ** n
    assert np.linalg.norm(st.vy + z.imag * u0) / np.linalg.norm(u0) > 0.5


def test_rotbouss_coriolis_is_twice_omega_cross_v():
    """The three Coriolis terms are the components of 2 Omega x v (checked
    against numpy's cross product on a random state and a generic Omega)."""
    g = grid(8, 8, 57, 25, 5)
    rng = np.random.default_rng(4)
    v = [rng.standard_normal(g.spec_shape) + 1j * rng.standard_normal(g.spec_shape)
         for _ in range(3)]
    z = zero_like(g)
    st = O.HDState(*v, z, z, z, z)
    omega = (0.3, -1.7, 2.9)
    got = np.stack(O.coriolis_terms(st, omega), axis=-1)
    ref = 2 * np.cross(np.array(omega), np.stack(v, axis=-1))
    assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()


def test_rotbouss_without_rotation_is_bouss_without_theta_filter():
    """With Omega = 0 a rotating stage followed by the Boussinesq stage's
    filter and real-space round trip of th (bouss_rkstep2.f90:54-59) is the
    Boussinesq stage, bit for bit: the two solvers share every other step."""
    g = grid(16, 16, 89, 25, 5)
    u = ic.spectrum_peaked_velocity(16, 16, 89, 25, 9)
    th = ic.thermal_perturbation(16, 16, 89, 25, 10)
    z = zero_like(g)
    base = O.HDState(*(O.fftp3d_real_to_complex(g, a.copy()) for a in u), z.copy(), z.copy(),
                     z.copy(), z.copy(), O.fftp3d_real_to_complex(g, th.copy()), z.copy())
    args = (3, 4, 1e-3, 2e-3, 3e-3, 1.5, 0.7)
    s1 = O.HDState(*(getattr(base, f).copy() for f in ("vx", "vy", "vz", "pr", "fx", "fy",
                                                        "fz", "th", "fs")))
    s2 = O.HDState(*(getattr(base, f).copy() for f in ("vx", "vy", "vz", "pr", "fx", "fy",
                                                        "fz", "th", "fs")))
    C = (base.vx.copy(), base.vy.copy(), base.vz.copy(), base.th.copy())
    O.rotbouss_substep(g, s1, *C, *args, (0.0, 0.0, 0.0))
    O.fc_filter(g, s1.th)
    R1 = O.fftp3d_complex_to_real(g, s1.th.copy()) / g.nx / g.ny / g.nz
    s1.th[...] = O.fftp3d_real_to_complex(g, R1)
    O.bouss_substep(g, s2, *C, *args)
    for f in ("vx", "vy", "vz", "pr", "th"):
        assert np.array_equal(getattr(s1, f), getattr(s2, f)), f


# ---------------------------------------------------------------------------
# the longdouble oracle (the extended-precision referee of the cfg1 fixture)
# ---------------------------------------------------------------------------
def test_longdouble_oracle_is_the_same_algorithm():
    """One HD RK2 step in longdouble and in float64 from the same fields:
    they differ only by float64 roundoff (a transposed operand or a dropped
    term in either would differ at O(1)); the longdouble run is closer to a
    second longdouble run with a different FFT order of operations than to
    float64 (it carries more digits)."""
    nx, ny, nz, C, d = 16, 16, 41, 25, 5
    g64 = grid(nx, ny, nz, C, d)
    gld = O.make_grid(nx, ny, nz, C, d, TABLES, precision="longdouble")
    assert gld.kx.dtype == np.longdouble and gld.dirm.dtype == np.longdouble
    # constants are the float64 ones, held exactly
    assert np.array_equal(gld.kk2.astype(np.float64), g64.kk2)
    assert np.array_equal(gld.dirm.astype(np.float64), g64.dirm)
    u = ic.spectrum_peaked_velocity(nx, ny, nz, C, 7)
    out = {}
    for g in (g64, gld):
        zero = np.zeros(g.spec_shape, g.cdtype)
        st = O.HDState(*(O.fftp3d_real_to_complex(g, a.astype(g.rdtype)) for a in u),
                       zero.copy(), zero.copy(), zero.copy(), zero.copy())
        O.hd_step(g, st, 2, 1e-3, 2e-3)
        assert st.vx.dtype == g.cdtype
        out[g.precision] = st
    for n in ("vx", "vy", "vz"):
        a = getattr(out["longdouble"], n)
        b = getattr(out["float64"], n).astype(np.clongdouble)
        rel = float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(a.ravel()))
        assert 1e-17 < rel < 1e-12, (n, rel)


def test_cfg1_longdouble_fixture():
    """The committed difference between the longdouble and float64 cfg1
    trajectories: finite, of the size the float64 drift has (2-5e-11)."""
    z = np.load(os.path.join(GOLDEN, "cfg1_longdouble_delta.npz"))
    for n in ("vx", "vy", "vz"):
        assert np.all(np.isfinite(z[f"{n}_re"])) and np.all(np.isfinite(z[f"{n}_im"]))
        assert 1e-11 < float(z[f"{n}_rel"]) < 1e-10
        assert np.abs(z[f"{n}_re"]).max() <= 1.0 and np.abs(z[f"{n}_im"]).max() <= 1.0
