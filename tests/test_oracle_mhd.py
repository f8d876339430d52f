"""Pinning of the oracle's MHD pieces (CPU only): the Neumann / Robin wall
reconstruction (fcgram_mod.f90:260-644), sol_project with the potential's
Dirichlet and Robin conditions and laplace_z's (0,0), (2,2), (0,2) cases
(boundary_mod.fpp:197-678), the magnetic boundary conditions
(bboundary.f90:100-344), prodre / vector (pseudospec_hd.f90:322-402,
pseudospec_mhd.f90) and the MHD stage (include/mhd/mhd_rkstep2.f90).

Checks fixed by the mathematics, not by the reference's formulas:
* the Neumann vectors are the one-sided finite-difference weights of the
  degree d-1 interpolant on d equispaced points (exact rationals), for the
  outward normal derivative (order 1) and the second derivative (order 2);
* sol_project with the potential's conditions removes a gradient field grad
  psi exactly when psi satisfies the chosen homogeneous conditions (and only
  then), and returns psi as the potential;
* prodre(v) = (v.grad)v - grad(|v|^2/2) against the independently pinned
  gradre; vector(A, B) = numpy.cross of the real fields;
* after a_imposebc_and_project the wall values satisfy the discrete
  conditions built from the exact weights;
* an MHD step: resistive decay of a (0,0)-line potential mode is the RK
  Taylor polynomial of -mu (pi/Lz)^2 dt; in an ideal step (nu = mu = 0) the
  kinetic energy gained equals the magnetic energy lost (a flipped Lorentz
  force or electromotive force breaks it).
"""
import math
from fractions import Fraction as Fr

import numpy as np
import pytest

from conftest import TABLES
from oracle import specter_ref as O


def fd_weights(d, m):
    """w_j with f^(m)(0) = sum_j w_j f(j) + O(h^d), j = 0..d-1 (exact)."""
    M = [[Fr(j) ** p for j in range(d)] + [Fr(math.factorial(m) if p == m else 0)]
         for p in range(d)]
    for c in range(d):
        piv = next(r for r in range(c, d) if M[r][c] != 0)
        M[c], M[piv] = M[piv], M[c]
        for r in range(d):
            if r != c and M[r][c] != 0:
                f = M[r][c] / M[c][c]
                M[r] = [x - f * y for x, y in zip(M[r], M[c])]
    return [M[i][d] / M[i][i] for i in range(d)]


def spec(g, R):
    return O.fftp3d_real_to_complex(g, R.copy())


def real(g, S):
    return O.fftp3d_complex_to_real(g, S.copy()) / (g.nx * g.ny * g.nz)


@pytest.mark.parametrize("C,d,nz", [(25, 5, 41), (25, 5, 89), (27, 8, 91)])
def test_neumann_vectors_are_one_sided_fd_weights(C, d, nz):
    g = O.make_grid(8, 8, nz, C, d, TABLES)
    h = float(g.dz)
    w1, w2 = fd_weights(d, 1), fd_weights(d, 2)
    # outward derivative g1 = -f'(0): f0 = -(h g1... ) solved for f0
    e1 = [float(-w1[d - 1 - k] / w1[0]) for k in range(d - 1)] + [float(-h / w1[0])]
    e2 = [float(-w2[d - 1 - k] / w2[0]) for k in range(d - 1)] + [float(h * h / w2[0])]
    assert np.abs(O.neumann_vector(g, TABLES, 1) - e1).max() < 1e-14 * max(map(abs, e1))
    assert np.abs(O.neumann_vector(g, TABLES, 2) - e2).max() < 1e-14 * max(map(abs, e2))


@pytest.mark.parametrize("pos", [0, 1])
def test_reconstruct_polynomials(pos):
    """Degree < d polynomials: Neumann (outward first, second derivative) and
    Robin (d_n f + kh f = 0) reconstruction return the wall value."""
    g = O.make_grid(4, 4, 41, 25, 5, TABLES)
    neu, neu2 = O.neumann_vector(g, TABLES, 1), O.neumann_vector(g, TABLES, 2)
    z = np.asarray(g.z, float)
    zw = z[0] if pos == 0 else z[g.np_ - 1]
    s = -1.0 if pos == 0 else 1.0            # outward normal
    c = np.array([0.3, -1.1, 0.7, 2.0, -0.4])  # degree 4 = d - 1
    t = z - zw
    p = sum(ck * t ** k for k, ck in enumerate(c))
    ind = 0 if pos == 0 else g.np_ - 1
    f = np.broadcast_to(p, g.spec_shape).astype(complex).copy()
    f[:, :, ind] = s * c[1]                  # outward first derivative
    O.neumann_reconstruct(g, f, pos, neu)
    assert np.abs(f[:, :, ind] - c[0]).max() < 1e-12
    f[:, :, ind] = 2 * c[2]                  # second derivative
    O.neumann_reconstruct(g, f, pos, neu2)
    assert np.abs(f[:, :, ind] - c[0]).max() < 1e-12
    # Robin: pick c0 so that s c1 + kh c0 = 0 on every line with kh > 0
    kh = g.khom
    m = kh > 0
    c0 = -s * c[1] / np.where(m, kh, 1.0)
    f = (p[None, None, :] - c[0] + c0[:, :, None]).astype(complex)
    f[:, :, ind] = 0.0
    O.robin_reconstruct(g, f, pos, neu)
    assert np.abs((f[:, :, ind] - c0)[m]).max() < 1e-12 * np.abs(c0).max()


def _psi_cases(Lz, kh):
    a3 = -kh * (2 + kh * Lz) / (Lz ** 2 * (3 + kh * Lz))
    a2 = -(1 + kh * Lz) / (2 * Lz + kh * Lz ** 2)
    return {
        (2, 2): (lambda z: 1 + kh * z + a3 * z ** 3, lambda z: kh + 3 * a3 * z ** 2),
        (0, 2): (lambda z: z + a2 * z ** 2, lambda z: 1 + 2 * a2 * z),
        (0, 0): (lambda z: z * (Lz - z), lambda z: Lz - 2 * z),
    }


@pytest.mark.parametrize("nz", [57, 89])
def test_sol_project_potential_conditions(nz):
    """v = grad psi with psi = cos(x) G(z): psi meets the homogeneous Robin
    (d_n + khom) / Dirichlet conditions of exactly one kind pair; only that
    pair removes v entirely, and returns psi as the potential."""
    Lz = 1.3
    g = O.make_grid(8, 8, nz, 25, 5, TABLES, Lz=Lz)
    z, y, x = np.meshgrid(g.z, g.y, g.x, indexing="ij")
    for kinds, (G, dG) in _psi_cases(Lz, 1.0).items():
        v = [spec(g, -np.sin(x) * G(z)), spec(g, 0 * x), spec(g, np.cos(x) * dG(z))]
        psi_xy = np.fft.rfft2(np.cos(x) * G(z), axes=(1, 2))  # (z, ky, kx)
        for try_kinds in _psi_cases(Lz, 1.0):
            a, b, c = (w.copy() for w in v)
            ph = np.zeros_like(a)
            O.sol_project(g, a, b, c, ph, 0, *try_kinds)
            resid = max(np.abs(real(g, w)[: g.np_]).max() for w in (a, b, c))
            if try_kinds == kinds:
                assert resid < 1e-12, (kinds, resid)
                got = ph[:, :, : g.np_]
                ref = psi_xy.transpose(2, 1, 0)[:, :, : g.np_]
                assert np.abs(got - ref).max() < 1e-11 * np.abs(ref).max()
            else:
                assert resid > 0.1, (kinds, try_kinds, resid)


def test_laplace_z_closed_forms():
    """The (2,2) and (0,2) harmonic solutions: a'' = khom^2 a, b = a', and the
    boundary data a' - kh a = -bc1 at z = 0 (Robin) or a = bc1 (Dirichlet),
    a' + kh a = bc2 at z = Lz; the kh = 0 line is linear."""
    g = O.make_grid(8, 8, 225, 25, 5, TABLES, Lz=1.7)
    rng = np.random.default_rng(2)
    bc = rng.standard_normal((g.nkx, g.ny, 2)) + 1j * rng.standard_normal((g.nkx, g.ny, 2))
    kh = g.khom
    top = g.np_ - 1
    for kinds in ((2, 2), (0, 2)):
        a, b = O.laplace_z(g, bc, *kinds)
        z = np.asarray(g.z)
        # b is the z derivative: centred differences of the closed form
        m = kh > 0
        lhs0 = b[:, :, 0] - kh * a[:, :, 0]
        if kinds[0] == 2:
            assert np.abs((lhs0 + bc[:, :, 0])[m]).max() < 1e-12 * np.abs(bc).max()
        else:
            assert np.abs((a[:, :, 0] - bc[:, :, 0])[m]).max() < 1e-12 * np.abs(bc).max()
        lhs1 = b[:, :, top] + kh * a[:, :, top]
        assert np.abs((lhs1 - bc[:, :, 1])[m]).max() < 1e-12 * np.abs(bc).max()
        # a'' = kh^2 a on the interior (second differences, O(dz^2))
        dz = float(z[1] - z[0])
        d2 = (a[:, :, 2:g.np_] - 2 * a[:, :, 1:g.np_ - 1] + a[:, :, :g.np_ - 2]) / dz ** 2
        ref = (kh ** 2)[:, :, None] * a[:, :, 1:g.np_ - 1]
        assert np.abs(d2 - ref).max() < 1e-3 * np.abs(ref).max()
        # kh = 0 (real parts only, as the reference): a linear with slope b
        assert np.allclose(np.diff(a[0, 0, :g.np_].real) / dz, b[0, 0, 0].real)
        assert abs(b[0, 0, 0] - (bc[0, 0, 1].real if kinds == (0, 2) else bc[0, 0, 0].real)) < 1e-14


def test_prodre_is_gradre_minus_kinetic_gradient():
    """(curl v) x v = (v.grad)v - grad(|v|^2/2) on a smooth field (x and y
    components to rounding, z to the FC derivative's accuracy at the walls,
    1e-4); a swapped factor order (v x curl v) fails."""
    g = O.make_grid(16, 16, 89, 25, 5, TABLES)
    z, y, x = np.meshgrid(g.z, g.y, g.x, indexing="ij")
    s = np.sin(np.pi * z) ** 2
    vr = [s * np.sin(y) + 0.3 * np.cos(z), s * np.sin(x + y), 0.5 * s * np.cos(x)]
    v = [spec(g, w) for w in vr]
    P = O.prodre(g, *v)
    G = O.gradre(g, *v)
    q = spec(g, 0.5 * sum(w * w for w in vr))
    K = [O.derivk(g, q, k) for k in (1, 2, 3)]
    scale = max(np.abs(real(g, w)[: g.np_]).max() for w in G)
    for p_, g_, k_ in zip(P, G, K):
        diff = real(g, p_ - (g_ - k_))[: g.np_]
        assert np.abs(diff).max() < 1e-3 * scale   # FC z derivative at the walls
    bad = [-p_ for p_ in P]
    assert np.abs(real(g, bad[0] - (G[0] - K[0]))[: g.np_]).max() > 0.1 * scale


def test_vector_is_numpy_cross():
    g = O.make_grid(8, 8, 41, 25, 5, TABLES)
    z, y, x = np.meshgrid(g.z, g.y, g.x, indexing="ij")
    # smooth fields (random ones continue into ~1e7 and lose the comparison
    # to the round trip's rounding)
    A = [np.sin(x + 2 * z), np.cos(y - z) * np.sin(x), 0.5 + np.cos(2 * z)]
    B = [np.cos(x) * np.cos(z), 0.3 * np.sin(y + z), np.sin(2 * x + y) + z]
    out = O.vector(g, *(spec(g, w) for w in A), *(spec(g, w) for w in B))
    cr = np.cross(np.stack(A, -1)[: g.np_], np.stack(B, -1)[: g.np_])
    for c in range(3):
        assert np.abs(real(g, out[c])[: g.np_] - cr[..., c]).max() < 1e-13 * np.abs(cr).max()


def _state(g, vr, ar):
    zero = lambda: np.zeros(g.spec_shape, complex)  # noqa: E731
    return O.MHDState(*(spec(g, w) for w in vr), zero(), zero(), zero(), zero(),
                      *(spec(g, w) for w in ar), zero(), zero(), zero(), zero())


@pytest.mark.parametrize("kinds", [(0, 0), (1, 1)])
def test_wall_values_after_a_imposebc(kinds):
    """Conducting: d2/dz2 ax = d2/dz2 ay = 0, d/dz az = 0 at the wall; insulating:
    d_n a + khom a = 0; discrete forms from the exact weights."""
    g = O.make_grid(8, 8, 57, 25, 5, TABLES)
    neu, neu2 = O.neumann_vector(g, TABLES, 1), O.neumann_vector(g, TABLES, 2)
    rng = np.random.default_rng(4)
    z, y, x = np.meshgrid(g.z, g.y, g.x, indexing="ij")
    ar = [np.cos(x + 2 * y) * np.sin(3 * z) + 0.1 * rng.standard_normal(g.real_shape)
          for _ in range(3)]
    st = _state(g, [0 * x] * 3, ar)
    O.a_imposebc_and_project(g, neu, neu2, st.ax, st.ay, st.az, st.ph, *kinds)
    h = float(g.dz)
    w1, w2 = ([float(w) for w in fd_weights(5, m)] for m in (1, 2))
    for pos in (0, 1):
        vals = []
        for f in (st.ax, st.ay, st.az):
            m = f.copy()
            O.goto_domain_w_boundaries(g, m)
            idx = list(range(5)) if pos == 0 else [g.np_ - 1 - j for j in range(5)]
            vals.append(m[:, :, idx])           # wall first, then inwards
        sc = max(np.abs(v).max() for v in vals)
        for comp, v in enumerate(vals):
            d1 = -sum(w1[j] * v[:, :, j] for j in range(5)) / h    # outward
            d2 = sum(w2[j] * v[:, :, j] for j in range(5)) / h ** 2
            if kinds[pos] == 0:
                r = d1 if comp == 2 else d2
                assert np.abs(r).max() * h ** (1 if comp == 2 else 2) < 1e-12 * sc
            else:
                r = d1 + g.khom * v[:, :, 0]
                assert np.abs(r).max() * h < 1e-12 * sc


@pytest.mark.parametrize("kind,prof,tol", [(0, "sin", 5e-7), (1, "cos", 5e-8)])
@pytest.mark.parametrize("ordr", [2, 4])
def test_resistive_decay_of_a_line_mode(kind, prof, tol, ordr):
    """v = 0, no uniform field: a = (P(z), 0, 0) on the kx = ky = 0 line with P
    = sin(pi z/Lz) (conducting walls) or cos (insulating: Robin with khom = 0)
    decays per step by sum_{j<=ORD} h^j/j!, h = -mu (pi/Lz)^2 dt, to the
    accuracy of the one-sided wall stencils (dt inside the explicit diffusion
    limit of the grid's largest kz); growth (a sign mistake in mu) fails."""
    Lz, mu, dt = 1.0, 0.05, 0.002
    g = O.make_grid(4, 4, 89, 25, 5, TABLES, Lz=Lz)
    neu, neu2 = O.neumann_vector(g, TABLES, 1), O.neumann_vector(g, TABLES, 2)
    z, y, x = np.meshgrid(g.z, g.y, g.x, indexing="ij")
    P = (np.sin if prof == "sin" else np.cos)(np.pi * z / Lz)
    st = _state(g, [0 * x] * 3, [P, 0 * x, 0 * x])
    a0 = st.ax.copy()
    O.mhd_step(g, st, ordr, dt, 0.0, mu, (0.0, 0.0, 0.0), neu, neu2, kind, kind)
    h = -mu * (np.pi / Lz) ** 2 * dt
    fac = sum(h ** j / math.factorial(j) for j in range(ordr + 1))
    got = real(g, st.ax)[: g.np_]
    ref = real(g, a0)[: g.np_]
    assert np.abs(got - fac * ref).max() < tol
    grow = sum((-h) ** j / math.factorial(j) for j in range(ordr + 1))
    assert np.abs(got - grow * ref).max() > 1e-3
    for f in (st.ay, st.az):
        assert np.abs(f).max() < 1e-12 * np.abs(a0).max()
    # j x b = (0, 0, -P'' P') is a gradient on this line: the projection
    # leaves only the FC residual in v
    for f in (st.vx, st.vy, st.vz):
        assert np.abs(f).max() < 1e-6 * np.abs(a0).max()


def _ideal_exchange(flip_lorentz=False, flip_emf=False, dt=1e-3):
    g = O.make_grid(16, 16, 89, 25, 5, TABLES)
    neu, neu2 = O.neumann_vector(g, TABLES, 1), O.neumann_vector(g, TABLES, 2)
    z, y, x = np.meshgrid(g.z, g.y, g.x, indexing="ij")
    s = np.sin(np.pi * z) ** 2
    A = 0.7
    st = _state(g, [s * np.sin(y), s * np.sin(x), 0 * x],
                [A * np.sin(np.pi * z) * np.cos(y) * np.cos(x),
                 A * np.sin(np.pi * z) * np.sin(x + y), A * np.cos(np.pi * z) * np.cos(x)])
    O.a_imposebc_and_project(g, neu, neu2, st.ax, st.ay, st.az, st.ph, 0, 0)
    O.v_imposebc_and_project(g, st.vx, st.vy, st.vz, st.pr, 1, 1)
    E = lambda: (O.energy(g, st.vx, st.vy, st.vz, 1), O.energy(g, st.ax, st.ay, st.az, 0))  # noqa: E731
    ek0, em0 = E()
    orig = O.vector

    def patched(g_, a, b, c, d, e, f):
        out = orig(g_, a, b, c, d, e, f)
        lorentz = a is st.ax
        if (lorentz and flip_lorentz) or (not lorentz and flip_emf):
            out = tuple(-w for w in out)
        return out

    O.vector = patched
    try:
        O.mhd_step(g, st, 4, dt, 0.0, 0.0, (0.0, 0.0, 0.0), neu, neu2, 0, 0)
    finally:
        O.vector = orig
    ek1, em1 = E()
    return ek1 - ek0, em1 - em0


def test_ideal_step_exchanges_kinetic_and_magnetic_energy():
    dk, dm = _ideal_exchange()
    assert abs(dk) > 5e-5
    assert abs(dk + dm) < 0.05 * abs(dk)
    for kw in ({"flip_lorentz": True}, {"flip_emf": True}):
        dk, dm = _ideal_exchange(**kw)
        assert abs(dk + dm) > 0.5 * max(abs(dk), abs(dm)), kw


def test_energy_of_curl_curl():
    """energy(kin=2) is <|curl curl a|^2>: for z-independent a = (cos 2y,
    sin 3x, 0), curl curl a = (4 cos 2y, 9 sin 3x, 0)."""
    g = O.make_grid(16, 16, 41, 25, 5, TABLES)
    z, y, x = np.meshgrid(g.z, g.y, g.x, indexing="ij")
    ax, ay, zero = np.cos(2 * y), np.sin(3 * x), 0 * x
    S = [spec(g, w) for w in (ax, ay, zero)]
    Z = spec(g, zero)
    ex = O.energy(g, S[0], Z, Z, 1)
    ey = O.energy(g, Z, S[1], Z, 1)
    e2 = O.energy(g, *S, 2)
    assert abs(e2 - (16 * ex + 81 * ey)) < 1e-10 * e2
    # kin=0 is the reference's enstrophy quirk (wy counted twice, wz never):
    # here curl a = (0, 0, 3 cos 3x + 2 sin 2y), so kin=0 sees nothing
    assert abs(O.energy(g, *S, 0)) < 1e-10 * e2
