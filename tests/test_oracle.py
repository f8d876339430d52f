"""Pinning of the float64 oracle (CPU only).

Each check is chosen so that a plausible mistake in the restatement (a dropped
term, a wrong sign, index or mirror, a transposed operand) fails it: several
tests plant such a mistake in a copy of the oracle and require the check to
catch it.  Sources of truth: the DFT definition (brute force), the reference's
own analytic test functions (src/tests/fc_dirichlet.f90, energy.f90,
poisson.f90), closed forms (laplace_z, derivatives of trig fields) and the
FC-Gram tables shipped with the reference.
"""
import numpy as np
import pytest

from conftest import TABLES, rel_err
from oracle import specter_ref as O


def grid(nx, ny, nz, C, d):
    return O.make_grid(nx, ny, nz, C, d, TABLES)


# ---------------------------------------------------------------------------
# tables and grid
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("C,d", [(25, 5), (27, 8)])
def test_tables_q_orthogonal_and_dir(C, d):
    A, Q = O.load_tables(TABLES, C, d)
    assert A.shape == (C, d) and Q.shape == (d, d)
    # Q holds an orthonormal (Gram polynomial) basis of the d matching points
    assert np.abs(Q.T @ Q - np.eye(d)).max() < 1e-13
    D = O.dir_matrix(A, Q)
    assert np.abs(D - A @ Q.T).max() < 1e-12 * np.abs(D).max()
    # a transposed Q (a plausible mistake) gives a different operator
    assert np.abs(A @ Q - D).max() > 1e-3


def test_tables_raw_layout_is_fortran_order():
    raw = np.fromfile(f"{TABLES}/A25-5.dat", dtype="<f8")
    A, _ = O.load_tables(TABLES, 25, 5)
    assert A[1, 0] == raw[1] and A[0, 1] == raw[25]


def test_wavenumbers():
    g = grid(8, 6, 41, 25, 5)
    assert list(g.kx) == [0, 1, 2, 3, -4]           # Nyquist kx is negative (specter.fpp:772-775)
    assert list(g.ky) == [0, 1, 2, -3, -2, -1]
    kz = g.kz / g.Dkz
    assert kz[20] == 20 and kz[21] == -20 and kz[40] == -1   # standard order for odd nz
    assert g.dz == 1.0 / 15 and abs(g.Dkz - 2 * np.pi / (g.dz * 41)) < 1e-15
    g2 = grid(8, 6, 40, 25, 5)
    kz = g2.kz / g2.Dkz
    assert kz[19] == 19 and kz[20] == -20 and kz[39] == -1   # = reference loop for even nz
    assert np.allclose(g.khom, np.sqrt(g.kx[:, None] ** 2 + g.ky[None, :] ** 2), rtol=0, atol=0)


# ---------------------------------------------------------------------------
# transforms against the DFT definition
# ---------------------------------------------------------------------------
def _dft_mat(n, sign):
    j = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(j, j) / n)


def test_r2c_3d_brute_force():
    g = grid(4, 6, 31, 25, 5)
    rng = np.random.default_rng(0)
    R = rng.standard_normal(g.real_shape)
    S = O.fftp3d_real_to_complex(g, R.copy())
    # 2D DFT over (x, y) per plane, keep kx <= nx/2, transpose to (kx, ky, z)
    X = np.einsum("zyx,ix,jy->ijz", R, _dft_mat(4, -1)[: g.nkx], _dft_mat(6, -1))
    O.fc_continue(g, X)
    ref = X @ _dft_mat(31, -1).T
    assert rel_err(S, ref) < 1e-13


def test_c2r_3d_brute_force_hermitian_semantics():
    g = grid(4, 6, 31, 25, 5)
    rng = np.random.default_rng(1)
    S = rng.standard_normal(g.spec_shape) + 1j * rng.standard_normal(g.spec_shape)
    R = O.fftp3d_complex_to_real(g, S.copy())
    A = S @ _dft_mat(31, +1).T                       # z backward, (kx, ky, z)
    A = np.einsum("ijz,yj->iyz", A, _dft_mat(6, +1))  # y backward
    # c2r: Im of kx = 0 and kx = nx/2 ignored, other kx counted twice
    A[0] = A[0].real
    A[2] = A[2].real
    w = np.array([1, 2, 1])
    ref = np.einsum("i,iyz,xi->zyx", w, A, _dft_mat(4, +1)[:, : g.nkx]).real
    assert rel_err(R, ref) < 1e-13


def test_z_transforms_brute_force():
    g = grid(4, 6, 31, 25, 5)
    rng = np.random.default_rng(2)
    S = rng.standard_normal(g.spec_shape) + 1j * rng.standard_normal(g.spec_shape)
    B = S.copy()
    O.fftp1d_complex_to_real_z(g, B)
    assert rel_err(B, S @ _dft_mat(31, +1).T) < 1e-13
    F = S.copy()
    O.fftp1d_real_to_complex_z(g, F)
    Sc = S.copy()
    O.fc_continue(g, Sc)
    assert rel_err(F, Sc @ _dft_mat(31, -1).T) < 1e-13
    # the physical points are untouched by the continuation
    assert np.array_equal(Sc[:, :, : g.np_], S[:, :, : g.np_])


def test_round_trip_physical_planes():
    g = grid(16, 12, 41, 25, 5)
    rng = np.random.default_rng(3)
    R = rng.standard_normal(g.real_shape)
    back = O.fftp3d_complex_to_real(g, O.fftp3d_real_to_complex(g, R.copy())) / (16 * 12 * 41)
    # white noise continues to values ~100x the data (|dir| up to ~1e2), so
    # the round trip is exact to ~1e-16 of the continued line, not of the data
    assert np.abs(back[: g.np_] - R[: g.np_]).max() < 1e-11


# ---------------------------------------------------------------------------
# FC-Gram continuation: the reference's own KAT (tests/fc_dirichlet.f90:14-89)
# ---------------------------------------------------------------------------
def _fc_derivative_error(g, cont=None):
    X = g.x[None, None, :]
    Y = g.y[None, :, None]
    Z = g.z[:, None, None]
    R = np.sin(4 * X) * np.cos(8 * Y) * np.sin(6 * Z) * np.ones(g.real_shape)
    S = O.fftp2d_real_to_complex_xy(g, R)
    (cont or O.fc_continue)(g, S)
    S = S @ _dft_mat(g.nz, -1).T
    r = O.fftp3d_complex_to_real(g, O.derivk(g, S, 3)) / (g.nx * g.ny * g.nz)
    exact = (6 * np.sin(4 * X) * np.cos(8 * Y) * np.cos(6 * Z)) * np.ones(g.real_shape)
    return np.abs(r[: g.np_] - exact[: g.np_]).max()


@pytest.mark.parametrize("nzp,C,d,bound", [(64, 25, 5, 2e-4), (128, 25, 5, 1.5e-5),
                                            (64, 27, 8, 1e-7)])
def test_fc_dirichlet_kat(nzp, C, d, bound):
    g = grid(32, 32, nzp + C, C, d)
    assert _fc_derivative_error(g) < bound


def test_fc_dirichlet_kat_converges():
    e64 = _fc_derivative_error(grid(32, 32, 64 + 25, 25, 5))
    e128 = _fc_derivative_error(grid(32, 32, 128 + 25, 25, 5))
    assert e64 / e128 > 8.0          # high-order convergence in z


def test_fc_planted_mistakes_fail_kat():
    g = grid(32, 32, 64 + 25, 25, 5)
    good = _fc_derivative_error(g)

    def no_mirror(g, S):   # bottom block read with dir(ii,.) instead of dir(C-ii+1,.)
        n, C, d = g.nz, g.Cz, g.d
        D = g.dirm
        for ii in range(C):
            acc = 0
            for jj in range(d):
                acc = acc + D[ii, jj] * S[:, :, n - C - d + jj] + D[ii, jj] * S[:, :, d - 1 - jj]
            S[:, :, n - C + ii] = acc

    def shifted(g, S):     # top matching block off by one point
        n, C, d = g.nz, g.Cz, g.d
        D = g.dirm
        for ii in range(C):
            acc = 0
            for jj in range(d):
                acc = acc + D[ii, jj] * S[:, :, n - C - d - 1 + jj]
                acc = acc + D[C - 1 - ii, jj] * S[:, :, d - 1 - jj]
            S[:, :, n - C + ii] = acc

    for bad in (no_mirror, shifted):
        assert _fc_derivative_error(g, bad) > 100 * good


# ---------------------------------------------------------------------------
# operators
# ---------------------------------------------------------------------------
def test_derivk_laplak_closed_form():
    g = grid(16, 16, 41, 25, 5)
    X, Y = g.x[None, None, :], g.y[None, :, None]
    R = np.sin(3 * X) * np.cos(2 * Y) * np.ones(g.real_shape)
    S = O.fftp3d_real_to_complex(g, R)
    N = g.nx * g.ny * g.nz
    dx = O.fftp3d_complex_to_real(g, O.derivk(g, S, 1)) / N
    dy = O.fftp3d_complex_to_real(g, O.derivk(g, S, 2)) / N
    lap = O.fftp3d_complex_to_real(g, O.laplak(g, S)) / N
    ones = np.ones(g.real_shape)
    p = g.np_
    assert np.abs(dx - 3 * np.cos(3 * X) * np.cos(2 * Y) * ones)[:p].max() < 1e-12
    assert np.abs(dy + 2 * np.sin(3 * X) * np.sin(2 * Y) * ones)[:p].max() < 1e-12
    # z is constant: its FC derivative vanishes to continuation accuracy
    assert np.abs(lap + 13 * R)[:p].max() < 1e-8
    # curl of (a, b, c) with a = b = c: closed form component by component
    c1 = O.fftp3d_complex_to_real(g, O.curlk(g, S, S, 3)) / N
    assert np.abs(c1 - (dx - dy))[:p].max() < 1e-12


def test_gradre_taylor_green():
    """u = (sin x cos y, -cos x sin y, 0): (u.grad)u = (sin 2x, sin 2y, 0)/2."""
    g = grid(32, 32, 41, 25, 5)
    X, Y = g.x[None, None, :], g.y[None, :, None]
    ones = np.ones(g.real_shape)
    u = np.sin(X) * np.cos(Y) * ones
    v = -np.cos(X) * np.sin(Y) * ones
    w = 0 * ones
    S = [O.fftp3d_real_to_complex(g, a) for a in (u, v, w)]
    nl = O.gradre(g, *S)
    N = g.nx * g.ny * g.nz
    r = [O.fftp3d_complex_to_real(g, a) / N for a in nl]
    p = g.np_
    assert np.abs(r[0] - 0.5 * np.sin(2 * X) * ones)[:p].max() < 1e-9
    assert np.abs(r[1] - 0.5 * np.sin(2 * Y) * ones)[:p].max() < 1e-9
    assert np.abs(r[2])[:p].max() < 1e-9


def test_filter_factors():
    g = grid(16, 16, 41, 25, 5)
    fx, fy, fz = O.filter_factors(g)
    assert fx[0] == 1.0 and fy[0] == 1.0 and fz[0] == 1.0
    # exp(-16 ln 10 * 1) = 1e-16 at the x/y Nyquist (|k| = n/2)
    assert abs(fx[-1] - 1e-16) < 1e-28 and abs(fy[8] - 1e-16) < 1e-28
    assert np.all(np.diff(fx[:-1]) <= 0)


# ---------------------------------------------------------------------------
# boundary conditions and projection (tests/poisson.f90)
# ---------------------------------------------------------------------------
def test_laplace_z_neumann_closed_form():
    g = grid(16, 16, 41, 25, 5)
    rng = np.random.default_rng(4)
    bc = rng.standard_normal((g.nkx, g.ny, 2)) + 1j * rng.standard_normal((g.nkx, g.ny, 2))
    a, b = O.laplace_z(g, bc, 1, 1)
    p = g.np_
    m = np.ones((g.nkx, g.ny), bool)
    m[0, 0] = False
    assert np.abs(b[:, :, 0] - bc[:, :, 0])[m].max() < 1e-12
    assert np.abs(b[:, :, p - 1] - bc[:, :, 1])[m].max() < 1e-12
    # a'' = khom^2 a: second difference on the physical grid
    h = g.dz
    d2 = (a[:, :, 2:p] - 2 * a[:, :, 1:p - 1] + a[:, :, : p - 2]) / h ** 2
    res = d2 - (g.khom ** 2)[:, :, None] * a[:, :, 1:p - 1]
    sc = (g.khom ** 2)[:, :, None] * np.abs(a[:, :, 1:p - 1]) + 1e-30
    small = (g.khom < 4) & m
    assert np.abs(res / sc)[small].max() < 0.05
    # zero mode is the linear profile bc1*z
    assert np.allclose(a[0, 0, :], bc[0, 0, 0].real * g.z, rtol=0, atol=1e-14)


def _poisson_field(g):
    X, Y, Z = g.x[None, None, :], g.y[None, :, None], g.z[:, None, None]
    R = np.sin(4 * X) * np.cos(8 * Y) * np.exp(0.4 * Z / g.Lz) * np.ones(g.real_shape)
    R[g.np_:] = 0
    return [O.fftp3d_real_to_complex(g, R.copy()) for _ in range(3)]


def test_sol_project_invariants():
    """tests/poisson.f90:60-83: after sol_project(bctarget=1) the field is
    solenoidal and its normal component vanishes at both walls."""
    g = grid(32, 32, 64 + 25, 25, 5)
    a, b, c = _poisson_field(g)
    d0 = O.divergence(g, a, b, c)
    pr = np.zeros(g.spec_shape, complex)
    O.sol_project(g, a, b, c, pr, 1, 0, 0)
    d1 = O.divergence(g, a, b, c)
    nb, nt = O.bouncheck_z(g, c)
    e = O.energy(g, a, b, c, 1)
    assert d1 < 1e-8 * d0
    assert nb < 1e-12 * e and nt < 1e-12 * e


def test_sol_project_planted_sign_error_fails():
    g = grid(32, 32, 64 + 25, 25, 5)
    a, b, c = _poisson_field(g)
    orig = O.poisson_inhomogeneous

    def wrong(g_, a_, b_, c_):
        return -orig(g_, a_, b_, c_)

    O.poisson_inhomogeneous = wrong
    try:
        pr = np.zeros(g.spec_shape, complex)
        d0 = O.divergence(g, a, b, c)
        O.sol_project(g, a, b, c, pr, 1, 0, 0)
        assert O.divergence(g, a, b, c) > 1e-3 * d0
    finally:
        O.poisson_inhomogeneous = orig


def test_noslip_walls():
    g = grid(16, 16, 41, 25, 5)
    rng = np.random.default_rng(5)
    vx, vy, vz, pr = (rng.standard_normal(g.spec_shape) + 1j * rng.standard_normal(g.spec_shape)
                      for _ in range(4))
    pr_mixed = pr.copy()
    O.goto_domain_w_boundaries(g, vx, vy)
    O.noslip_z(g, 1, 2, vx, vy, pr_mixed, (0.25, -0.5), 0)
    # tau = (o+1)/o = 2 for o=1 < ord=2
    ref = 2.0 * (1j * g.kx[:, None] * pr_mixed[:, :, 0])
    ref[0, 0] = 16 * 16 * 0.25
    assert np.abs(vx[:, :, 0] - ref).max() < 1e-12
    assert vy[0, 0, 0] == 16 * 16 * -0.5


# ---------------------------------------------------------------------------
# diagnostics (tests/energy.f90)
# ---------------------------------------------------------------------------
def test_energy_parseval_kat():
    g = grid(32, 32, 64 + 25, 25, 5)
    X, Y, Z = g.x[None, None, :], g.y[None, :, None], g.z[:, None, None]
    R = np.sin(4 * X) * np.cos(8 * Y) * np.sin(6 * Z) * np.ones(g.real_shape)
    S = O.fftp3d_real_to_complex(g, R.copy())
    e = O.energy(g, S, S, S, 1)
    ref = 3 * np.mean(R[: g.np_] ** 2)
    assert abs(e - ref) / ref < 1e-13


def test_enstrophy_reference_quirk_and_cross():
    g = grid(32, 32, 41, 25, 5)
    X, Y = g.x[None, None, :], g.y[None, :, None]
    ones = np.ones(g.real_shape)
    u = np.sin(X) * np.cos(Y) * ones
    v = -np.cos(X) * np.sin(Y) * ones
    w = np.cos(2 * X) * ones
    S = [O.fftp3d_real_to_complex(g, a) for a in (u, v, w)]
    # omega = (dy w - dz v, dz u - dx w, dx v - dy u) = (0, 2 sin 2x, 2 sin x sin y)
    p = g.np_
    wy2 = np.mean((2 * np.sin(2 * X) * ones)[:p] ** 2)
    # energy(kin=0) sums wx^2 + wy^2 + wy^2 (pseudospec_hd.f90:527,540)
    assert abs(O.energy(g, *S, 0) - 2 * wy2) / wy2 < 1e-9
    # cross(v, v) = <|v|^2> = energy(kin=1)
    assert abs(O.cross(g, *S, *S) - O.energy(g, *S, 1)) < 1e-14


def _kz0_spectral(g, R2):
    """Spectral array of the z-independent field R2(y, x): only the kz = 0
    coefficient of each line, nz times its 2D DFT (no continuation involved,
    so z derivatives vanish exactly)."""
    X = np.fft.rfft2(R2)                              # (ny, nkx)
    S = np.zeros(g.spec_shape, complex)
    S[:, :, 0] = g.nz * X.T
    return S


@pytest.mark.parametrize("k", [1, 2, 3])
def test_helicity_beltrami(k):
    """v = (sin ky, cos kx, cos ky + sin kx) has curl v = -k v, so
    <v . curl v> = -k <|v|^2>; a swapped component pair or a lost sign breaks it."""
    g = grid(32, 24, 41, 25, 5)
    X, Y = g.x[None, :], g.y[:, None]
    one = np.ones((g.ny, g.nx))
    u, v, w = np.sin(k * Y) * one, np.cos(k * X) * one, (np.cos(k * Y) + np.sin(k * X)) * one
    S = [_kz0_spectral(g, a) for a in (u, v, w)]
    e = O.energy(g, *S, 1)
    assert abs(e - 2.0) < 1e-13
    assert abs(O.helicity(g, *S) - (-k * e)) < 1e-12 * k
    # planted mistakes: the enstrophy quirk's pairing (w_y twice, no w_z) and a
    # flipped curl sign both give other values
    S1 = [_kz0_spectral(g, a) for a in (u, v, 0 * w)]
    bad = (O.cross(g, *S, O.curlk(g, S[1], S[2], 1), O.curlk(g, S[0], S[2], 2),
                   O.curlk(g, S[0], S[2], 2)))
    assert abs(bad - (-k * e)) > 0.1
    assert abs(O.helicity(g, *S1) - (-k * O.energy(g, *S1, 1))) > 0.1 or \
        abs(O.helicity(g, S[0], S[1], -S[2]) - (-k * e)) > 0.1


def test_helicity_parseval_kat():
    """Mixed-domain helicity equals the physical-space mean of v . w with w
    the curl computed from the same spectral fields (tests/energy.f90:20-32
    style: Parseval over x, y and physical z)."""
    g = grid(24, 16, 57, 25, 5)
    rng = np.random.default_rng(3)
    S = [O.fftp3d_real_to_complex(g, rng.standard_normal(g.real_shape)) for _ in range(3)]
    for a in S:  # no Nyquist content (the x Nyquist column carries weight 2)
        a[g.nkx - 1] = 0.0
        a[:, g.ny // 2] = 0.0
    W = [O.curlk(g, S[1], S[2], 1), O.curlk(g, S[0], S[2], 2), O.curlk(g, S[0], S[1], 3)]
    Rv = [O.fftp3d_complex_to_real(g, a.copy()) for a in S]
    Rw = [O.fftp3d_complex_to_real(g, a.copy()) for a in W]
    n = float(g.nx * g.ny * g.nz)
    ref = sum(np.mean(rv[: g.np_] * rw[: g.np_]) for rv, rw in zip(Rv, Rw)) / n ** 2
    assert abs(O.helicity(g, *S) - ref) < 1e-12 * abs(ref) + 1e-12


def test_divergence_kat():
    """tests/poisson.f90:36-55: numeric vs analytic mean-square divergence."""
    g = grid(32, 32, 64 + 25, 25, 5)
    X, Y, Z = g.x[None, None, :], g.y[None, :, None], g.z[:, None, None]
    a, b, c = _poisson_field(g)
    div = np.exp(0.4 * Z) * (4 * np.cos(4 * X) * np.cos(8 * Y) - 8 * np.sin(4 * X)
                             * np.sin(8 * Y) + 0.4 * np.sin(4 * X) * np.cos(8 * Y)) * np.ones(g.real_shape)
    ref = np.mean(div[: g.np_] ** 2)
    assert abs(O.divergence(g, a, b, c) - ref) / ref < 1e-6


# ---------------------------------------------------------------------------
# scalar (pseudospec_phd.f90) and wall diagnostics
# ---------------------------------------------------------------------------
def test_variance_product_closed_forms():
    g = grid(32, 32, 64 + 25, 25, 5)
    X, Y, Z = g.x[None, None, :], g.y[None, :, None], g.z[:, None, None]
    ones = np.ones(g.real_shape)
    th = np.cos(2 * X) * np.cos(3 * Y) * ones           # constant in z
    S = O.fftp3d_real_to_complex(g, th.copy())
    zero = np.zeros_like(S)
    # kin=1 is the one-component energy; product(a, a) is the same sum
    assert O.variance(g, S, 1) == O.energy(g, S, zero, zero, 1)
    assert abs(O.product(g, S, S) - O.variance(g, S, 1)) < 1e-15
    assert abs(O.variance(g, S, 1) - 0.25) < 1e-12
    # kin=0: kk2 = 2^2 + 3^2 on a z-constant field (its continuation is flat)
    assert abs(O.variance(g, S, 0) - 169 * 0.25) / (169 * 0.25) < 1e-6
    # product of two different modes vanishes, of a mode with itself scaled
    other = O.fftp3d_real_to_complex(g, (np.sin(X) * ones).copy())
    assert abs(O.product(g, S, other)) < 1e-14
    assert abs(O.product(g, S, 3 * S) - 0.75) < 1e-12
    # pscheck = (variance 1, variance 0, product)
    assert O.pscheck(g, S, other) == (O.variance(g, S, 1), O.variance(g, S, 0),
                                      O.product(g, S, other))


def test_wall_shear_nusselt_divmax_closed_forms():
    g = grid(32, 32, 64 + 25, 25, 5)
    X, Y, Z = g.x[None, None, :], g.y[None, :, None], g.z[:, None, None]
    ones = np.ones(g.real_shape)
    Lz = g.Lz
    # u_x = 2 z (1 + 0.3 cos x): mean shear nu*2 at both walls (cos x averages out)
    u = 2 * Z * (1 + 0.3 * np.cos(X)) * ones
    v = -0.5 * Z * Z * ones                             # d/dz = -z: 0 at bottom, -Lz at top
    Su = O.fftp3d_real_to_complex(g, u.copy())
    Sv = O.fftp3d_real_to_complex(g, v.copy())
    tb_x, tb_y, tt_x, tt_y = O.wall_shear(g, Su, Sv, 0.01)
    assert abs(tb_x - 0.02) < 1e-8 and abs(tt_x - 0.02) < 1e-8
    assert abs(tb_y) < 1e-8 and abs(tt_y + 0.01 * Lz) < 1e-8
    # theta' = a z (Lz - z): d/dz = a Lz at z=0, -a Lz at z=Lz
    a = 0.7
    th = a * Z * (Lz - Z) * np.cos(0 * X) * ones
    nb, nt = O.nusselt(g, O.fftp3d_real_to_complex(g, th.copy()))
    assert abs(nb - (1 - a * Lz * Lz)) < 1e-8 and abs(nt - (1 + a * Lz * Lz)) < 1e-8
    # div of (sin 2x, 0, 0) is 2 cos 2x: max 2
    S1 = O.fftp3d_real_to_complex(g, (np.sin(2 * X) * ones).copy())
    zero = np.zeros_like(S1)
    assert abs(O.divergence_max(g, S1, zero, zero) - 2.0) < 1e-12
    # a planted sign error in the z derivative flips the shear
    assert O.wall_shear(g, -Su, Sv, 0.01)[0] < 0


def test_energy_spectrum_h():
    g = grid(32, 32, 41, 25, 5)
    X, Y = g.x[None, None, :], g.y[None, :, None]
    ones = np.ones(g.real_shape)
    S = O.fftp3d_real_to_complex(g, (np.cos(3 * X) * ones).copy())
    zero = np.zeros_like(S)
    sp = O.energy_spectrum_h(g, S, zero, zero, 24)
    assert abs(sp[3] - 0.5) < 1e-13 and abs(sp.sum() - sp[3]) < 1e-13
    # every shell: the bins add up to the energy
    rng = np.random.default_rng(3)
    comps = [O.fftp3d_real_to_complex(g, rng.standard_normal(g.real_shape)) for _ in range(3)]
    sp = O.energy_spectrum_h(g, *comps, 40)
    assert abs(sp.sum() - O.energy(g, *comps, 1)) / O.energy(g, *comps, 1) < 1e-12
