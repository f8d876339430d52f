"""Pinning of the oracle's non-periodic y transforms (CPU only): the 3D
transforms with an FC-Gram continuation along y as well as z
(fftp.fpp:388-425 with Cy > 0: fftp1d_real_to_complex_x + fftp2d_real_to_complex_yz,
fftp2d_complex_to_real_yz + fftp1d_complex_to_real_x; grid specter.fpp:705-723).

* the reference's fc_dirichlet KAT (tests/fc_dirichlet.f90) in y: spectral
  x, y and z derivatives of sin(4x) cos(8y) sin(6z) on the physical points,
  converging at high order in the number of physical y points;
* brute force on tiny shapes: the continued field built from the
  continuation matrices (y first on the physical z planes, then z), then a
  dense 3D DFT;
* round trip r2c -> c2r = N on the physical region;
* a planted mirror mistake in the y continuation fails the KAT.
"""
import numpy as np
import pytest

from conftest import TABLES
from oracle import specter_ref as O


def grid(nx, npy, npz, Cy=25, oy=5, Cz=25, oz=5):
    return O.make_grid(nx, npy + Cy, npz + Cz, Cz, oz, TABLES, Cy=Cy, oy=oy)


def field(g, fn):
    R = np.zeros(g.real_shape)
    z, y, x = np.meshgrid(g.z, g.y, g.x, indexing="ij")
    R[:, :, :] = fn(x, y, z)
    return R


def kat_errors(g):
    R = field(g, lambda x, y, z: np.sin(4 * x) * np.cos(8 * y) * np.sin(6 * z))
    S = O.fftp3d_real_to_complex(g, R.copy())
    exact = {1: lambda x, y, z: 4 * np.cos(4 * x) * np.cos(8 * y) * np.sin(6 * z),
             2: lambda x, y, z: -8 * np.sin(4 * x) * np.sin(8 * y) * np.sin(6 * z),
             3: lambda x, y, z: 6 * np.sin(4 * x) * np.cos(8 * y) * np.cos(6 * z)}
    out = {}
    for dirn, fn in exact.items():
        d = O.fftp3d_complex_to_real(g, O.derivk(g, S, dirn)) / (g.nx * g.ny * g.nz)
        ref = field(g, fn)
        sl = (slice(0, g.np_), slice(0, g.npy), slice(None))
        out[dirn] = np.abs(d[sl] - ref[sl]).max() / np.abs(ref[sl]).max()
    return out


def test_grid_non_periodic_y():
    g = grid(8, 40, 40)
    assert g.ny == 65 and g.npy == 40
    dy = 1.0 / (40 - 1)
    assert np.allclose(g.y[:40], dy * np.arange(40)) and abs(g.y[39] - 1.0) < 1e-15
    assert abs(g.Dky - 2 * np.pi / (dy * 65)) < 1e-12
    # standard DFT order for odd ny: 0..32, then -32..-1
    m = np.round(g.ky / g.Dky).astype(int)
    assert list(m[:33]) == list(range(33)) and list(m[33:]) == list(range(-32, 0))


def test_y_continuation_kat_converges():
    """Derivative errors on the physical points (relative to the derivative's
    maximum) at 64 and 128 physical points in y and z: all three are small
    and the y and z errors drop by more than 10x per doubling (fourth-order
    convergence of the d = 5 continuation gives about 16x); x is spectral."""
    e64 = kat_errors(grid(16, 64, 64))
    e128 = kat_errors(grid(16, 128, 128))
    print(e64, e128)
    for dirn in (1, 2, 3):
        assert e64[dirn] < 1e-4 and e128[dirn] < 1e-4, (dirn, e64[dirn], e128[dirn])
    for dirn in (2, 3):
        assert e64[dirn] > 10 * e128[dirn], dirn
    assert e64[1] < 1e-13 and e128[1] < 1e-13


def _cont_matrix(n_phys, C, d):
    A, Q = O.load_tables(TABLES, C, d)
    D = A @ Q.T
    E = np.zeros((n_phys + C, n_phys))
    E[:n_phys] = np.eye(n_phys)
    for ii in range(C):
        for jj in range(d):
            E[n_phys + ii, n_phys - d + jj] += D[ii, jj]
            E[n_phys + ii, d - 1 - jj] += D[C - 1 - ii, jj]
    return E


@pytest.mark.parametrize("nx,npy,npz", [(4, 6, 7), (6, 9, 5)])
def test_brute_force_tiny(nx, npy, npz):
    g = grid(nx, npy, npz)
    rng = np.random.default_rng(3)
    R = np.zeros(g.real_shape)
    R[:npz, :npy, :] = rng.standard_normal((npz, npy, nx))
    S = O.fftp3d_real_to_complex(g, R.copy())
    Ey, Ez = _cont_matrix(npy, 25, 5), _cont_matrix(npz, 25, 5)
    F = R[:npz, :npy, :]                                  # (z, y, x)
    H = np.einsum("zc,cyx->zyx", Ez, np.einsum("yb,zbx->zyx", Ey, F))
    # dense DFTs: x (half spectrum), y, z
    def W(n):
        m = np.arange(n)
        return np.exp(-2j * np.pi * np.outer(m, m) / n)
    X = np.einsum("kx,zyx->zyk", W(nx)[: nx // 2 + 1], H)
    Y = np.einsum("jy,zyk->zjk", W(g.ny), X)
    Z = np.einsum("lz,zjk->kjl", W(g.nz), Y)
    assert np.abs(S - Z).max() <= 1e-12 * np.abs(Z).max()


def test_round_trip():
    g = grid(8, 20, 24)
    rng = np.random.default_rng(4)
    R = rng.standard_normal(g.real_shape)
    back = O.fftp3d_complex_to_real(g, O.fftp3d_real_to_complex(g, R.copy()))
    back /= g.nx * g.ny * g.nz
    sl = (slice(0, g.np_), slice(0, g.npy), slice(None))
    # random data continues into values ~1e7 in the (y, z) corner, so the
    # rounding scale is the continued field's maximum, not the input's
    assert np.abs(back[sl] - R[sl]).max() < 1e-14 * np.abs(back).max()


def test_planted_mirror_mistake_fails(monkeypatch):
    good = kat_errors(grid(16, 64, 64))[2]

    def bad(g, S):
        ny, C, d, p = g.ny, g.Cy, g.oy, g.np_
        D = g.diry
        top = ny - C - d
        for ii in range(C):
            acc = D[ii, 0] * S[:, top, :p] + D[C - 1 - ii, 0] * S[:, top, :p]
            for jj in range(1, d):
                acc = acc + D[ii, jj] * S[:, top + jj, :p]
                acc = acc + D[C - 1 - ii, jj] * S[:, top + jj, :p]
            S[:, ny - C + ii, :p] = acc

    monkeypatch.setattr(O, "fc_continue_y", bad)
    worse = kat_errors(grid(16, 64, 64))[2]
    assert worse > 100 * good, (good, worse)


def test_backward_pieces_compose_to_3d():
    """fftp2d_complex_to_real_yz + fftp1d_complex_to_real_x (fftp.fpp:922-1056)
    equal the 3D backward transform (z IFFT + 2D xy c2r, :817-818), and
    fftp1d_real_to_complex_x + fftp2d_real_to_complex_yz the forward one."""
    g = grid(8, 20, 24)
    rng = np.random.default_rng(5)
    R = rng.standard_normal(g.real_shape)
    S = O.fftp3d_real_to_complex(g, R.copy())
    S2 = O.fftp1d_real_to_complex_x(g, R.copy())
    O.fftp2d_real_to_complex_yz(g, S2)
    assert np.abs(S2 - S).max() <= 1e-14 * np.abs(S).max()
    B = O.fftp3d_complex_to_real(g, S.copy())
    T = S.copy()
    O.fftp2d_complex_to_real_yz(g, T)
    B2 = O.fftp1d_complex_to_real_x(g, T)
    assert np.abs(B2 - B).max() <= 1e-14 * np.abs(B).max()
