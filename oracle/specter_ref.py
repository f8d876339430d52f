"""Float64 CPU restatement of SPECTER's per-RK-substep hot path (oracle).

TEST INFRASTRUCTURE ONLY -- imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.

Layout convention (byte-identical to the reference's Fortran arrays):
  spectral  Fortran (nz, ny, nkx) -> numpy (nkx, ny, nz) complex128, index [i, j, k]
  real      Fortran (nx, ny, nz)  -> numpy (nz, ny, nx)  float64,    index [k, j, i]
  mixed (z,ky,kx) domain = spectral layout after a z-only inverse transform.

All transforms are unnormalised, as in the reference (callers scale).

Pinning of each function (tests/test_oracle_*.py):
  * Grid, wavenumbers          closed form (specter.fpp:683-800) + DFT frequency order
  * dir_matrix                 Q orthogonal, A*Q^T reproduced from the raw tables; the
                               FC extension of a smooth function is spectrally accurate
  * fc_continue                FC-Gram derivative KAT of sin(4x)cos(8y)sin(6z)
                               (tests/fc_dirichlet.f90:14-89) -- a wrong index or a
                               missing mirror destroys the convergence
  * fftp* transforms           brute-force DFT on tiny shapes (sign, axes, layout,
                               Hermitian handling of c2r), round trip = N * identity
  * derivk/laplak/curlk        closed form on trig fields
  * gradre / advect            analytic (v.grad)v of trig fields (periodic z limit via
                               the continuation), products only on physical planes
  * poisson/laplace_z/sol_project  invariants: div ~ 0 after projection
                               (tests/poisson.f90:70-76), v_n ~ 0 at walls (:78-83),
                               laplace_z satisfies d2/dz2 - khom^2 = 0 and the
                               Neumann data in closed form
  * energy/cross/divergence    Parseval vs real-space sums (tests/energy.f90:20-32)
  * helicity                   Beltrami fields: <v.curl v> = -k <|v|^2> exactly
  * hd_substep / bouss_substep composition of the above in the order of include/hd/hd_rkstep2.f90
    hd_step / bouss_step        and include/bouss/bouss_rkstep2.f90, pinned by
                               tests/test_oracle_steps.py: a step on an eigenmode of the
                               wall-projected Laplacian is the Taylor polynomial of its
                               eigenvalue; tangential wall slip is O(dt^2) only with the
                               pressure extrapolation factor of vboundary.f90:195-196; the
                               Boussinesq rest state is hydrostatic; planted composition
                               mistakes fail
  * rotbouss_substep / _step    tests/test_oracle_steps.py: with Omega = Omega_z e_z the
    coriolis_terms              horizontal mean flow is an inertial oscillation -- the line
                               follows a matrix model built from the DFT, the continuation
                               and the filter to 1e-9 and the RK polynomial of the rotation
                               to FC accuracy (a sign, order or filter mistake fails); the
                               Coriolis terms are 2 Omega x v (numpy cross product); with
                               Omega = 0 the stage plus the th filter and round trip is the
                               Boussinesq stage bit for bit
  * sol_project pressure d      the removed velocity is (i kx d, i ky d, d/dz d)
  * non-periodic y (fc_continue_y and the x / yz transforms)  tests/test_oracle_ycont.py:
                               the fc_dirichlet KAT with y continued, brute-force DFT of the
                               continued field, round trip, planted mirror mistake
  * MHD: neumann_vector,       tests/test_oracle_mhd.py: the Neumann vectors equal the exact
    neumann/robin_reconstruct, one-sided finite-difference weights (outward first and
    laplace_z (0,0)/(2,2)/(0,2), second derivative) of the degree d-1 interpolant; polynomial
    sol_project (potential),   walls reconstruct exactly; grad psi is removed exactly by the
    a_imposebc_and_project,    matching potential kinds only, with psi returned; laplace_z
    prodre / vector,           meets its boundary data and a'' = khom^2 a; prodre = gradre -
    mhd_substep / mhd_step     grad |v|^2/2; vector = numpy cross; wall values after the
                               magnetic BCs satisfy the discrete conditions; resistive decay of
                               a line mode; kinetic and magnetic energy exchange in an ideal
                               step (a flipped Lorentz or electromotive force fails)
  * precision="longdouble"      the same routines in extended precision (constants formed
                               in float64): the exact-trajectory referee of the float64 runs

Parity notes (deliberate, documented):
  * kz ordering for odd nz: the reference loop specter.fpp:783-786 leaves kz(nz)
    uninitialised and shifts the negative half by one when nz is odd.  This
    restatement uses the standard DFT order kz[m] = m if m < (nz+1)//2 else m-nz,
    which is identical to the reference for every even nz.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np
import scipy.fft as sfft

PI = 3.14159265358979323846  # pseudospec_mod.fpp:71 (module var)


def _workers() -> int:
    return int(os.environ.get("SPX_ORACLE_WORKERS", str(os.cpu_count() or 1)))


class _NumpyFFT:
    """numpy.fft backend with the scipy.fft call signature (used to measure the
    spread between two independent float64 DFT implementations)."""

    @staticmethod
    def fft(x, axis=-1, workers=None):
        return np.fft.fft(x, axis=axis)

    @staticmethod
    def ifft(x, axis=-1, norm=None, workers=None):
        return np.fft.ifft(x, axis=axis, norm=norm)

    @staticmethod
    def rfftn(x, axes=None, workers=None):
        return np.fft.rfftn(x, axes=axes)

    @staticmethod
    def irfftn(x, s=None, axes=None, norm=None, workers=None):
        return np.fft.irfftn(x, s=s, axes=axes, norm=norm)


if os.environ.get("SPX_ORACLE_FFT", "scipy") == "numpy":
    sfft = _NumpyFFT  # noqa: F811


# ---------------------------------------------------------------------------
# FC-Gram tables  (fcgram_mod.f90:180-257)
# ---------------------------------------------------------------------------
def load_tables(tdir: str, C: int, d: int):
    """Read A{C}-{d}.dat (C x d) and Q{d}.dat (d x d): raw float64 stream in
    Fortran order (fcgram_mod.f90:221-231, 238-244)."""
    A = np.fromfile(os.path.join(tdir, f"A{C}-{d}.dat"), dtype="<f8")
    Q = np.fromfile(os.path.join(tdir, f"Q{d}.dat"), dtype="<f8")
    if A.size != C * d or Q.size != d * d:
        raise ValueError("FC-Gram table size mismatch")
    return A.reshape(d, C).T.copy(), Q.reshape(d, d).T.copy()


def dir_matrix(A: np.ndarray, Q: np.ndarray) -> np.ndarray:
    """dir = A * Q^T  (fcgram_mod.f90:250-252), always float64 (:26), as the
    textbook sum over k in increasing order."""
    C, d = A.shape
    D = np.zeros((C, d))
    for k in range(d):
        D = D + A[:, k][:, None] * Q[:, k][None, :]
    return D


# ---------------------------------------------------------------------------
# Grid and wavenumbers  (specter.fpp:683-800; pseudospec_mod.fpp:13-91)
# ---------------------------------------------------------------------------
@dataclass
class Grid:
    nx: int
    ny: int
    nz: int          # total z points including the Cz continuation points
    Cz: int
    d: int
    Lx: float = 1.0  # periodic lengths in units of 2*pi
    Ly: float = 1.0
    Lz: float = 1.0  # physical wall-to-wall length
    dirm: np.ndarray = field(default=None, repr=False)
    # arithmetic type of the run: float64 (the reference's) or longdouble (an
    # extended-precision referee of the same discrete problem: every constant
    # is formed in float64 as the reference forms it, then held exactly in
    # longdouble, and the arithmetic -- transforms included -- runs in
    # longdouble)
    precision: str = "float64"
    # non-periodic y (FC-Gram continuation along y in the 3D transforms):
    # Cy continuation points, matching order oy; Ly is then the physical length
    Cy: int = 0
    oy: int = 0
    diry: np.ndarray = field(default=None, repr=False)

    def __post_init__(self):
        nx, ny, nz, Cz = self.nx, self.ny, self.nz, self.Cz
        self.nkx = nx // 2 + 1
        self.np_ = nz - Cz
        self.npy = ny - self.Cy
        # specter.fpp:683-689 (x periodic), 705-723 (y periodic or FC), 743-749 (z FC)
        rmp = self.Lx * 2.0 * PI / nx
        self.Dkx = 1.0 / self.Lx
        self.x = rmp * np.arange(nx, dtype=np.float64)
        if self.Cy == 0:
            rmp = self.Ly * 2.0 * PI / ny
            self.Dky = 1.0 / self.Ly
        else:
            rmp = self.Ly / (ny - self.Cy - 1)
            self.Dky = 2.0 * PI / (rmp * ny)
        self.y = rmp * np.arange(ny, dtype=np.float64)
        rmp = self.Lz / (nz - Cz - 1)
        self.dz = rmp
        self.Dkz = 2.0 * PI / (rmp * nz)
        self.z = rmp * np.arange(nz, dtype=np.float64)
        # specter.fpp:772-789: FFT order, Nyquist of even n is negative.
        kx = np.arange(nx, dtype=np.float64)
        kx[nx // 2:] -= nx
        ky = np.arange(ny, dtype=np.float64)
        ky[(ny + 1) // 2:] -= ny  # parity note: standard order for odd ny
        kz = np.arange(nz, dtype=np.float64)
        kz[(nz + 1) // 2:] -= nz  # parity note: standard order for odd nz
        self.kx = (kx * self.Dkx)[: self.nkx]
        self.ky = ky * self.Dky
        self.kz = kz * self.Dkz
        # specter.fpp:791-800: kk2 = kx^2+ky^2+kz^2 ; khom = sqrt(kk2(1,j,i))
        kxy = self.kx[:, None] ** 2 + self.ky[None, :] ** 2
        self.kk2 = kxy[:, :, None] + (self.kz ** 2)[None, None, :]
        self.khom = np.sqrt(self.kk2[:, :, 0])
        if self.precision not in ("float64", "longdouble"):
            raise ValueError(f"unknown oracle precision {self.precision!r}")
        self.rdtype = np.float64 if self.precision == "float64" else np.longdouble
        self.cdtype = np.complex128 if self.precision == "float64" else np.clongdouble
        for name in ("x", "y", "z", "kx", "ky", "kz", "kk2", "khom"):
            setattr(self, name, getattr(self, name).astype(self.rdtype))

    @property
    def spec_shape(self):
        return (self.nkx, self.ny, self.nz)

    @property
    def real_shape(self):
        return (self.nz, self.ny, self.nx)


def make_grid(nx, ny, nz, Cz, d, tdir, Lx=1.0, Ly=1.0, Lz=1.0,
              precision="float64", Cy=0, oy=0) -> Grid:
    """fcgram_create_plan (fcgram_mod.f90:49-144) + grid setup.  dir = A*Q^T
    is formed in float64 in both precisions (fcgram_mod.f90:26); with Cy > 0
    the y tables A{Cy}-{oy}, Q{oy} as well (:136-151)."""
    g = Grid(nx, ny, nz, Cz, d, Lx, Ly, Lz, precision=precision, Cy=Cy, oy=oy)
    A, Q = load_tables(tdir, Cz, d)
    g.dirm = dir_matrix(A, Q).astype(g.rdtype)
    if Cy > 0:
        Ay, Qy = load_tables(tdir, Cy, oy)
        g.diry = dir_matrix(Ay, Qy).astype(g.rdtype)
    return g


# ---------------------------------------------------------------------------
# Transforms  (fftp.fpp)
# ---------------------------------------------------------------------------
def fc_continue(g: Grid, S: np.ndarray) -> None:
    """FC-Gram Dirichlet continuation of every z line, in place
    (fftp.fpp:755-773).  Overwrites the C continuation slots from the first d
    and the last d physical points, accumulating in the reference's order."""
    n, C, d = g.nz, g.Cz, g.d
    D = g.dirm
    top = n - C - d   # 0-based start of the top matching block
    for ii in range(C):
        acc = D[ii, 0] * S[:, :, top] + D[C - 1 - ii, 0] * S[:, :, d - 1]
        for jj in range(1, d):
            acc = acc + D[ii, jj] * S[:, :, top + jj]
            acc = acc + D[C - 1 - ii, jj] * S[:, :, d - 1 - jj]
        S[:, :, n - C + ii] = acc


def fftp1d_real_to_complex_z(g: Grid, S: np.ndarray) -> None:
    """Continuation then forward z DFT in place (fftp.fpp:720-786)."""
    fc_continue(g, S)
    S[...] = sfft.fft(S, axis=2, workers=_workers())


def fftp1d_complex_to_real_z(g: Grid, S: np.ndarray) -> None:
    """Unnormalised backward z DFT in place (fftp.fpp:1060-1094)."""
    S[...] = sfft.ifft(S, axis=2, norm="forward", workers=_workers())


def fftp2d_real_to_complex_xy(g: Grid, R: np.ndarray) -> np.ndarray:
    """Per-plane 2D r2c over (x,y) then transpose to (z,ky,kx)
    (fftp.fpp:428-524)."""
    X = sfft.rfftn(R, axes=(1, 2), workers=_workers())      # (nz, ny, nkx)
    return np.ascontiguousarray(X.transpose(2, 1, 0))        # (nkx, ny, nz)


def fftp2d_complex_to_real_xy(g: Grid, S: np.ndarray) -> np.ndarray:
    """Transpose (z,ky,kx)->(kx,ky,z) then per-plane 2D c2r, unnormalised
    (fftp.fpp:824-919).  Like FFTW's c2r, the imaginary parts of the kx=0 and
    kx=nx/2 columns (after the y transform) are ignored."""
    X = np.ascontiguousarray(S.transpose(2, 1, 0))           # (nz, ny, nkx)
    return sfft.irfftn(X, s=(g.ny, g.nx), axes=(1, 2), norm="forward",
                       workers=_workers())


def fc_continue_y(g: Grid, S: np.ndarray) -> None:
    """FC-Gram Dirichlet continuation along y of the physical z planes, in
    place on S (nkx, ny, nz) (fftp.fpp:565-581): the Cy slots from the first
    oy and the last oy physical y points, in the reference's order."""
    ny, C, d, p = g.ny, g.Cy, g.oy, g.np_
    D = g.diry
    top = ny - C - d
    for ii in range(C):
        acc = D[ii, 0] * S[:, top, :p] + D[C - 1 - ii, 0] * S[:, d - 1, :p]
        for jj in range(1, d):
            acc = acc + D[ii, jj] * S[:, top + jj, :p]
            acc = acc + D[C - 1 - ii, jj] * S[:, d - 1 - jj, :p]
        S[:, ny - C + ii, :p] = acc


def fftp1d_real_to_complex_x(g: Grid, R: np.ndarray) -> np.ndarray:
    """x r2c of every row, transposed to (kx, y, z) (fftp.fpp:620-716)."""
    X = sfft.rfft(R, axis=2, workers=_workers())             # (nz, ny, nkx)
    return np.ascontiguousarray(X.transpose(2, 1, 0))        # (nkx, ny, nz)


def fftp2d_real_to_complex_yz(g: Grid, S: np.ndarray) -> None:
    """y continuation (physical z planes), z continuation (every y), then the
    2D (y, z) FFT, in place (fftp.fpp:527-616)."""
    fc_continue_y(g, S)
    fc_continue(g, S)
    S[...] = sfft.fftn(S, axes=(1, 2), workers=_workers())


def fftp2d_complex_to_real_yz(g: Grid, S: np.ndarray) -> None:
    """Unnormalised 2D (y, z) backward DFT in place (fftp.fpp:922-956)."""
    S[...] = sfft.ifftn(S, axes=(1, 2), norm="forward", workers=_workers())


def fftp1d_complex_to_real_x(g: Grid, S: np.ndarray) -> np.ndarray:
    """Transpose to (z, y, kx) and unnormalised x c2r of every row; like
    FFTW's c2r the imaginary parts of kx = 0 and nx/2 are ignored
    (fftp.fpp:961-1056)."""
    X = np.ascontiguousarray(S.transpose(2, 1, 0))
    return sfft.irfft(X, n=g.nx, axis=2, norm="forward", workers=_workers())


def fftp3d_real_to_complex(g: Grid, R: np.ndarray) -> np.ndarray:
    """fftp.fpp:388-425: Cy = 0 (:419-421): 2D xy r2c, continuation, z FFT;
    Cy > 0 (:416-418): x r2c, y and z continuation, 2D yz FFT.
    Continuation planes and rows of the real input are ignored."""
    if g.Cy > 0:
        S = fftp1d_real_to_complex_x(g, R)
        fftp2d_real_to_complex_yz(g, S)
        return S
    S = fftp2d_real_to_complex_xy(g, R)
    fftp1d_real_to_complex_z(g, S)
    return S


def fftp3d_complex_to_real(g: Grid, S: np.ndarray) -> np.ndarray:
    """fftp.fpp:789-821: z IFFT then 2D xy c2r for every Cy (:817-818); in
    place on the input, destroying it."""
    fftp1d_complex_to_real_z(g, S)
    return fftp2d_complex_to_real_xy(g, S)


# ---------------------------------------------------------------------------
# Pointwise spectral operators  (pseudospec_hd.f90)
# ---------------------------------------------------------------------------
def _ik(kvec: np.ndarray, a: np.ndarray) -> np.ndarray:
    """(0,1)*k*a evaluated as the reference's complex product: (-k*Im a, k*Re a)."""
    out = np.empty_like(a)
    out.real = -(kvec * a.imag)
    out.imag = kvec * a.real
    return out


def derivk(g: Grid, a: np.ndarray, dirn: int) -> np.ndarray:
    """b = i*k_dir*a  (pseudospec_hd.f90:28-94)."""
    if dirn == 1:
        k = g.kx[:, None, None]
    elif dirn == 2:
        k = g.ky[None, :, None]
    else:
        k = g.kz[None, None, :]
    return _ik(k, a)


def laplak(g: Grid, a: np.ndarray) -> np.ndarray:
    """b = -kk2*a  (pseudospec_hd.f90:97-127)."""
    return -g.kk2 * a


def curlk(g: Grid, a, b, dirn):
    """curl component from two derivk calls (pseudospec_hd.f90:130-206)."""
    if dirn == 1:
        return derivk(g, b, 2) - derivk(g, a, 3)
    if dirn == 2:
        return derivk(g, a, 3) - derivk(g, b, 1)
    return derivk(g, b, 1) - derivk(g, a, 2)


FILTER_ALPHA = 16 * np.log(10.0)   # pseudospec_hd.f90:1099
FILTER_P = 50.0


def filter_factors(g: Grid):
    """Per-axis exponential filter factors of fc_filter (pseudospec_hd.f90:1103-1105):
    exp(-alpha*(2*k_i/n_i/Dk_i)**(2p))."""
    f64 = lambda k: np.asarray(k, np.float64)
    fx = np.exp(-FILTER_ALPHA * (2 * f64(g.kx) / g.nx / g.Dkx) ** (2 * FILTER_P))
    fy = np.exp(-FILTER_ALPHA * (2 * f64(g.ky) / g.ny / g.Dky) ** (2 * FILTER_P))
    fz = np.exp(-FILTER_ALPHA * (2 * f64(g.kz) / g.nz / g.Dkz) ** (2 * FILTER_P))
    return fx.astype(g.rdtype), fy.astype(g.rdtype), fz.astype(g.rdtype)


def fc_filter(g: Grid, a: np.ndarray) -> None:
    """a <- ((a*fx)*fy)*fz in place (pseudospec_hd.f90:1082-1115)."""
    fx, fy, fz = filter_factors(g)
    a *= fx[:, None, None]
    a *= fy[None, :, None]
    a *= fz[None, None, :]


def _phys(g: Grid, R: np.ndarray) -> np.ndarray:
    return R[: g.np_]


def gradre(g: Grid, a, b, c):
    """(A.grad)A in advective form (pseudospec_hd.f90:209-319): three stages of
    4 c2r each, products on physical planes only, scale 1/(nx*ny*nz)^2, 3 r2c."""
    npz = g.np_
    tmp = 1.0 / (float(g.nx) * float(g.ny) * float(g.nz)) ** 2
    rx = np.zeros(g.real_shape, g.rdtype)
    ry = np.zeros(g.real_shape, g.rdtype)
    rz = np.zeros(g.real_shape, g.rdtype)
    for stage, vcomp in enumerate((a, b, c), start=1):
        r1 = fftp3d_complex_to_real(g, vcomp.copy())
        r2 = fftp3d_complex_to_real(g, derivk(g, a, stage))
        r3 = fftp3d_complex_to_real(g, derivk(g, b, stage))
        r4 = fftp3d_complex_to_real(g, derivk(g, c, stage))
        p1, p2, p3, p4 = r1[:npz], r2[:npz], r3[:npz], r4[:npz]
        if stage == 1:
            rx[:npz] = p1 * p2
            ry[:npz] = p1 * p3
            rz[:npz] = p1 * p4
        elif stage == 2:
            rx[:npz] = rx[:npz] + p1 * p2
            ry[:npz] = ry[:npz] + p1 * p3
            rz[:npz] = rz[:npz] + p1 * p4
        else:
            rx[:npz] = (rx[:npz] + p1 * p2) * tmp
            ry[:npz] = (ry[:npz] + p1 * p3) * tmp
            rz[:npz] = (rz[:npz] + p1 * p4) * tmp
    return (fftp3d_real_to_complex(g, rx), fftp3d_real_to_complex(g, ry),
            fftp3d_real_to_complex(g, rz))


def advect(g: Grid, a, b, c, d):
    """(A.grad)B for a scalar B (pseudospec_phd.f90:23-113)."""
    npz = g.np_
    tmp = 1.0 / (float(g.nx) * float(g.ny) * float(g.nz)) ** 2
    r3 = np.zeros(g.real_shape, g.rdtype)
    for stage, vcomp in enumerate((a, b, c), start=1):
        r1 = fftp3d_complex_to_real(g, vcomp.copy())
        r2 = fftp3d_complex_to_real(g, derivk(g, d, stage))
        if stage == 1:
            r3[:npz] = r1[:npz] * r2[:npz]
        elif stage == 2:
            r3[:npz] = r3[:npz] + r1[:npz] * r2[:npz]
        else:
            r3[:npz] = (r3[:npz] + r1[:npz] * r2[:npz]) * tmp
    return fftp3d_real_to_complex(g, r3)


# ---------------------------------------------------------------------------
# Boundary conditions and projection  (boundary_mod.fpp, vboundary.f90, sboundary.f90)
# ---------------------------------------------------------------------------
def goto_domain_w_boundaries(g: Grid, *arrs) -> None:
    """z IFFT then x 1/nz on physical points (boundary_mod.fpp:72-150)."""
    tmp = 1.0 / float(g.nz)
    for a in arrs:
        fftp1d_complex_to_real_z(g, a)
        a[:, :, : g.np_] *= tmp


def goto_3d_fourier(g: Grid, *arrs) -> None:
    """Continuation + z FFT (boundary_mod.fpp:153-194)."""
    for a in arrs:
        fftp1d_real_to_complex_z(g, a)


def noslip_z(g: Grid, o: int, ordr: int, vx, vy, pr, vbound, pos: int) -> None:
    """Wall values vx = i*kx*pr*tau, vy = i*ky*pr*tau, tau = 1/o or (o+1)/o
    (vboundary.f90:154-211); mode (0,0) <- nx*ny*vbound."""
    ind = 0 if pos == 0 else g.np_ - 1
    tmp = 1.0 / float(o)
    if o != ordr:
        tmp = float(o + 1) * tmp
    p = pr[:, :, ind]
    vx[:, :, ind] = _ik(g.kx[:, None], p) * tmp
    vy[:, :, ind] = _ik(g.ky[None, :], p) * tmp
    vx[0, 0, ind] = float(g.nx * g.ny) * vbound[0]
    vy[0, 0, ind] = float(g.nx * g.ny) * vbound[1]


def poisson_inhomogeneous(g: Grid, a, b, c) -> np.ndarray:
    """d = -i*(kx a + ky b + kz c)/kk2, d(0,0,0)=0 (boundary_mod.fpp:405-448)."""
    s = g.kx[:, None, None] * a + g.ky[None, :, None] * b
    s = s + g.kz[None, None, :] * c
    d = np.empty_like(s)
    with np.errstate(divide="ignore", invalid="ignore"):
        d.real = s.imag / g.kk2
        d.imag = -s.real / g.kk2
    d[0, 0, 0] = 0.0
    return d


def laplace_z(g: Grid, bc: np.ndarray, bczsta: int, bczend: int):
    """Harmonic solution a (and its z derivative b) in the mixed domain
    (boundary_mod.fpp:451-678).  bc has shape (nkx, ny, 2); kinds 0
    Dirichlet, 1 Neumann, 2 Robin with constant khom; the combinations (0,0)
    (:497-526), (1,1) (:528-559), (2,2) (:561-590) and (0,2) (:592-621)."""
    khom, Lz = g.khom, g.Lz
    coef1 = np.empty(khom.shape, g.cdtype)
    coef2 = np.empty(khom.shape, g.cdtype)
    b1, b2 = bc[:, :, 0], bc[:, :, 1]
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        e1 = np.exp(-khom * Lz)
        if bczsta == 1 and bczend == 1:
            tmp = 1.0 / (khom * (1 - np.exp(-2 * khom * Lz)))
            coef1[...] = (b2 - b1 * e1) * tmp
            coef2[...] = (-b1 + b2 * e1) * tmp
            coef1[0, 0] = b1[0, 0]
            coef2[0, 0] = 0.0
        elif bczsta == 0 and bczend == 0:
            tmp = 1.0 / (1 - np.exp(-2 * khom * Lz))
            coef1[...] = (b2 - b1 * e1) * tmp
            coef2[...] = (b1 - b2 * e1) * tmp
            coef1[0, 0] = (b2[0, 0] - b1[0, 0]) / Lz
            coef2[0, 0] = b1[0, 0]
        elif bczsta == 2 and bczend == 2:
            tmp = 1.0 / (2 * khom)
            coef1[...] = b2 * tmp
            coef2[...] = b1 * tmp
            coef1[0, 0] = b1[0, 0]
            coef2[0, 0] = 0.0
        elif bczsta == 0 and bczend == 2:
            tmp = 1.0 / (2 * khom)
            coef1[...] = b2 * tmp
            coef2[...] = (b1 * 2 * khom - b2 * e1) * tmp
            coef1[0, 0] = b2[0, 0]
            coef2[0, 0] = b1[0, 0]
        else:
            raise ValueError("unsupported BC combination in laplace_z")
    z = g.z[None, None, :]
    kh = khom[:, :, None]
    with np.errstate(over="ignore", invalid="ignore"):
        ep = np.exp(kh * (z - Lz))
        em = np.exp(-kh * z)
    c1 = coef1[:, :, None]
    c2 = coef2[:, :, None]
    a = c1 * ep + c2 * em
    b = kh * (c1 * ep - c2 * em)
    a[0, 0, :] = coef1[0, 0].real * g.z + coef2[0, 0].real
    b[0, 0, :] = coef1[0, 0].real
    return a, b


def sol_project(g: Grid, a, b, c, d_out, bctarget: int, bczsta: int = 0, bczend: int = 0):
    """Solenoidal projection with a harmonic correction
    (boundary_mod.fpp:197-402).  a,b,c updated in place, d_out (pressure
    potential, mixed domain) overwritten."""
    d = poisson_inhomogeneous(g, a, b, c)
    a -= _ik(g.kx[:, None, None], d)
    b -= _ik(g.ky[None, :, None], d)
    c -= _ik(g.kz[None, None, :], d)
    tmp = 1.0 / g.nz
    if bctarget == 0:
        C1 = d * tmp
    else:
        C1 = c * tmp
    robin = bczsta == 2 or bczend == 2
    if robin:
        if bctarget != 0:
            raise ValueError("Robin data only for the potential (bctarget 0)")
        C2r = derivk(g, C1, 3)                       # (:275-278)
        fftp1d_complex_to_real_z(g, C2r)
    fftp1d_complex_to_real_z(g, C1)
    bc = np.empty((g.nkx, g.ny, 2), g.cdtype)
    top = g.np_ - 1
    # (:280-336): Dirichlet data +-C1 at the wall, Robin data from C1 and C2
    if bczsta == 0:
        bc[:, :, 0] = C1[:, :, 0] if bctarget == 1 else -C1[:, :, 0]
    elif bczsta == 2:
        bc[:, :, 0] = C2r[:, :, 0] - g.khom * C1[:, :, 0]
    else:
        raise ValueError("unsupported BC kind in sol_project")
    if bczend == 0:
        bc[:, :, 1] = C1[:, :, top] if bctarget == 1 else -C1[:, :, top]
    elif bczend == 2:
        bc[:, :, 1] = -(C2r[:, :, top] + g.khom * C1[:, :, top])
    else:
        raise ValueError("unsupported BC kind in sol_project")
    C2, C3 = laplace_z(g, bc, bczsta + bctarget, bczend + bctarget)
    if bctarget == 0:
        d_out[...] = C1 + C2
    else:
        fftp1d_complex_to_real_z(g, d)
        d_out[...] = d * tmp + C2
    fftp1d_real_to_complex_z(g, C2)
    fftp1d_real_to_complex_z(g, C3)
    a -= _ik(g.kx[:, None, None], C2)
    b -= _ik(g.ky[None, :, None], C2)
    c -= C3


def v_imposebc_and_project(g: Grid, vx, vy, vz, pr, rki: int, ordr: int,
                           v_zsta=(0.0, 0.0), v_zend=(0.0, 0.0)) -> None:
    """vboundary.f90:67-151 for no-slip walls at z=0 and z=Lz."""
    goto_domain_w_boundaries(g, vx, vy)
    noslip_z(g, rki, ordr, vx, vy, pr, v_zsta, 0)
    noslip_z(g, rki, ordr, vx, vy, pr, v_zend, 1)
    goto_3d_fourier(g, vx, vy)
    sol_project(g, vx, vy, vz, pr, 1, 0, 0)


def s_imposebc(g: Grid, th) -> None:
    """Constant (zero) scalar at both walls (sboundary.f90:67-165)."""
    goto_domain_w_boundaries(g, th)
    th[:, :, 0] = 0.0
    th[:, :, g.np_ - 1] = 0.0
    goto_3d_fourier(g, th)


# ---------------------------------------------------------------------------
# Neumann / Robin wall reconstruction and the magnetic boundary conditions
# (fcgram_mod.f90:260-644, boundary/bboundary.f90:100-344)
# ---------------------------------------------------------------------------
def load_neumann_table(tdir: str, d: int, ordn: int):
    """Q{d}.dat and Q{ordn}n{d}.dat: a float64 dxp, then Qn (d x d, Fortran
    order) as one raw stream (fcgram_mod.f90:319-345)."""
    Q = np.fromfile(os.path.join(tdir, f"Q{d}.dat"), dtype="<f8")
    raw = np.fromfile(os.path.join(tdir, f"Q{ordn}n{d}.dat"), dtype="<f8")
    if Q.size != d * d or raw.size != 1 + d * d:
        raise ValueError("Neumann table size mismatch")
    return Q.reshape(d, d).T.copy(), raw[1:].reshape(d, d).T.copy(), float(raw[0])


def neumann_vector(g: Grid, tdir: str, ordn: int) -> np.ndarray:
    """neu (ordn 1) / neu2 (ordn 2) of the z coordinate: Q(d,:) Qn^T, the last
    entry scaled by (dz/dxp)^ordn for the grid's spacing
    (fcgram_mod.f90:347-355), float64."""
    d = g.d
    Q, Qn, dxp = load_neumann_table(tdir, d, ordn)
    neu = np.zeros(d)
    for j in range(d):
        acc = 0.0
        for k in range(d):
            acc = acc + Q[d - 1, k] * Qn[j, k]
        neu[j] = acc
    if ordn == 1:
        neu[d - 1] = neu[d - 1] * (float(g.dz) / dxp)
    else:
        neu[d - 1] = neu[d - 1] * (float(g.dz) / dxp) ** 2
    return neu


def _wall_stencil(g: Grid, f, pos: int, vec):
    """vec(d)*f(wall) + sum_{k<d} vec(k)*f(interior), in the reference's
    order (fcgram_mod.f90:437-468)."""
    d = g.d
    if pos == 0:
        acc = vec[d - 1] * f[:, :, 0]
        for k in range(d - 1):
            acc = acc + vec[k] * f[:, :, d - 1 - k]
    else:
        t = g.np_ - 1
        acc = vec[d - 1] * f[:, :, t]
        for k in range(d - 1):
            acc = acc + vec[k] * f[:, :, t - d + 1 + k]
    return acc


def neumann_reconstruct(g: Grid, f, pos: int, vec) -> None:
    """Wall value from its normal derivative held at the wall point, z walls
    (boun 5 + pos; fcgram_mod.f90:437-468), in place on the mixed domain."""
    f[:, :, 0 if pos == 0 else g.np_ - 1] = _wall_stencil(g, f, pos, vec)


def robin_reconstruct(g: Grid, f, pos: int, neu) -> None:
    """Wall value with d_n f + khom f = 0 (fcgram_mod.f90:550-568)."""
    acc = _wall_stencil(g, f, pos, neu)
    f[:, :, 0 if pos == 0 else g.np_ - 1] = acc / (g.khom * neu[g.d - 1] + 1)


def int_conducting_z(g: Grid, ax, ay, pos: int) -> None:
    """bboundary.f90:192-236."""
    ind = 0 if pos == 0 else g.np_ - 1
    ax[:, :, ind] = 0.0
    ay[:, :, ind] = 0.0


def conducting_z(g: Grid, neu, neu2, ax, ay, az, pos: int) -> None:
    """bboundary.f90:239-290: zero wall values, then d2/dz2 ax = d2/dz2 ay = 0
    and d/dz az = 0 at the wall."""
    ind = 0 if pos == 0 else g.np_ - 1
    for f in (ax, ay, az):
        f[:, :, ind] = 0.0
    neumann_reconstruct(g, ax, pos, neu2)
    neumann_reconstruct(g, ay, pos, neu2)
    neumann_reconstruct(g, az, pos, neu)


def insulating_z(g: Grid, neu, ax, ay, az, pos: int) -> None:
    """bboundary.f90:294-344: zero wall values, then the Robin condition with
    khom for each component."""
    ind = 0 if pos == 0 else g.np_ - 1
    for f in (ax, ay, az):
        f[:, :, ind] = 0.0
    for f in (ax, ay, az):
        robin_reconstruct(g, f, pos, neu)


def a_imposebc_and_project(g: Grid, neu, neu2, ax, ay, az, ph, bczsta: int,
                           bczend: int) -> None:
    """Magnetic vector potential: walls 0 conducting, 1 insulating (vacuum)
    (bboundary.f90:100-189)."""
    if bczsta not in (0, 1) or bczend not in (0, 1):
        raise ValueError("unsupported magnetic boundary kind")
    az[0, 0, 0] = 0.0
    if bczsta == 0 or bczend == 0:
        goto_domain_w_boundaries(g, ax, ay)
        if bczsta == 0:
            int_conducting_z(g, ax, ay, 0)
        if bczend == 0:
            int_conducting_z(g, ax, ay, 1)
        goto_3d_fourier(g, ax, ay)
    sol_project(g, ax, ay, az, ph, 0, 2 * bczsta, 2 * bczend)
    goto_domain_w_boundaries(g, ax, ay, az)
    for pos, kind in ((0, bczsta), (1, bczend)):
        if kind == 0:
            conducting_z(g, neu, neu2, ax, ay, az, pos)
        else:
            insulating_z(g, neu, ax, ay, az, pos)
    goto_3d_fourier(g, ax, ay, az)


def _cross_real(g: Grid, A, B):
    """A x B through real space on the physical planes, scaled by
    1/(nx ny nz)^2, in the reference's operation order
    (pseudospec_hd.f90:383-399, pseudospec_mhd.f90 vector)."""
    npz = g.np_
    tmp = 1.0 / (float(g.nx) * float(g.ny) * float(g.nz)) ** 2
    r1, r2, r3 = (fftp3d_complex_to_real(g, x.copy())[:npz] for x in A)
    r4, r5, r6 = (fftp3d_complex_to_real(g, x.copy())[:npz] for x in B)
    outs = [np.zeros(g.real_shape, g.rdtype) for _ in range(3)]
    outs[0][:npz] = (r2 * r6 - r5 * r3) * tmp
    outs[1][:npz] = (r3 * r4 - r6 * r1) * tmp
    outs[2][:npz] = (r1 * r5 - r4 * r2) * tmp
    return tuple(fftp3d_real_to_complex(g, o) for o in outs)


def prodre(g: Grid, a, b, c):
    """(curl A) x A, the rotational form of the advective term
    (pseudospec_hd.f90:322-402)."""
    w = (curlk(g, b, c, 1), curlk(g, a, c, 2), curlk(g, a, b, 3))
    return _cross_real(g, w, (a, b, c))


def vector(g: Grid, a, b, c, d, e, f):
    """(a,b,c) x (d,e,f) (pseudospec_mhd.f90: SUBROUTINE vector)."""
    return _cross_real(g, (a, b, c), (d, e, f))


# ---------------------------------------------------------------------------
# RK stage bodies  (include/hd/hd_rkstep{1,2}.f90, include/bouss/bouss_rkstep{1,2}.f90)
# ---------------------------------------------------------------------------
@dataclass
class HDState:
    vx: np.ndarray
    vy: np.ndarray
    vz: np.ndarray
    pr: np.ndarray
    fx: np.ndarray
    fy: np.ndarray
    fz: np.ndarray
    th: np.ndarray | None = None
    fs: np.ndarray | None = None

    def copy(self):
        return HDState(*(None if v is None else v.copy() for v in
                         (self.vx, self.vy, self.vz, self.pr, self.fx, self.fy,
                          self.fz, self.th, self.fs)))


def hd_substep(g: Grid, st: HDState, C1, C2, C3, o: int, ordr: int, dt: float, nu: float,
               v_zsta=(0.0, 0.0), v_zend=(0.0, 0.0)) -> None:
    """One RK stage o of the HD solver (include/hd/hd_rkstep2.f90:3-36)."""
    rmp = 1.0 / float(o)
    C4, C5, C6 = gradre(g, st.vx, st.vy, st.vz)
    fc_filter(g, C4)
    fc_filter(g, C5)
    fc_filter(g, C6)
    lx, ly, lz = laplak(g, st.vx), laplak(g, st.vy), laplak(g, st.vz)
    st.vx[...] = C1 + dt * (nu * lx - C4 + st.fx) * rmp
    st.vy[...] = C2 + dt * (nu * ly - C5 + st.fy) * rmp
    st.vz[...] = C3 + dt * (nu * lz - C6 + st.fz) * rmp
    v_imposebc_and_project(g, st.vx, st.vy, st.vz, st.pr, o, ordr, v_zsta, v_zend)


def hd_step(g: Grid, st: HDState, ordr: int, dt: float, nu: float) -> None:
    """hd_rkstep1 copy then o = ord..1 (specter.fpp:1142-1161)."""
    C1, C2, C3 = st.vx.copy(), st.vy.copy(), st.vz.copy()
    for o in range(ordr, 0, -1):
        hd_substep(g, st, C1, C2, C3, o, ordr, dt, nu)


def bouss_substep(g: Grid, st: HDState, C1, C2, C3, C7, o: int, ordr: int, dt: float,
                  nu: float, kappa: float, xmom: float, xtemp: float) -> None:
    """One RK stage of the Boussinesq solver (include/bouss/bouss_rkstep2.f90:3-59)."""
    rmp = 1.0 / float(o)
    C4, C5, C6 = gradre(g, st.vx, st.vy, st.vz)
    C8 = advect(g, st.vx, st.vy, st.vz, st.th)
    C6 = C6 - xmom * st.th
    C8 = C8 - xtemp * st.vz
    for arr in (C4, C5, C6, C8):
        fc_filter(g, arr)
    lx, ly, lz, lt = (laplak(g, st.vx), laplak(g, st.vy), laplak(g, st.vz),
                      laplak(g, st.th))
    st.vx[...] = C1 + dt * (nu * lx - C4 + st.fx) * rmp
    st.vy[...] = C2 + dt * (nu * ly - C5 + st.fy) * rmp
    st.vz[...] = C3 + dt * (nu * lz - C6 + st.fz) * rmp
    st.th[...] = C7 + dt * (kappa * lt - C8 + st.fs) * rmp
    v_imposebc_and_project(g, st.vx, st.vy, st.vz, st.pr, o, ordr)
    s_imposebc(g, st.th)
    fc_filter(g, st.th)
    R1 = fftp3d_complex_to_real(g, st.th.copy())
    R1 = R1 / g.nx / g.ny / g.nz
    st.th[...] = fftp3d_real_to_complex(g, R1)


def bouss_step(g: Grid, st: HDState, ordr, dt, nu, kappa, xmom, xtemp) -> None:
    C1, C2, C3, C7 = st.vx.copy(), st.vy.copy(), st.vz.copy(), st.th.copy()
    for o in range(ordr, 0, -1):
        bouss_substep(g, st, C1, C2, C3, C7, o, ordr, dt, nu, kappa, xmom, xtemp)


def coriolis_terms(st: HDState, omega):
    """The Coriolis contributions 2 Omega x v added to the nonlinear terms
    C4, C5, C6 (include/rotbouss/rotbouss_rkstep2.f90:13-19), in the
    reference's operation order."""
    ox, oy, oz = omega
    return (2 * (oy * st.vz - oz * st.vy), 2 * (oz * st.vx - ox * st.vz),
            2 * (ox * st.vy - oy * st.vx))


def rotbouss_substep(g: Grid, st: HDState, C1, C2, C3, C7, o: int, ordr: int, dt: float,
                     nu: float, kappa: float, xmom: float, xtemp: float, omega) -> None:
    """One RK stage of the rotating Boussinesq solver
    (include/rotbouss/rotbouss_rkstep2.f90:3-50): the Boussinesq stage with
    2 Omega x v added to the nonlinear term before the filter (:13-21), and
    without the Boussinesq stage's filter and real-space round trip of th
    (the stage ends with s_imposebc, :49-50)."""
    rmp = 1.0 / float(o)
    C4, C5, C6 = gradre(g, st.vx, st.vy, st.vz)
    C8 = advect(g, st.vx, st.vy, st.vz, st.th)
    r4, r5, r6 = coriolis_terms(st, omega)
    C4 = C4 + r4
    C5 = C5 + r5
    C6 = C6 + r6 - xmom * st.th
    C8 = C8 - xtemp * st.vz
    for arr in (C4, C5, C6, C8):
        fc_filter(g, arr)
    lx, ly, lz, lt = (laplak(g, st.vx), laplak(g, st.vy), laplak(g, st.vz),
                      laplak(g, st.th))
    st.vx[...] = C1 + dt * (nu * lx - C4 + st.fx) * rmp
    st.vy[...] = C2 + dt * (nu * ly - C5 + st.fy) * rmp
    st.vz[...] = C3 + dt * (nu * lz - C6 + st.fz) * rmp
    st.th[...] = C7 + dt * (kappa * lt - C8 + st.fs) * rmp
    v_imposebc_and_project(g, st.vx, st.vy, st.vz, st.pr, o, ordr)
    s_imposebc(g, st.th)


def rotbouss_step(g: Grid, st: HDState, ordr, dt, nu, kappa, xmom, xtemp, omega) -> None:
    """rkstep1 copies (include/rotbouss/rotbouss_rkstep1.f90) + stages ORD..1."""
    C1, C2, C3, C7 = st.vx.copy(), st.vy.copy(), st.vz.copy(), st.th.copy()
    for o in range(ordr, 0, -1):
        rotbouss_substep(g, st, C1, C2, C3, C7, o, ordr, dt, nu, kappa, xmom, xtemp, omega)


@dataclass
class MHDState:
    """Fields of the MHD solver: velocity, pressure, forcing, the magnetic
    vector potential a, its projection potential ph and the electromotive
    forcing m (specter.fpp with MAGFIELD_)."""
    vx: np.ndarray
    vy: np.ndarray
    vz: np.ndarray
    pr: np.ndarray
    fx: np.ndarray
    fy: np.ndarray
    fz: np.ndarray
    ax: np.ndarray
    ay: np.ndarray
    az: np.ndarray
    ph: np.ndarray
    mx: np.ndarray
    my: np.ndarray
    mz: np.ndarray

    def copy(self):
        return MHDState(*(v.copy() for v in (
            self.vx, self.vy, self.vz, self.pr, self.fx, self.fy, self.fz,
            self.ax, self.ay, self.az, self.ph, self.mx, self.my, self.mz)))


def mhd_substep(g: Grid, st: MHDState, C1, C2, C3, C9, C10, C11, o: int, ordr: int,
                dt: float, nu: float, mu: float, b0, neu, neu2, bczsta: int = 0,
                bczend: int = 0) -> None:
    """One RK stage o of the MHD solver (include/mhd/mhd_rkstep2.f90:3-84)."""
    rmp = 1.0 / float(o)
    # b = curl a, plus the uniform field in the (0,0,0) mode (:6-16)
    C12 = curlk(g, st.ay, st.az, 1)
    C13 = curlk(g, st.ax, st.az, 2)
    C14 = curlk(g, st.ax, st.ay, 3)
    C12[0, 0, 0] = b0[0] * float(g.nx) * float(g.ny) * float(g.nz)
    C13[0, 0, 0] = b0[1] * float(g.nx) * float(g.ny) * float(g.nz)
    C14[0, 0, 0] = b0[2] * float(g.nx) * float(g.ny) * float(g.nz)
    # j = curl b, held in a (:18-21)
    jx = curlk(g, C13, C14, 1)
    jy = curlk(g, C12, C14, 2)
    jz = curlk(g, C12, C13, 3)
    st.ax[...], st.ay[...], st.az[...] = jx, jy, jz
    # (curl v) x v - j x b, filtered (:28-43)
    C4, C5, C6 = prodre(g, st.vx, st.vy, st.vz)
    C15, C16, C17 = vector(g, st.ax, st.ay, st.az, C12, C13, C14)
    C4 = C4 - C15
    C5 = C5 - C16
    C6 = C6 - C17
    for arr in (C4, C5, C6):
        fc_filter(g, arr)
    # electromotive force v x b, filtered (:45-51)
    C15, C16, C17 = vector(g, st.vx, st.vy, st.vz, C12, C13, C14)
    for arr in (C15, C16, C17):
        fc_filter(g, arr)
    lx, ly, lz = laplak(g, st.vx), laplak(g, st.vy), laplak(g, st.vz)
    st.vx[...] = C1 + dt * (nu * lx - C4 + st.fx) * rmp
    st.vy[...] = C2 + dt * (nu * ly - C5 + st.fy) * rmp
    st.vz[...] = C3 + dt * (nu * lz - C6 + st.fz) * rmp
    st.ax[...] = C9 + dt * (-mu * st.ax + C15 + st.mx) * rmp
    st.ay[...] = C10 + dt * (-mu * st.ay + C16 + st.my) * rmp
    st.az[...] = C11 + dt * (-mu * st.az + C17 + st.mz) * rmp
    v_imposebc_and_project(g, st.vx, st.vy, st.vz, st.pr, o, ordr)
    a_imposebc_and_project(g, neu, neu2, st.ax, st.ay, st.az, st.ph, bczsta, bczend)


def mhd_step(g: Grid, st: MHDState, ordr: int, dt: float, nu: float, mu: float, b0, neu,
             neu2, bczsta: int = 0, bczend: int = 0) -> None:
    """mhd_rkstep1 copies (include/mhd/mhd_rkstep1.f90) + stages ORD..1."""
    C1, C2, C3 = st.vx.copy(), st.vy.copy(), st.vz.copy()
    C9, C10, C11 = st.ax.copy(), st.ay.copy(), st.az.copy()
    for o in range(ordr, 0, -1):
        mhd_substep(g, st, C1, C2, C3, C9, C10, C11, o, ordr, dt, nu, mu, b0, neu, neu2,
                    bczsta, bczend)


# ---------------------------------------------------------------------------
# Global diagnostics  (pseudospec_hd.f90:405-1235, boundary_mod.fpp:681-801)
# ---------------------------------------------------------------------------
def _mean_phys(g: Grid, R1: np.ndarray) -> float:
    """Weighted mean over physical z in the mixed domain: weight 1 for kx=0,
    2 for every other kx including Nyquist (pseudospec_hd.f90:602-628)."""
    tmp = 1.0 / (float(g.nx) * float(g.ny) * float(g.nz)) ** 2 / float(g.nz - g.Cz)
    part = R1[:, :, : g.np_] * tmp
    return float(part[0].sum() + 2.0 * part[1:].sum())


def _zinv(g: Grid, a):
    c = a.copy()
    fftp1d_complex_to_real_z(g, c)
    return c


def energy(g: Grid, a, b, c, kin: int) -> float:
    """<|v|^2> (kin=1), the reference's enstrophy (kin=0, which sums
    wx^2 + wy^2 + wy^2, pseudospec_hd.f90:527,540) or <|curl curl v|^2>
    (kin=2, the current of mhdcheck, :555-590) (pseudospec_hd.f90:405-635)."""
    if kin == 1:
        comps = (a, b, c)
    elif kin == 0:
        comps = (curlk(g, b, c, 1), curlk(g, a, c, 2), curlk(g, a, c, 2))
    elif kin == 2:
        w1, w2, w3 = curlk(g, b, c, 1), curlk(g, a, c, 2), curlk(g, a, b, 3)
        comps = (curlk(g, w2, w3, 1), curlk(g, w1, w3, 2), curlk(g, w1, w2, 3))
    else:
        raise ValueError("kin must be 0, 1 or 2")
    R1 = None
    for comp in comps:
        z = _zinv(g, comp)
        sq = z.real ** 2 + z.imag ** 2
        R1 = sq if R1 is None else R1 + sq
    return _mean_phys(g, R1)


def cross(g: Grid, a, b, c, d, e, f) -> float:
    """<v.f> for kin=1 (pseudospec_hd.f90:778-940)."""
    R1 = None
    for p, q in ((a, d), (b, e), (c, f)):
        zp, zq = _zinv(g, p), _zinv(g, q)
        r = (zp * np.conj(zq)).real
        R1 = r if R1 is None else R1 + r
    return _mean_phys(g, R1)


def helicity(g: Grid, a, b, c) -> float:
    """Kinetic helicity <v . curl v> over physical z (pseudospec_hd.f90:638-775):
    Re(IFFT(v_i) conj(IFFT(w_i))) with w = (curlk(b,c,1), curlk(a,c,2),
    curlk(a,b,3)), summed over the components in that order."""
    R1 = None
    for v, w in ((a, curlk(g, b, c, 1)), (b, curlk(g, a, c, 2)), (c, curlk(g, a, b, 3))):
        zv, zw = _zinv(g, v), _zinv(g, w)
        r = (zv * np.conj(zw)).real
        R1 = r if R1 is None else R1 + r
    return _mean_phys(g, R1)


def divergence(g: Grid, a, b, c) -> float:
    """<|div v|^2> over the physical domain (pseudospec_hd.f90:1118-1235)."""
    C2 = derivk(g, a, 1)
    C2 = C2 + derivk(g, b, 2)
    C2 = C2 + derivk(g, c, 3)
    z = _zinv(g, C2)
    return _mean_phys(g, z.real ** 2 + z.imag ** 2)


def bouncheck_z(g: Grid, a, b=None):
    """Mean |a|^2 (+|b|^2) on the z=0 and z=Lz planes (boundary_mod.fpp:681-801)."""
    tmp = (1.0 / (g.nx * g.ny * g.nz)) ** 2
    za = _zinv(g, a)
    R1 = za.real[:, :, 0] ** 2 + za.imag[:, :, 0] ** 2
    R2 = za.real[:, :, g.np_ - 1] ** 2 + za.imag[:, :, g.np_ - 1] ** 2
    if b is not None:
        zb = _zinv(g, b)
        R1 = R1 + zb.real[:, :, 0] ** 2 + zb.imag[:, :, 0] ** 2
        R2 = R2 + zb.real[:, :, g.np_ - 1] ** 2 + zb.imag[:, :, g.np_ - 1] ** 2
    bot = float((R1[0] * tmp).sum() + (2 * R1[1:] * tmp).sum())
    top = float((R2[0] * tmp).sum() + (2 * R2[1:] * tmp).sum())
    return bot, top


def mhdcheck(g: Grid, st: "MHDState"):
    """The values mhdcheck writes with hel = crs = 1 (pseudospec_mhd.f90
    mhdcheck): balance.txt (eng, ens, cur), energy.txt (engk, engm),
    helicity.txt (helk, helm), cross.txt (crh, asq)."""
    engk = energy(g, st.vx, st.vy, st.vz, 1)
    ens = energy(g, st.vx, st.vy, st.vz, 0)
    engm = energy(g, st.ax, st.ay, st.az, 0)
    cur = energy(g, st.ax, st.ay, st.az, 2)
    helk = helicity(g, st.vx, st.vy, st.vz)
    helm = helicity(g, st.ax, st.ay, st.az)
    asq = energy(g, st.ax, st.ay, st.az, 1)
    c1, c2, c3 = derivk(g, st.ax, 1), derivk(g, st.ay, 2), derivk(g, st.az, 3)
    crh = cross(g, st.vx, st.vy, st.vz, c1, c2, c3)
    return (engk + engm, ens, cur, engk, engm, helk, helm, crh, asq)


def hdcheck(g: Grid, st: HDState):
    """energy, enstrophy and injection rate of balance.txt (pseudospec_hd.f90:943-1005)."""
    return (energy(g, st.vx, st.vy, st.vz, 1), energy(g, st.vx, st.vy, st.vz, 0),
            cross(g, st.vx, st.vy, st.vz, st.fx, st.fy, st.fz))


def hdcheck_helicity(g: Grid, st: HDState) -> float:
    """The helicity.txt value hdcheck writes with hel=1 (pseudospec_hd.f90:980-1003;
    include/hd/hd_global.f90:2 passes hel=1)."""
    return helicity(g, st.vx, st.vy, st.vz)


def vdiagnostic(g: Grid, st: HDState):
    """<div^2>, wall <v_t^2> (bottom, top), wall <v_n^2> (bottom, top)
    (vboundary.f90:214-269)."""
    dv = divergence(g, st.vx, st.vy, st.vz)
    tb, tt = bouncheck_z(g, st.vx, st.vy)
    nb, nt = bouncheck_z(g, st.vz)
    return dv, tb, tt, nb, nt


# ---------------------------------------------------------------------------
# Scalar diagnostics  (pseudospec_phd.f90:116-321)
# ---------------------------------------------------------------------------
def variance(g: Grid, a, kin: int) -> float:
    """<|a|^2> (kin=1) or <|kk2 a|^2> (kin=0) over physical z
    (pseudospec_phd.f90:116-176)."""
    at = a if kin == 1 else g.kk2 * a
    z = _zinv(g, at)
    return _mean_phys(g, z.real ** 2 + z.imag ** 2)


def product(g: Grid, a, b) -> float:
    """<Re(a conj b)> over physical z (pseudospec_phd.f90:178-235)."""
    za, zb = _zinv(g, a), _zinv(g, b)
    return _mean_phys(g, (za * np.conj(zb)).real)


def pscheck(g: Grid, th, fs):
    """scalar.txt quantities (pseudospec_phd.f90:237-262): <th^2>,
    <|kk2 th|^2>, <th fs>."""
    return variance(g, th, 1), variance(g, th, 0), product(g, th, fs)


# ---------------------------------------------------------------------------
# Wall and solution diagnostics the reference does not compute (defined here)
# ---------------------------------------------------------------------------
def _plane_mean_dz(g: Grid, a):
    """Horizontal mean of d a/dz at z = 0 and z = Lz in physical units:
    the (kx, ky) = (0, 0) line of IFFT_z(i kz a) / (nx ny nz)."""
    line = _ik(g.kz, a[0, 0, :])
    z = np.fft.ifft(line) * g.nz          # unnormalised backward DFT
    s = 1.0 / (g.nx * g.ny * g.nz)
    return z[0].real * s, z[g.np_ - 1].real * s


def wall_shear(g: Grid, a, b, nu: float):
    """Mean wall shear stress nu <d(u_x, u_y)/dz> at z = 0 and z = Lz:
    (tau_x bottom, tau_y bottom, tau_x top, tau_y top)."""
    ab, at = _plane_mean_dz(g, a)
    bb, bt = _plane_mean_dz(g, b)
    return nu * ab, nu * bb, nu * at, nu * bt


def nusselt(g: Grid, th):
    """Nusselt number at z = 0 and z = Lz for theta_total = 1 - z/Lz + th
    (heated bottom wall): Nu = 1 - Lz <d th/dz>_wall."""
    b, t = _plane_mean_dz(g, th)
    return 1.0 - g.Lz * b, 1.0 - g.Lz * t


def divergence_max(g: Grid, a, b, c) -> float:
    """max |div v| over the physical points, physical units."""
    d = derivk(g, a, 1) + derivk(g, b, 2) + derivk(g, c, 3)
    R = fftp3d_complex_to_real(g, d.copy())[: g.np_]
    return float(np.abs(R).max() / (g.nx * g.ny * g.nz))


def energy_spectrum_h(g: Grid, a, b, c, nbins: int):
    """Horizontal shell spectrum of <|v|^2> over physical z: bin s collects
    the (kx, ky) lines with s - 1/2 <= |(kx, ky)|/Dk < s + 1/2 (Dk = 1/Lx);
    the bins sum to energy(kin=1) when nbins covers every shell."""
    tmp = 1.0 / (float(g.nx) * float(g.ny) * float(g.nz)) ** 2 / float(g.nz - g.Cz)
    R1 = None
    for comp in (a, b, c):
        z = _zinv(g, comp)
        sq = z.real ** 2 + z.imag ** 2
        R1 = sq if R1 is None else R1 + sq
    lines = R1[:, :, : g.np_].sum(axis=2) * tmp
    lines[1:] *= 2.0
    kh = np.sqrt(g.kx[:, None] ** 2 + g.ky[None, :] ** 2) / g.Dkx
    s = np.floor(kh + 0.5).astype(int)
    out = np.zeros(nbins)
    m = s < nbins
    np.add.at(out, s[m], lines[m])
    return out
