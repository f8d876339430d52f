"""Initial conditions and forcing of the benchmark workloads (host, numpy).

These are the project's own generators (the reference selects compile-time
include files instead, src/initialv.f90, src/initialfv.f90).  They produce
REAL fields on the physical grid in the reference's real layout
(numpy (nz, ny, nx), planes k >= nz-Cz zero); the solver moves them to the
spectral domain with the plan's fftp3d_real_to_complex, exactly as a
reference restart (specter.fpp:885-934) does.

Random numbers come from numpy's counter-based Philox generator, so a seed
fixes the field on every machine.
"""
from __future__ import annotations

import numpy as np

PI = 3.14159265358979323846


def physical_grid(nx, ny, nz, Cz, Lx=1.0, Ly=1.0, Lz=1.0):
    """x, y, z coordinates of the reference grid (specter.fpp:683-749)."""
    x = (Lx * 2.0 * PI / nx) * np.arange(nx)
    y = (Ly * 2.0 * PI / ny) * np.arange(ny)
    z = (Lz / (nz - Cz - 1)) * np.arange(nz)
    return x, y, z


def wall_envelope(nz, Cz):
    """sin(pi z/Lz) on the physical points, exactly 0 at both walls."""
    npz = nz - Cz
    env = np.sin(PI * np.arange(npz) / (npz - 1))
    env[0] = 0.0
    env[-1] = 0.0
    return env


def spectrum_peaked_velocity(nx, ny, nz, Cz, seed, Lx=1.0, Ly=1.0, Lz=1.0, kpeak=3.0):
    """Random-phase solenoidal field with shell spectrum E(k) ~ k^4 exp(-2 (k/kpeak)^2),
    multiplied by sin(pi z/Lz) and normalised to <|u|^2> = 1 over the physical
    points.  Returns (ux, uy, uz) as float64 (nz, ny, nx) arrays."""
    npz = nz - Cz
    dz = Lz / (npz - 1)
    rng = np.random.Generator(np.random.Philox(seed))
    kx = np.fft.fftfreq(nx, d=1.0 / nx) / Lx
    ky = np.fft.fftfreq(ny, d=1.0 / ny) / Ly
    kz = np.fft.rfftfreq(npz, d=dz) * 2.0 * PI     # periodic box of length npz*dz
    KZ, KY, KX = np.meshgrid(kz, ky, kx, indexing="ij")  # (nkz, ny, nx) -- z halved
    K2 = KX ** 2 + KY ** 2 + KZ ** 2
    K = np.sqrt(K2)
    with np.errstate(divide="ignore", invalid="ignore"):
        amp = np.where(K > 0, K * np.exp(-(K / kpeak) ** 2), 0.0)
    shape = K.shape
    comps = []
    for _ in range(3):
        phase = rng.uniform(0.0, 2.0 * PI, size=shape)
        comps.append(amp * np.exp(1j * phase))
    ux, uy, uz = comps
    with np.errstate(divide="ignore", invalid="ignore"):
        kdotu = (KX * ux + KY * uy + KZ * uz) / np.where(K2 > 0, K2, 1.0)
    ux, uy, uz = ux - KX * kdotu, uy - KY * kdotu, uz - KZ * kdotu
    out = []
    env = wall_envelope(nz, Cz)
    for c in (ux, uy, uz):
        # axes (z, y, x) with z halved: irfftn over z last needs axes reordering
        r = np.fft.irfftn(c, s=(ny, nx, npz), axes=(1, 2, 0))
        full = np.zeros((nz, ny, nx))
        full[:npz] = r * env[:, None, None]
        out.append(full)
    e = sum(float(np.mean(f[:npz] ** 2)) for f in out)
    s = 1.0 / np.sqrt(e)
    return tuple(f * s for f in out)


def large_scale_forcing(nx, ny, nz, Cz, amp=0.5, Lx=1.0, Ly=1.0, Lz=1.0):
    """f = amp * (sin 2y, sin 2x, 0) * sin(pi z/Lz) on the physical points."""
    x, y, _ = physical_grid(nx, ny, nz, Cz, Lx, Ly, Lz)
    env = wall_envelope(nz, Cz)
    npz = nz - Cz
    fx = np.zeros((nz, ny, nx))
    fy = np.zeros((nz, ny, nx))
    fz = np.zeros((nz, ny, nx))
    fx[:npz] = amp * np.sin(2 * y)[None, :, None] * env[:, None, None] * np.ones(nx)[None, None, :]
    fy[:npz] = amp * np.sin(2 * x)[None, None, :] * env[:, None, None] * np.ones(ny)[None, :, None]
    return fx, fy, fz


def thermal_perturbation(nx, ny, nz, Cz, seed, amp=1e-3):
    """theta' = amp * white noise on the interior, 0 on the walls."""
    rng = np.random.Generator(np.random.Philox(seed))
    npz = nz - Cz
    th = np.zeros((nz, ny, nx))
    th[1:npz - 1] = amp * rng.standard_normal((npz - 2, ny, nx))
    return th
