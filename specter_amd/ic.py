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


def spectrum_peaked_velocity_device(nx, ny, nz, Cz, seed, Lx=1.0, Ly=1.0, Lz=1.0,
                                    kpeak=3.0):
    """The same construction as spectrum_peaked_velocity, on the GPU with
    torch (for grids whose host generation would take minutes and tens of
    GB): phases from torch's Philox generator on the device, so a seed fixes
    the field for a given device type and torch version (not bit-identical
    to the numpy generator).  Returns three float64 (nz, ny, nx) cuda
    tensors, computed one component at a time to bound the memory."""
    import torch
    dev = torch.device("cuda")
    npz = nz - Cz
    dz = Lz / (npz - 1)
    g = torch.Generator(device=dev)
    g.manual_seed(int(seed))
    f64 = torch.float64
    kx = torch.fft.fftfreq(nx, d=1.0 / nx, dtype=f64, device=dev) / Lx
    ky = torch.fft.fftfreq(ny, d=1.0 / ny, dtype=f64, device=dev) / Ly
    kz = torch.fft.rfftfreq(npz, d=dz, dtype=f64, device=dev) * 2.0 * PI
    KX, KY, KZ = kx[None, None, :], ky[None, :, None], kz[:, None, None]
    K2 = KZ ** 2 + KY ** 2 + KX ** 2                       # (nkz, ny, nx)
    K = torch.sqrt(K2)
    amp = torch.where(K > 0, K * torch.exp(-(K / kpeak) ** 2), torch.zeros_like(K))
    del K
    comps = []
    for _ in range(3):
        phase = torch.rand(K2.shape, generator=g, dtype=f64, device=dev) * (2.0 * PI)
        comps.append(torch.polar(amp, phase))
        del phase
    del amp
    K2s = torch.where(K2 > 0, K2, torch.ones_like(K2))
    kdotu = (KX * comps[0] + KY * comps[1] + KZ * comps[2]) / K2s
    del K2s, K2
    comps = [comps[0] - KX * kdotu, comps[1] - KY * kdotu, comps[2] - KZ * kdotu]
    del kdotu
    env = torch.as_tensor(wall_envelope(nz, Cz), device=dev)
    out = []
    for i in range(3):
        c = comps[i]
        comps[i] = None
        r = torch.fft.irfftn(c, s=(npz, ny, nx), dim=(0, 1, 2))
        del c
        full = torch.zeros((nz, ny, nx), dtype=f64, device=dev)
        full[:npz] = r * env[:, None, None]
        del r
        out.append(full)
    e = sum(float(torch.mean(f[:npz] ** 2)) for f in out)
    s = 1.0 / e ** 0.5
    for f in out:
        f.mul_(s)
    return tuple(out)


def large_scale_forcing_device(nx, ny, nz, Cz, amp=0.5, Lx=1.0, Ly=1.0, Lz=1.0):
    """large_scale_forcing on the GPU (float64 cuda tensors)."""
    import torch
    dev = torch.device("cuda")
    x, y, _ = physical_grid(nx, ny, nz, Cz, Lx, Ly, Lz)
    env = torch.as_tensor(wall_envelope(nz, Cz), device=dev)
    npz = nz - Cz
    xs = torch.as_tensor(np.sin(2 * x), device=dev)
    ys = torch.as_tensor(np.sin(2 * y), device=dev)
    fx = torch.zeros((nz, ny, nx), dtype=torch.float64, device=dev)
    fy = torch.zeros_like(fx)
    fz = torch.zeros_like(fx)
    fx[:npz] = amp * ys[None, :, None] * env[:, None, None] * torch.ones_like(xs)[None, None, :]
    fy[:npz] = amp * xs[None, None, :] * env[:, None, None] * torch.ones_like(ys)[None, :, None]
    return fx, fy, fz
