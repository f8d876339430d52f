"""z lines longer than 256 threads x 8 points: nz = 1049 (the 1024^3
workload's prime length; 1048 = 8*131 is not smooth, so Bluestein with an
in-LDS FFT of length 2100 on one 512-thread workgroup per line)."""
import numpy as np
import pytest

from conftest import TABLES, rel_err, rel_l2

pytestmark = pytest.mark.gpu

SHAPE = (8, 8, 1049, 25, 5)


def test_long_z_transforms_and_step():
    import torch
    from oracle import specter_ref as O
    from specter_amd import FCPlan, Solver, SolverParams, ic
    nx, ny, nz, C, d = SHAPE
    g = O.make_grid(nx, ny, nz, C, d, TABLES)
    p = FCPlan(nx, ny, nz, C, d)
    assert p.z_fft_len == 2100 and p.z_algorithm == 1
    rng = np.random.default_rng(11)
    R = rng.standard_normal(g.real_shape)
    ref = O.fftp3d_real_to_complex(g, R.copy())
    out = p.empty_spectral()
    p.fftp3d_real_to_complex(torch.as_tensor(R).cuda(), out)
    assert rel_err(out.cpu().numpy(), ref) < 1e-12
    S = rng.standard_normal(g.spec_shape) + 1j * rng.standard_normal(g.spec_shape)
    ref = O.fftp3d_complex_to_real(g, S.copy())
    outr = p.empty_real()
    p.fftp3d_complex_to_real(torch.as_tensor(S).cuda(), outr)
    assert rel_err(outr.cpu().numpy(), ref) < 1e-12
    # one HD RK2 step (walls, projection, gradre) from a smooth field, both precisions
    u = ic.spectrum_peaked_velocity(nx, ny, nz, C, 3)
    zero = np.zeros(g.spec_shape, complex)
    st = O.HDState(*(O.fftp3d_real_to_complex(g, a.copy()) for a in u), zero.copy(),
                   zero.copy(), zero.copy(), zero.copy())
    O.hd_step(g, st, 2, 1e-3, 2e-3)
    for prec, tol in ((0, 1e-10), (1, 1e-5)):
        sol = Solver(FCPlan(nx, ny, nz, C, d, precision=prec),
                     SolverParams(solver="HD", ord=2, dt=1e-3, nu=2e-3))
        for name, arr in zip(("vx", "vy", "vz"), u):
            sol.set_real_field(name, arr)
        sol.step(1)
        for name in ("vx", "vy", "vz"):
            err = rel_l2(sol.spectral(name), getattr(st, name))
            print(f"nz=1049 precision {prec} {name}: {err:.2e}")
            assert err < tol, (prec, name)
