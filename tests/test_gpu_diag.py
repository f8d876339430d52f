"""GPU scalar, wall and spectrum diagnostics against the float64 oracle
(variance / product / pscheck of pseudospec_phd.f90, plus the wall shear,
Nusselt number, max |div v| and horizontal energy spectrum this library
defines)."""
import numpy as np
import pytest

from conftest import TABLES

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(16, 16, 41, 25, 5), (32, 24, 57, 25, 5),
                                   (12, 20, 91, 27, 8)],
                         ids=lambda s: "x".join(map(str, s)))
def test_scalar_wall_spectrum_diagnostics(shape):
    import torch
    from oracle import specter_ref as O
    from specter_amd import FCPlan, ic
    nx, ny, nz, C, d = shape
    g = O.make_grid(nx, ny, nz, C, d, TABLES)
    u = ic.spectrum_peaked_velocity(nx, ny, nz, C, 5)
    th = ic.spectrum_peaked_velocity(nx, ny, nz, C, 6)[0]
    fs = ic.spectrum_peaked_velocity(nx, ny, nz, C, 7)[1]
    S = [O.fftp3d_real_to_complex(g, a.copy()) for a in (*u, th, fs)]
    plan = FCPlan(nx, ny, nz, C, d)
    D = [torch.as_tensor(a).cuda() for a in S]
    vx, vy, vz, tth, tfs = D
    nu = 3e-3

    def close(a, b, tol):
        a, b = np.asarray(a, float), np.asarray(b, float)
        scale = max(np.abs(b).max(), 1e-300)
        assert np.abs(a - b).max() <= tol * scale, (a, b)

    close(plan.variance(tth, 1), O.variance(g, S[3], 1), 1e-12)
    close(plan.variance(tth, 0), O.variance(g, S[3], 0), 1e-12)
    close(plan.product(tth, tfs), O.product(g, S[3], S[4]), 1e-12)
    close(plan.wall_shear(vx, vy, nu), O.wall_shear(g, S[0], S[1], nu), 1e-10)
    close(plan.nusselt(tth), O.nusselt(g, S[3]), 1e-10)
    close(plan.divergence_max(vx, vy, vz), O.divergence_max(g, *S[:3]), 1e-10)
    nb = max(nx, ny)
    close(plan.energy_spectrum(vx, vy, vz, nb), O.energy_spectrum_h(g, *S[:3], nb), 1e-12)
