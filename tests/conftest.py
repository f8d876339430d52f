import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

TABLES = os.path.join(ROOT, "specter_amd", "tables")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and libspx.so")


@pytest.fixture(scope="session")
def tables_dir():
    return TABLES


def rel_err(a, b):
    """max|a-b| / max|b| (0 if both are zero)."""
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.max(np.abs(b))
    num = np.max(np.abs(a - b))
    if den == 0:
        return float(num)
    return float(num / den)


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den else 1.0))


# HD 64x64x(64+25), d=5, C=25, RK2, dt=1e-3, nu=2e-3, unforced, 100 steps
CFG1 = dict(shape=(64, 64, 89, 25, 5), prm=dict(solver="HD", ord=2, dt=1e-3, nu=2e-3),
            nsteps=100, diag_every=10, seed=1000)


def oracle_trajectory(shape, prm, nsteps, seed=1000, diag_every=0):
    """Float64 oracle run from the seeded spectrum-peaked field: final state
    and the (energy, enstrophy, injection, <div^2>, wall values) series."""
    from oracle import specter_ref as O
    from specter_amd import ic
    nx, ny, nz, C, d = shape
    g = O.make_grid(nx, ny, nz, C, d, TABLES)
    u = ic.spectrum_peaked_velocity(nx, ny, nz, C, seed)
    zero = np.zeros(g.spec_shape, complex)
    st = O.HDState(*(O.fftp3d_real_to_complex(g, a.copy()) for a in u), zero.copy(),
                   zero.copy(), zero.copy(), zero.copy())
    series = []
    for t in range(nsteps):
        if diag_every and t % diag_every == 0:
            series.append(O.hdcheck(g, st) + O.vdiagnostic(g, st))
        O.hd_step(g, st, prm["ord"], prm["dt"], prm["nu"])
    return g, st, np.array(series)


def gpu_trajectory(shape, prm, nsteps, seed=1000, diag_every=0, precision=0):
    """The same run on the GPU through libspx (precision: SPX_FP64 = 0,
    SPX_FP32_MIXED = 1)."""
    from specter_amd import FCPlan, Solver, SolverParams, ic
    nx, ny, nz, C, d = shape
    plan = FCPlan(nx, ny, nz, C, d, precision=precision)
    sol = Solver(plan, SolverParams(**prm))
    for name, arr in zip(("vx", "vy", "vz"), ic.spectrum_peaked_velocity(nx, ny, nz, C, seed)):
        sol.set_real_field(name, arr)
    series = []
    for t in range(nsteps):
        if diag_every and t % diag_every == 0:
            series.append(sol.hdcheck() + sol.vdiagnostic())
        sol.step(1)
    return sol, np.array(series)


GOLDEN = os.path.join(ROOT, "tests", "golden")


def cfg1_longdouble(st):
    """The longdouble trajectory of cfg1 after 100 steps, rebuilt from the
    float64 oracle's final state st and the committed difference
    (tests/golden/make_cfg1_longdouble.py)."""
    z = np.load(os.path.join(GOLDEN, "cfg1_longdouble_delta.npz"))
    out = {}
    for n in ("vx", "vy", "vz"):
        s = float(z[f"{n}_scale"])
        dl = (z[f"{n}_re"].astype(np.float64) + 1j * z[f"{n}_im"].astype(np.float64)) * s
        out[n] = getattr(st, n).astype(np.clongdouble) + dl.astype(np.clongdouble)
        # the fixture matches this float64 trajectory
        rel = float(np.linalg.norm(dl.ravel()) / np.linalg.norm(out[n].ravel()))
        assert abs(rel / float(z[f"{n}_rel"]) - 1) < 1e-2, (n, rel, float(z[f"{n}_rel"]))
    return out


@pytest.fixture(scope="session")
def cfg1_oracle():
    return oracle_trajectory(CFG1["shape"], CFG1["prm"], CFG1["nsteps"], CFG1["seed"],
                             CFG1["diag_every"])
