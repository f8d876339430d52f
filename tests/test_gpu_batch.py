"""Batched plans: several independent runs of one grid stepped together.

Every member of a batched plan is computed by the same per-line / per-row
kernel code as a single run (the members are extra (kx, ky) lines for the z
passes and extra planes for the y and x passes), so a member's fields and
diagnostics must equal the single run's bit for bit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPE = (32, 32, 57, 25, 5)
RUNS = [dict(seed=11, nu=2e-3), dict(seed=12, nu=5e-3), dict(seed=13, nu=8e-4)]


def _ic(seed):
    from specter_amd import ic
    nx, ny, nz, C, _ = SHAPE
    return ic.spectrum_peaked_velocity(nx, ny, nz, C, seed)


def _single(run, precision, nsteps):
    from specter_amd import FCPlan, Solver, SolverParams
    plan = FCPlan(*SHAPE, precision=precision)
    sol = Solver(plan, SolverParams(solver="HD", ord=2, dt=1e-3, nu=run["nu"]))
    for name, arr in zip(("vx", "vy", "vz"), _ic(run["seed"])):
        sol.set_real_field(name, arr)
    sol.step(nsteps)
    out = {n: sol.spectral(n) for n in ("vx", "vy", "vz", "pr")}
    out["diag"] = sol.hdcheck() + sol.vdiagnostic()
    return out


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_batched_members_equal_single_runs(precision):
    from specter_amd import FCPlan, Solver, SolverParams
    nsteps = 6
    plan = FCPlan(*SHAPE, precision=precision, members=len(RUNS))
    sol = Solver(plan, SolverParams(solver="HD", ord=2, dt=1e-3, nu=1e-3))
    for b, run in enumerate(RUNS):
        sol.set_member(b, run["nu"])
        for name, arr in zip(("vx", "vy", "vz"), _ic(run["seed"])):
            sol.set_real_field(name, arr, member=b)
    sol.step(nsteps)
    diag = np.array(sol.hdcheck() + sol.vdiagnostic())      # (quantities, members)
    for b, run in enumerate(RUNS):
        ref = _single(run, precision, nsteps)
        for name in ("vx", "vy", "vz", "pr"):
            assert np.array_equal(sol.spectral(name, member=b), ref[name]), (b, name)
        assert np.array_equal(diag[:, b], np.array(ref["diag"])), b


def test_compaction_and_active_count():
    """A finished member is replaced by the last active one and the batch
    continues with fewer members; every remaining run still equals its
    single-run trajectory."""
    from specter_amd import FCPlan, Solver, SolverParams
    plan = FCPlan(*SHAPE, members=3)
    sol = Solver(plan, SolverParams(solver="HD", ord=2, dt=1e-3, nu=1e-3))
    for b, run in enumerate(RUNS):
        sol.set_member(b, run["nu"])
        for name, arr in zip(("vx", "vy", "vz"), _ic(run["seed"])):
            sol.set_real_field(name, arr, member=b)
    sol.step(3)
    # run 0 leaves: run 2 moves into slot 0, two members stay active
    sol.copy_member(0, 2)
    plan.set_active(2)
    sol.step(4)
    for slot, run in ((0, RUNS[2]), (1, RUNS[1])):
        ref = _single(run, 0, 7)
        for name in ("vx", "vy", "vz", "pr"):
            assert np.array_equal(sol.spectral(name, member=slot), ref[name]), (slot, name)
