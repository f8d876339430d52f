"""Ensemble server: continuous batching of independent runs.  Every result
must be the same as the run computed alone (bit for bit: a member of a
batched plan runs the single run's per-line / per-row code), whatever it was
batched with and whenever it was admitted."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _alone(req):
    from specter_amd import FCPlan, Solver, SolverParams, ic
    nx, ny, nz, C, d = req.grid
    plan = FCPlan(nx, ny, nz, C, d, precision=req.precision)
    sol = Solver(plan, SolverParams(solver="HD", ord=req.ord, dt=req.dt, nu=req.nu))
    # the server generates its initial fields on the GPU
    u = ic.spectrum_peaked_velocity_device(nx, ny, nz, C, req.seed)
    for name, arr in zip(("vx", "vy", "vz"), u):
        sol.set_real_field(name, arr)
    sol.step(req.nsteps)
    return {n: sol.spectral(n) for n in ("vx", "vy", "vz")}


@pytest.mark.parametrize("precision", [1, 0], ids=["mixed", "fp64"])
def test_results_equal_alone_and_in_the_mix(precision):
    from specter_amd.ensemble import EnsembleServer, RunRequest, latency_stats
    reqs = []
    spec = [(16, 5, 30), (16, 5, 12), (32, 8, 20), (16, 5, 45), (32, 8, 9), (16, 8, 17),
            (16, 5, 25), (32, 8, 31)]
    for i, (n, d, steps) in enumerate(spec):
        reqs.append(RunRequest(rid=i, n=n, d=d, nu=1e-3 * (1 + 0.37 * i), seed=100 + i,
                               nsteps=steps, arrival=0.02 * i, precision=precision))
    srv = EnsembleServer(max_slots=3, chunk=7, keep_fields=True)
    results = srv.run(reqs, realtime=True)
    assert sorted(r.rid for r in results) == list(range(len(reqs)))
    assert srv.peak_classes >= 2
    for res in results:
        ref = _alone(reqs[res.rid])
        for name in ("vx", "vy", "vz"):
            assert np.array_equal(res.fields[name], ref[name]), (res.rid, name)
        assert res.finished >= res.admitted >= 0.0
    st = latency_stats(results)
    print(st)
    assert 0 < st["p50_s"] <= st["p99_s"]
