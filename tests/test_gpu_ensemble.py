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


def test_failed_run_is_reported_and_leaves_its_batch_alone():
    """A run whose fields go non-finite (here nu = NaN) is returned with
    status "failed" after its first chunk; the runs batched with it finish
    with the same bits as alone."""
    from specter_amd.ensemble import EnsembleServer, RunRequest
    reqs = [RunRequest(rid=0, n=16, d=5, nu=2e-3, seed=11, nsteps=20, precision=1),
            RunRequest(rid=1, n=16, d=5, nu=float("nan"), seed=12, nsteps=20, precision=1),
            RunRequest(rid=2, n=16, d=5, nu=3e-3, seed=13, nsteps=20, precision=1)]
    srv = EnsembleServer(max_slots=3, chunk=5, keep_fields=True)
    results = {r.rid: r for r in srv.run(reqs, realtime=False)}
    assert results[1].status == "failed" and "non-finite" in results[1].error
    assert results[1].steps_done < 20
    for rid in (0, 2):
        assert results[rid].status == "ok"
        ref = _alone(reqs[rid])
        for name in ("vx", "vy", "vz"):
            assert np.array_equal(results[rid].fields[name], ref[name]), (rid, name)


def test_admission_within_the_hbm_budget():
    """Requests wait for memory: with room for one class at a time two grids
    are served one after the other; a budget no class fits is an error."""
    from specter_amd.ensemble import EnsembleServer, RunRequest, _Class
    reqs = [RunRequest(rid=0, n=32, d=5, nu=2e-3, seed=1, nsteps=6, precision=0),
            RunRequest(rid=1, n=16, d=5, nu=2e-3, seed=2, nsteps=6, precision=0)]
    need = []
    for r in reqs:
        c = _Class(r, 1)
        need.append(c.bytes)
        c.close()
    one = max(need) + min(need) // 2      # room for either class, not both
    srv = EnsembleServer(hbm_budget=one, max_slots=1, chunk=3)
    res = srv.run(reqs, realtime=False)
    assert sorted(r.rid for r in res) == [0, 1] and all(r.status == "ok" for r in res)
    assert srv.peak_classes == 1 and srv.peak_used <= one
    with pytest.raises(RuntimeError, match="HBM budget"):
        EnsembleServer(hbm_budget=1 << 20, max_slots=1).run(
            [RunRequest(rid=0, n=32, d=5, nu=2e-3, seed=1, nsteps=2, precision=0)],
            realtime=False)
