"""Parity of the z-line kernel paths the full-size workloads select, on small
(x, y) grids with the workloads' own z lengths, against the float64 oracle:

* nz = 281 (HD 256^3, d=5): Rader 280 = 8*5*7, convolution tables in global
  memory in float64 plans;
* nz = 539 (Boussinesq 512^3, d=8): direct 7*7*11, 128 threads per line,
  float64 and float32-mixed Boussinesq;
* nz = 537 (HD 512^3, d=5): 3*179, Bluestein 1080;
* nz = 153, 155, 217, 219 (ensemble grids 128^3 / 192^3, d = 5 / 8);
* 1024-point x and y lines (HD 1024^3).
"""
import numpy as np
import pytest

from conftest import TABLES, rel_err, rel_l2

pytestmark = pytest.mark.gpu


def _grid_plan(shape, precision=0):
    from oracle import specter_ref as O
    from specter_amd import FCPlan
    nx, ny, nz, C, d = shape
    return O, O.make_grid(nx, ny, nz, C, d, TABLES), FCPlan(nx, ny, nz, C, d,
                                                             precision=precision)


def _cu(a):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a)).cuda()


Z_PATHS = [
    # shape, z algorithm (0 direct, 1 Bluestein, 2 Rader), tables in global memory
    ((8, 8, 281, 25, 5), 2, 1),
    ((8, 8, 539, 27, 8), 0, 0),
    ((8, 8, 537, 25, 5), 1, None),
    ((8, 8, 153, 25, 5), 1, None),
    ((8, 8, 155, 27, 8), None, None),
    ((8, 8, 217, 25, 5), None, None),
    ((8, 8, 219, 27, 8), None, None),
]


@pytest.mark.parametrize("shape,alg,gtab", Z_PATHS, ids=lambda s: "x".join(map(str, s))
                         if isinstance(s, tuple) else str(s))
def test_z_paths_transforms_and_projection(shape, alg, gtab):
    O, g, p = _grid_plan(shape)
    info = p.info()
    print(f"{shape}: z algorithm {info.z_algorithm}, FFT length {info.z_fft_len}, "
          f"{info.z_threads_per_line} threads/line, {info.z_lines_per_block} lines/block, "
          f"global tables {info.z_tables_global}")
    if alg is not None:
        assert info.z_algorithm == alg
    if gtab is not None:
        assert info.z_tables_global == gtab
    rng = np.random.default_rng(5)
    R = rng.standard_normal(g.real_shape)
    out = p.empty_spectral()
    p.fftp3d_real_to_complex(_cu(R), out)
    assert rel_err(out.cpu().numpy(), O.fftp3d_real_to_complex(g, R.copy())) < 1e-12
    S = rng.standard_normal(g.spec_shape) + 1j * rng.standard_normal(g.spec_shape)
    outr = p.empty_real()
    p.fftp3d_complex_to_real(_cu(S), outr)
    assert rel_err(outr.cpu().numpy(), O.fftp3d_complex_to_real(g, S.copy())) < 1e-12
    # walls + projection from random spectral fields.  The pressure assembly
    # and (d=8) the continuation of white noise amplify roundoff, so the bound
    # is 10x the spread of the oracle's own result under a 1-ulp perturbation
    # of its inputs (and at least 1e-12)
    vx, vy, vz, pr = (rng.standard_normal(g.spec_shape) + 1j * rng.standard_normal(g.spec_shape)
                      for _ in range(4))
    r = [x.copy() for x in (vx, vy, vz, pr)]
    O.v_imposebc_and_project(g, *r, 3, 4, (0.0, 0.0), (0.0, 0.0))
    r2 = [x * (1 + 2.0 ** -53 * rng.standard_normal(x.shape)) for x in (vx, vy, vz, pr)]
    O.v_imposebc_and_project(g, *r2, 3, 4, (0.0, 0.0), (0.0, 0.0))
    ts = [_cu(x) for x in (vx, vy, vz, pr)]
    p.v_imposebc_and_project(*ts, 3, 4)
    for name, t, rr, rp in zip(("vx", "vy", "vz", "pr"), ts, r, r2):
        err, fl = rel_err(t.cpu().numpy(), rr), rel_err(rp, rr)
        print(f"{shape} {name}: gpu-vs-oracle {err:.2e}, oracle 1-ulp spread {fl:.2e}")
        assert err <= max(1e-12, 10 * fl)


def _one_step(shape, precision, bouss, seed=3, perturb=False):
    from oracle import specter_ref as O
    from specter_amd import FCPlan, Solver, SolverParams, ic
    nx, ny, nz, C, d = shape
    g = O.make_grid(nx, ny, nz, C, d, TABLES)
    if bouss:
        prm = dict(solver="BOUSS", ord=4, dt=1e-3, nu=2e-3, kappa=2e-3, xmom=1.0, xtemp=1.0)
    else:
        prm = dict(solver="HD", ord=4, dt=1e-3, nu=2e-3)
    sp = SolverParams(**prm)
    sol = Solver(FCPlan(nx, ny, nz, C, d, precision=precision), sp)
    u = ic.spectrum_peaked_velocity(nx, ny, nz, C, seed)
    zero = np.zeros(g.spec_shape, complex)
    st = O.HDState(*(O.fftp3d_real_to_complex(g, a.copy()) for a in u), zero.copy(),
                   zero.copy(), zero.copy(), zero.copy())
    for name, arr in zip(("vx", "vy", "vz"), u):
        sol.set_real_field(name, arr)
    names = ["vx", "vy", "vz"]
    if bouss:
        th = ic.thermal_perturbation(nx, ny, nz, C, seed + 1)
        sol.set_real_field("th", th)
        st.th = O.fftp3d_real_to_complex(g, th.copy())
        st.fs = zero.copy()
        names.append("th")
        O.bouss_step(g, st, sp.ord, sp.dt, sp.nu, sp.kappa, sp.xmom, sp.xtemp)
    else:
        O.hd_step(g, st, sp.ord, sp.dt, sp.nu)
    sol.step(1)
    errs = {n: rel_l2(sol.spectral(n), getattr(st, n)) for n in names}
    print(f"{shape} precision {precision} bouss {bouss}: " +
          ", ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    if not perturb:
        return errs
    # the oracle's own spread: the same step from inputs perturbed by one ulp
    rng = np.random.default_rng(17)
    st2 = O.HDState(*(O.fftp3d_real_to_complex(g, a.copy()) for a in u), zero.copy(),
                    zero.copy(), zero.copy(), zero.copy())
    for n_ in ("vx", "vy", "vz"):
        a = getattr(st2, n_)
        setattr(st2, n_, a * (1 + 2.0 ** -53 * rng.standard_normal(a.shape)))
    if bouss:
        st2.th = O.fftp3d_real_to_complex(g, th.copy())
        st2.th = st2.th * (1 + 2.0 ** -53 * rng.standard_normal(st2.th.shape))
        st2.fs = zero.copy()
        O.bouss_step(g, st2, sp.ord, sp.dt, sp.nu, sp.kappa, sp.xmom, sp.xtemp)
    else:
        O.hd_step(g, st2, sp.ord, sp.dt, sp.nu)
    floor = {n: rel_l2(getattr(st2, n), getattr(st, n)) for n in names}
    print("  oracle 1-ulp spread: " + ", ".join(f"{k} {v:.2e}" for k, v in floor.items()))
    return errs, floor


@pytest.mark.parametrize("shape", [(8, 8, 281, 25, 5), (8, 8, 537, 25, 5), (8, 8, 217, 25, 5)],
                         ids=lambda s: "x".join(map(str, s)))
def test_z_paths_hd_rk4_step(shape):
    for prec, tol in ((0, 1e-11), (1, 1e-5)):
        for n, e in _one_step(shape, prec, False).items():
            assert e < tol, (prec, n, e)


@pytest.mark.parametrize("shape", [(8, 8, 539, 27, 8), (8, 8, 155, 27, 8)],
                         ids=lambda s: "x".join(map(str, s)))
def test_z_paths_bouss_rk4_step(shape):
    """d=8 Boussinesq from a unit-rms turbulent field: float64 within 10x the
    oracle's 1-ulp spread (the d=8 continuation and the pressure assembly
    amplify roundoff); float32-mixed (the Boussinesq 512^3 workload's mode)
    states its error against the float64 oracle, bounded by 1e-3 (a d=8
    continuation multiplies the float32 noise of the nonlinear term by
    |A Q^T| = 2.2e6)."""
    errs, floor = _one_step(shape, 0, True, perturb=True)
    for n, e in errs.items():
        assert e <= max(1e-12, 10 * floor[n]), (n, e, floor[n])
    for n, e in _one_step(shape, 1, True).items():
        assert e < 1e-3, ("mixed", n, e)


@pytest.mark.parametrize("shape", [(1024, 8, 41, 25, 5), (8, 1024, 41, 25, 5)],
                         ids=lambda s: "x".join(map(str, s)))
def test_1024_point_xy_lines(shape):
    O, g, p = _grid_plan(shape)
    rng = np.random.default_rng(8)
    R = rng.standard_normal(g.real_shape)
    out = p.empty_spectral()
    p.fftp3d_real_to_complex(_cu(R), out)
    assert rel_err(out.cpu().numpy(), O.fftp3d_real_to_complex(g, R.copy())) < 1e-12
    S = rng.standard_normal(g.spec_shape) + 1j * rng.standard_normal(g.spec_shape)
    outr = p.empty_real()
    p.fftp3d_complex_to_real(_cu(S), outr)
    assert rel_err(outr.cpu().numpy(), O.fftp3d_complex_to_real(g, S.copy())) < 1e-12
    for prec, tol in ((0, 1e-12), (1, 1e-5)):
        _, _, q = _grid_plan(shape, prec)
        a, b, c = (rng.standard_normal(g.spec_shape) + 1j * rng.standard_normal(g.spec_shape)
                   for _ in range(3))
        ref = O.gradre(g, a, b, c)
        outs = [q.empty_nonlinear() for _ in range(3)]
        q.gradre(_cu(a), _cu(b), _cu(c), *outs)
        for o, rr in zip(outs, ref):
            e = rel_err(o.cpu().numpy().astype(complex), rr)
            print(f"{shape} precision {prec} gradre: {e:.2e}")
            assert e < tol


def test_plan_device_bytes_match_allocations():
    """device_bytes counts every allocation of the plan (checked against the
    device's free memory), work_bytes what the solver adds on first use."""
    import torch
    from specter_amd import FCPlan, Solver, SolverParams
    # load every code object of the library first (module loading maps its
    # static device arrays)
    for prec in (0, 1):
        w = FCPlan(16, 16, 41, 25, 5, precision=prec)
        Solver(w, SolverParams(solver="HD", ord=2, dt=1e-3, nu=2e-3)).step(1)
        w.sync()
    torch.cuda.synchronize()
    for prec in (0, 1):
        free0 = torch.cuda.mem_get_info()[0]
        p = FCPlan(128, 128, 153, 25, 5, precision=prec, members=4)
        info = p.info()
        free1 = torch.cuda.mem_get_info()[0]
        used = free0 - free1
        print(f"precision {prec}: device_bytes {info.device_bytes}, free-memory drop {used}")
        # small tables share pooled pages, large arrays round up to pages
        assert abs(used - info.device_bytes) <= 0.05 * info.device_bytes + (16 << 20)
        before = info.device_bytes + info.work_bytes + info.state_bytes
        sol = Solver(p, SolverParams(solver="HD", ord=2, dt=1e-3, nu=2e-3))
        after = p.info()
        assert after.work_bytes == 0
        assert after.device_bytes == info.device_bytes + info.work_bytes
        assert before == after.device_bytes + after.state_bytes
        sol.close()
        p.close()


@pytest.mark.parametrize("shape,plan", [((8, 8, 537, 25, 5), (0, 0)), ((8, 8, 539, 27, 8), (1, 1)),
                                        ((8, 8, 1049, 25, 5), (-1, 2))],
                         ids=["537-bluestein", "539-direct", "1049-bluestein-float32"])
def test_wave_plan_matches_engine(monkeypatch, shape, plan):
    """nz = 537 (the 512^3 grids, d = 5), 539 (d = 8) and 1049 (1024^3; the
    float32 lines only: eight float64 lines of 2100 points do not fit in
    LDS) run their z transforms on the compiled one-wavefront-per-line plans
    (Bluestein 1080 = 18*6*10; direct 539 = 11*7*7; Bluestein 2100 =
    15*20*7); with SPX_WAVE=0 at plan creation the general engine runs them.  Both match the oracle, and each other to
    rounding, in both precisions (the state transforms in float64, gradre's
    transforms in the plan's precision; d = 8 gradre in float32 carries the
    continuation's amplification of float32 noise)."""
    O, g, _ = _grid_plan(shape)
    rng = np.random.default_rng(11)
    R = rng.standard_normal(g.real_shape)
    S = rng.standard_normal(g.spec_shape) + 1j * rng.standard_normal(g.spec_shape)
    a, b, c = (rng.standard_normal(g.spec_shape) + 1j * rng.standard_normal(g.spec_shape)
               for _ in range(3))
    ref_r2c = O.fftp3d_real_to_complex(g, R.copy())
    ref_c2r = O.fftp3d_complex_to_real(g, S.copy())
    ref_grad = O.gradre(g, a, b, c)
    res = {}
    for wave in (True, False):
        if wave:
            monkeypatch.delenv("SPX_WAVE", raising=False)
        else:
            monkeypatch.setenv("SPX_WAVE", "0")
        for prec in (0, 1):
            _, _, p = _grid_plan(shape, prec)
            info = p.info()
            assert info.z_wave_plan == (plan[0] if wave else -1)
            assert info.z_wave_plan_lo == (plan[1] if wave and prec == 1 else -1)
            out = p.empty_spectral()
            p.fftp3d_real_to_complex(_cu(R), out)
            outr = p.empty_real()
            p.fftp3d_complex_to_real(_cu(S), outr)
            outs = [p.empty_nonlinear() for _ in range(3)]
            p.gradre(_cu(a), _cu(b), _cu(c), *outs)
            res[wave, prec] = (out.cpu().numpy(), outr.cpu().numpy(),
                               [o.cpu().numpy().astype(complex) for o in outs])
    d = shape[4]
    for prec, tol in ((0, 1e-12), (1, 1e-5 if d == 5 else 1e-2)):
        w, e = res[True, prec], res[False, prec]
        assert rel_err(w[0], ref_r2c) < 1e-12 and rel_err(w[1], ref_c2r) < 1e-12
        for o, r in zip(w[2], ref_grad):
            assert rel_err(o, r) < tol
        agree = max(rel_err(x, y) for x, y in zip(w[2], e[2]))
        print(f"precision {prec}: wave vs engine r2c {rel_err(w[0], e[0]):.2e}, "
              f"c2r {rel_err(w[1], e[1]):.2e}, gradre {agree:.2e}")
        assert rel_err(w[0], e[0]) < 1e-14 and rel_err(w[1], e[1]) < 1e-14
        assert agree < (1e-13 if prec == 0 else (2e-6 if d == 5 else 1e-3))
