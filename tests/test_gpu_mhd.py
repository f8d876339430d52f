"""GPU parity of the MHD pieces and the MHD solver against the float64 oracle
(include/mhd/mhd_rkstep2.f90, pseudo/pseudospec_hd.f90 curlk / prodre,
pseudo/pseudospec_mhd.f90 vector, boundary/bboundary.f90 and sol_project with
the potential's conditions, boundary_mod.fpp:197-678), through the C ABI.

Multi-step parity uses the bound of the other solver tests: the oracle run
from inputs perturbed by one ulp of their rms sets the spread, and the GPU
must stay within 10x that spread (or 1e-12)."""
import numpy as np
import pytest

from conftest import TABLES, rel_err, rel_l2

pytestmark = pytest.mark.gpu
ULP = 2.0 ** -53
SHAPES = [(16, 16, 41, 25, 5), (16, 12, 57, 25, 5), (12, 16, 91, 27, 8)]


def _ids(s):
    return "x".join(map(str, s))


def _tol(shape):
    """d = 8 continuations amplify roundoff (row sums ~2e6): 1e-10 there."""
    return 1e-12 if shape[4] <= 5 else 1e-10


def _check(t, ref, tol):
    err = rel_err(t.cpu().numpy(), ref)
    assert err < tol, err


def _setup(shape):
    import torch
    from oracle import specter_ref as O
    from specter_amd import FCPlan
    nx, ny, nz, C, d = shape
    g = O.make_grid(nx, ny, nz, C, d, TABLES)
    p = FCPlan(nx, ny, nz, C, d)
    assert p.neumann
    neu, neu2 = O.neumann_vector(g, TABLES, 1), O.neumann_vector(g, TABLES, 2)
    return torch, O, g, p, neu, neu2


def _cu(torch, a):
    return torch.as_tensor(np.ascontiguousarray(a)).cuda()


def _rand_spec(g, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(g.spec_shape) + 1j * rng.standard_normal(g.spec_shape)


def _smooth(O, g, seed, amp=1.0):
    """Smooth real fields (transformed), so the products do not amplify
    the continuation of white noise."""
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(g.z, g.y, g.x, indexing="ij")
    out = []
    for c in range(3):
        a, b, ph = rng.uniform(0.5, 1.5, 3)
        r = amp * (a * np.sin(x + ph) * np.cos(2 * y) * np.sin(3 * z + c)
                   + b * np.cos(2 * x - y) * np.sin(np.pi * z) + 0.3 * np.cos(z + ph))
        out.append(O.fftp3d_real_to_complex(g, r))
    return out


@pytest.mark.parametrize("shape", SHAPES[:2], ids=_ids)
def test_curlk(shape):
    torch, O, g, p, *_ = _setup(shape)
    a, b = _rand_spec(g, 1), _rand_spec(g, 2)
    for d in (1, 2, 3):
        out = p.empty_spectral()
        p.curlk(_cu(torch, a), _cu(torch, b), out, d)
        diff = np.abs(out.cpu().numpy() - O.curlk(g, a, b, d)).max()
        assert diff == 0.0, (d, diff)


@pytest.mark.parametrize("shape", SHAPES, ids=_ids)
def test_prodre_and_vector(shape):
    torch, O, g, p, *_ = _setup(shape)
    v = _smooth(O, g, 3)
    w = _smooth(O, g, 4)
    ref = O.prodre(g, *v)
    outs = [p.empty_spectral() for _ in range(3)]
    p.prodre(*(_cu(torch, x) for x in v), *outs)
    for o, r in zip(outs, ref):
        _check(o, r, _tol(shape))
    ref = O.vector(g, *v, *w)
    outs = [p.empty_spectral() for _ in range(3)]
    p.vector(*(_cu(torch, x) for x in v), *(_cu(torch, x) for x in w), *outs)
    for o, r in zip(outs, ref):
        _check(o, r, _tol(shape))


@pytest.mark.parametrize("kinds", [(0, 0), (2, 2), (0, 2)])
@pytest.mark.parametrize("shape", SHAPES, ids=_ids)
def test_sol_project_potential(shape, kinds):
    torch, O, g, p, *_ = _setup(shape)
    a, b, c = (_rand_spec(g, s) for s in (5, 6, 7))
    ra, rb, rc_ = a.copy(), b.copy(), c.copy()
    rd = np.zeros_like(a)
    O.sol_project(g, ra, rb, rc_, rd, 0, *kinds)
    ta, tb, tc = (_cu(torch, x) for x in (a, b, c))
    td = p.empty_spectral()
    p.sol_project(ta, tb, tc, td, 0, *kinds)
    for t, r in ((ta, ra), (tb, rb), (tc, rc_), (td, rd)):
        _check(t, r, _tol(shape))


def test_sol_project_refuses_robin_bottom_dirichlet_top():
    torch, O, g, p, *_ = _setup(SHAPES[0])
    from specter_amd.lib import SpxError
    t = [p.empty_spectral() for _ in range(4)]
    with pytest.raises(SpxError):
        p.sol_project(*t, 0, 2, 0)


@pytest.mark.parametrize("kinds", [(0, 0), (1, 1), (0, 1)])
@pytest.mark.parametrize("shape", SHAPES, ids=_ids)
def test_a_imposebc_and_project(shape, kinds):
    """Against the oracle, within 10x the oracle's own spread under a 1-ulp
    input perturbation (two wall round trips through the continuation
    amplify roundoff by its row sums)."""
    torch, O, g, p, neu, neu2 = _setup(shape)
    a = _smooth(O, g, 8)

    def run(inp):
        ref = [x.copy() for x in inp]
        rph = np.zeros_like(inp[0])
        O.a_imposebc_and_project(g, neu, neu2, *ref, rph, *kinds)
        return ref + [rph]

    rng = np.random.default_rng(5)
    ref, ref2 = run(a), run([_perturb(x, rng) for x in a])
    t = [_cu(torch, x) for x in a]
    ph = p.empty_spectral()
    p.a_imposebc_and_project(*t, ph, *kinds)
    for x, r, r2 in zip(t + [ph], ref, ref2):
        err, fl = rel_err(x.cpu().numpy(), r), rel_err(r2, r)
        print(f"gpu-vs-oracle {err:.2e}  oracle 1-ulp spread {fl:.2e}")
        assert err <= max(1e-12, 10 * fl), (err, fl)


def _perturb(a, rng):
    s = np.sqrt(np.mean(np.abs(a) ** 2))
    return a + ULP * s * (rng.standard_normal(a.shape) + 1j * rng.standard_normal(a.shape))


def _mhd_pair(shape, prm, nsteps, seed=11):
    from oracle import specter_ref as O
    from specter_amd import FCPlan, Solver, SolverParams, ic
    nx, ny, nz, C, d = shape
    g = O.make_grid(nx, ny, nz, C, d, TABLES)
    neu, neu2 = O.neumann_vector(g, TABLES, 1), O.neumann_vector(g, TABLES, 2)
    plan = FCPlan(nx, ny, nz, C, d)
    sp = SolverParams(solver="MHD", **prm)
    sol = Solver(plan, sp)
    u = ic.spectrum_peaked_velocity(nx, ny, nz, C, seed)
    z, y, x = np.meshgrid(g.z, g.y, g.x, indexing="ij")
    L = float(g.z[g.np_ - 1])
    s = np.sin(np.pi * z / L)
    ar = [0.3 * s * np.cos(y), 0.3 * s * np.sin(x + y), 0.3 * np.cos(np.pi * z / L) * np.cos(x)]
    fr = ic.large_scale_forcing(nx, ny, nz, C)
    mr = [0.05 * s * np.sin(y), 0 * x, 0 * x]
    for name, arr in zip(("vx", "vy", "vz", "fx", "fy", "fz", "ax", "ay", "az", "mx", "my",
                          "mz"), (*u, *fr, *ar, *mr)):
        sol.set_real_field(name, arr)
    sp_ = lambda r: O.fftp3d_real_to_complex(g, r.copy())  # noqa: E731
    zero = np.zeros(g.spec_shape, complex)
    st = O.MHDState(*(sp_(r) for r in u), zero.copy(), *(sp_(r) for r in fr),
                    *(sp_(r) for r in ar), zero.copy(), *(sp_(r) for r in mr))
    st2 = st.copy()
    rng = np.random.default_rng(99)
    for name in ("vx", "vy", "vz", "ax", "ay", "az"):
        setattr(st2, name, _perturb(getattr(st2, name), rng))
    for _ in range(nsteps):
        sol.step(1)
        for s_ in (st, st2):
            O.mhd_step(g, s_, sp.ord, sp.dt, sp.nu, sp.mu, sp.b0, neu, neu2, sp.bzsta, sp.bzend)
    return sol, st, st2


@pytest.mark.parametrize("walls", [(0, 0), (1, 1), (0, 1)])
@pytest.mark.parametrize("shape", SHAPES[:2], ids=_ids)
def test_mhd_steps(shape, walls):
    prm = dict(ord=2, dt=2e-3, nu=5e-2, mu=4e-2, b0=(0.4, -0.2, 0.3), bzsta=walls[0],
               bzend=walls[1])
    sol, st, st2 = _mhd_pair(shape, prm, 3)
    for name in ("vx", "vy", "vz", "ax", "ay", "az", "pr", "ph"):
        err = rel_l2(sol.spectral(name), getattr(st, name))
        fl = rel_l2(getattr(st2, name), getattr(st, name))
        print(f"{name}: gpu-vs-oracle {err:.2e}  oracle 1-ulp spread {fl:.2e}")
        assert err <= max(1e-12, 10 * fl), name


def test_mhd_rk4_d8_and_graph_replays():
    """RK4 at d = 8: the captured step equals the eager one bit for bit, over
    replays with eager work in between, and both follow the oracle."""
    prm = dict(ord=4, dt=1e-3, nu=5e-2, mu=4e-2, b0=(0.0, 0.3, 0.0), bzsta=0, bzend=0)
    sol, st, st2 = _mhd_pair(SHAPES[2], prm, 2)
    for name in ("vx", "ax"):
        err = rel_l2(sol.spectral(name), getattr(st, name))
        fl = rel_l2(getattr(st2, name), getattr(st, name))
        assert err <= max(1e-12, 10 * fl), name
    from specter_amd import FCPlan, Solver, SolverParams
    nx, ny, nz, C, d = SHAPES[2]
    runs = []
    for graph in (False, True):
        plan = FCPlan(nx, ny, nz, C, d)
        s = Solver(plan, SolverParams(solver="MHD", **prm))
        s.use_graph(graph)
        for name in ("vx", "vy", "vz", "ax", "ay", "az"):
            s.field(name).copy_(sol.field(name))
        out = []
        for _ in range(3):
            s.step(1)
            plan.sync()
            out.append([s.spectral(n) for n in ("vx", "ax", "ph")])
        runs.append(out)
    for a, b in zip(*runs):
        for x, y in zip(a, b):
            assert np.abs(x - y).max() == 0.0


def test_mhd_state_refusals():
    from specter_amd import FCPlan, Solver, SolverParams
    from specter_amd import lib as L
    with pytest.raises(L.SpxError):
        Solver(FCPlan(16, 16, 41, 25, 5, precision=L.SPX_FP32_MIXED),
               SolverParams(solver="MHD", mu=0.1))
    with pytest.raises(L.SpxError):
        Solver(FCPlan(16, 16, 41, 25, 5), SolverParams(solver="MHD", bzsta=1, bzend=0))


@pytest.mark.parametrize("shape", SHAPES[:2], ids=_ids)
def test_mhdcheck_matches_oracle(shape):
    """The mhdcheck values (energy kin 0/1/2, helicities, cross helicity) of
    the same state on the GPU and in the oracle."""
    from oracle import specter_ref as O
    from specter_amd import FCPlan, Solver, SolverParams
    nx, ny, nz, C, d = shape
    g = O.make_grid(nx, ny, nz, C, d, TABLES)
    sol = Solver(FCPlan(nx, ny, nz, C, d), SolverParams(solver="MHD", mu=0.1))
    v, a = _smooth(O, g, 21), _smooth(O, g, 22, amp=0.5)
    for name, x in zip(("vx", "vy", "vz", "ax", "ay", "az"), (*v, *a)):
        sol.field(name).copy_(_cu(__import__("torch"), x))
    zero = np.zeros(g.spec_shape, complex)
    st = O.MHDState(*v, zero, zero, zero, zero, *a, zero, zero, zero, zero)
    got, ref = sol.mhdcheck(), O.mhdcheck(g, st)
    for k, (x, r) in enumerate(zip(got, ref)):
        assert abs(x - r) <= 1e-12 * max(abs(r), max(map(abs, ref))), (k, x, r)
