#!/usr/bin/env python3
"""RK-substep throughput of the SPECTER hot path on one MI355X, mixed
precision (float32 nonlinear term, float64 state / continuation / pressure /
walls) against the float64 path.

A "step" is one Runge-Kutta stage (one pass of the hot path: the nonlinear
term (v.grad)v with its z/y/x transforms and products, filter + Laplacian +
RK update, no-slip walls, pressure projection) over the whole grid.  Inputs
are resident in HBM before the timed region.  Both precisions are timed on
the same workload; ``value`` is the mixed-precision rate, and the line also
carries the float64 rate, the speed-up and the relative L2 field error of the
mixed run against the float64 run after ``--err-steps`` full RK steps from
the same initial field.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload hd256|hd64|hd1024|bouss512]
                    [--precision both|mixed|fp64] [--err-steps N]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E

WORKLOADS = {
    # HD 256^3 (z physical) d=5 C=25 RK4 nu=5e-4, forced
    "hd256": dict(nx=256, ny=256, nzp=256, C=25, d=5, ord=4, nu=5e-4, dt=5e-4, forcing=True,
                  label="HD 256x256x(256+25) FC-Gram d=5 C=25, RK4, nu=5e-4, dt=5e-4, "
                        "forcing 0.5(sin2y,sin2x,0)sin(pi z)"),
    # HD 1024^3 (one GPU's share of a production run) d=5 C=25 RK4 nu=1e-4, forced
    "hd1024": dict(nx=1024, ny=1024, nzp=1024, C=25, d=5, ord=4, nu=1e-4, dt=1e-4, forcing=True,
                   label="HD 1024x1024x(1024+25) FC-Gram d=5 C=25, RK4, nu=1e-4, dt=1e-4, "
                         "forcing 0.5(sin2y,sin2x,0)sin(pi z)"),
    # Boussinesq 512^3 d=8 C=27 RK4 nu=kappa=2e-4, heated bottom wall: theta is the
    # deviation from the conduction profile 1 - z/Lz (0 on both walls), gravity along
    # z (buoyancy xmom*theta in the vz equation, heat current xtemp*vz)
    "bouss512": dict(nx=512, ny=512, nzp=512, C=27, d=8, ord=4, nu=2e-4, kappa=2e-4, dt=1e-4,
                     forcing=False, solver="BOUSS", xmom=1.0, xtemp=1.0,
                     label="Boussinesq 512x512x(512+27) FC-Gram d=8 C=27, RK4, nu=kappa=2e-4, "
                           "dt=1e-4, heated bottom wall (theta = 1 - z/Lz + theta'), "
                           "theta' = 1e-3 white noise, v = 0"),
    # HD 64^3 d=5 C=25 RK2 nu=2e-3 dt=1e-3, unforced
    "hd64": dict(nx=64, ny=64, nzp=64, C=25, d=5, ord=2, nu=2e-3, dt=1e-3, forcing=False,
                 label="HD 64x64x(64+25) FC-Gram d=5 C=25, RK2, nu=2e-3, dt=1e-3"),
}


def _fields(w, seed=1000):
    import numpy as np
    from specter_amd import ic
    nx, ny, nz, C = w["nx"], w["ny"], w["nzp"] + w["C"], w["C"]
    u = ic.spectrum_peaked_velocity(nx, ny, nz, C, seed)
    zero = np.zeros((nz, ny, nx))
    f = ic.large_scale_forcing(nx, ny, nz, C) if w["forcing"] else (zero, zero, zero)
    return u, f


def _plan(w, precision):
    from specter_amd import FCPlan
    return FCPlan(w["nx"], w["ny"], w["nzp"] + w["C"], w["C"], w["d"], precision=precision)


def _solver(plan, w, u, f):
    from specter_amd import Solver, SolverParams
    sol = Solver(plan, SolverParams(solver="HD", ord=w["ord"], dt=w["dt"], nu=w["nu"]))
    for name, arr in zip(("vx", "vy", "vz", "fx", "fy", "fz"), (*u, *f)):
        sol.set_real_field(name, arr)
    return sol


def _run_substeps(sol, ordr, nsub, start=0):
    """nsub RK stages cycling o = ord..1, with the rkstep1 copy before o = ord.
    Whole steps go through spx_step (one captured HIP graph per step)."""
    if start % ordr == 0 and nsub % ordr == 0:
        sol.step(nsub // ordr)
        return start + nsub
    for s in range(start, start + nsub):
        o = ordr - (s % ordr)
        if o == ordr:
            sol.rkstep1()
        sol.substep(o)
    return start + nsub


def _profile(plan, sol, ordr, nsub, start):
    lib = plan._lib
    from specter_amd.lib import check
    check(lib.spx_prof_enable(plan.handle, 1))
    _run_substeps(sol, ordr, nsub, start)
    nk = ctypes.c_int()
    check(lib.spx_prof_resolve(plan.handle, ctypes.byref(nk)))
    rows = []
    for i in range(nk.value):
        name = ctypes.create_string_buffer(64)
        n = ctypes.c_long()
        ms = ctypes.c_double()
        by = ctypes.c_double()
        check(lib.spx_prof_get(plan.handle, i, name, 64, ctypes.byref(n), ctypes.byref(ms),
                               ctypes.byref(by)))
        rows.append((name.value.decode(), n.value, ms.value, by.value))
    check(lib.spx_prof_enable(plan.handle, 0))
    return rows


def _cpu_baseline(w, u, f, nsub):
    """The float64 oracle (numpy/scipy restatement) on the host cores, on a
    bounded sample of the same workload: nsub RK stages."""
    import numpy as np
    workers = int(os.environ.get("OMP_NUM_THREADS", str(os.cpu_count() or 1)))
    os.environ["SPX_ORACLE_WORKERS"] = str(workers)
    from oracle import specter_ref as O
    nx, ny, nz, C, d = w["nx"], w["ny"], w["nzp"] + w["C"], w["C"], w["d"]
    g = O.make_grid(nx, ny, nz, C, d, os.path.join(ROOT, "specter_amd", "tables"))
    st = O.HDState(*(O.fftp3d_real_to_complex(g, a.copy()) for a in u),
                   np.zeros(g.spec_shape, complex),
                   *(O.fftp3d_real_to_complex(g, a.copy()) for a in f))
    C1, C2, C3 = st.vx.copy(), st.vy.copy(), st.vz.copy()
    t0 = time.perf_counter()
    for s in range(nsub):
        O.hd_substep(g, st, C1, C2, C3, w["ord"] - (s % w["ord"]), w["ord"], w["dt"], w["nu"])
    dt = (time.perf_counter() - t0) / nsub
    return dt, workers


def _time(sol, w, steps, warmup):
    """ms per RK stage over `steps` stages after `warmup` untimed ones (HIP
    events on the solver's stream, which is torch's current stream)."""
    import torch
    pos = _run_substeps(sol, w["ord"], warmup)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    pos = _run_substeps(sol, w["ord"], steps, pos)
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps, time.perf_counter() - t0, pos


def _kernels(plan, sol, w, steps, pos):
    rows = _profile(plan, sol, w["ord"], steps, pos)
    rows.sort(key=lambda r: -r[2])
    table = {r[0]: {"launches_per_step": round(r[1] / steps, 2),
                    "ms_per_step": round(r[2] / steps, 4),
                    "GBps": round(r[3] / (r[2] * 1e-3) / 1e9, 1)} for r in rows}
    return rows, table


def _measured_traffic(workload, mode):
    """HBM bytes per launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of the 256^3 workload (read + write, calibrated on a
    plain copy); {} for other workloads or when the summary is absent."""
    path = os.path.join(ROOT, "profiles", "hd256_hbm_traffic.md")
    if workload != "hd256" or not os.path.exists(path):
        return {}
    out = {}
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 4 and cells[1] == mode:
            try:
                out[cells[0]] = (float(cells[2]) + float(cells[3])) * 2 ** 20
            except ValueError:
                pass
    return out


def _roofline(rows, traffic):
    name, n, tot_ms, tot_b = rows[0]
    avg_ms = tot_ms / n
    ach = (tot_b / n) / (avg_ms * 1e-3) / 1e9
    t = traffic.get(name)
    return {"kernel": name, "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
            "traffic": round(t) if t else None,
            "avg_launch_ms": round(avg_ms, 5), "algorithmic_bytes_per_launch": tot_b / n}


def _field_error(w, u, f, nsteps, plans):
    """Errors of the mixed run against the float64 run after nsteps full RK
    steps from the same fields: relative L2 of each velocity component, of
    the horizontal energy spectrum, of the dissipation nu*<|omega|^2> (the
    reference's enstrophy) and of the mean wall shear."""
    import numpy as np
    sols = {k: _solver(p, w, u, f) for k, p in plans.items()}
    for s in sols.values():
        s.step(nsteps)
    out = {}
    for name in ("vx", "vy", "vz"):
        a = sols["mixed"].spectral(name)
        b = sols["fp64"].spectral(name)
        out[name] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    q = {}
    for k, s in sols.items():
        p = s.plan
        vx, vy, vz = s.field("vx"), s.field("vy"), s.field("vz")
        q[k] = (np.array(p.energy_spectrum(vx, vy, vz, w["nx"] // 2)),
                w["nu"] * p.energy(vx, vy, vz, 0),
                np.array(p.wall_shear(vx, vy, w["nu"])))
    (sm, dm, tm), (sf, df, tf) = q["mixed"], q["fp64"]
    out["energy_spectrum"] = float(np.linalg.norm(sm - sf) / np.linalg.norm(sf))
    out["dissipation"] = float(abs(dm - df) / abs(df))
    out["wall_shear"] = float(np.abs(tm - tf).max() / np.abs(tf).max())
    for s in sols.values():
        s.close()
    return out


def _ensemble(nruns, max_slots, chunk):
    """The ensemble workload (Poisson(4/s) arrivals of HD RK2 runs, 300 steps,
    grids {32,64,128,192}^3, d in {5,8}, nu log-uniform [5e-4,5e-3], float32
    mixed) through the continuous-batching server, against the same stream
    served one run at a time in float64 (FIFO, service times measured per
    grid class on the GPU)."""
    import numpy as np
    import torch
    from specter_amd.ensemble import EnsembleServer, latency_stats, poisson_stream
    from specter_amd import FCPlan, Solver, SolverParams, ic
    from specter_amd.lib import SPX_FP64
    reqs = poisson_stream(nruns=nruns)
    work = sum(r.points * r.nsteps * r.ord for r in reqs)
    srv = EnsembleServer(max_slots=max_slots, chunk=chunk)
    t0 = time.perf_counter()
    res = srv.run(reqs, realtime=True)
    wall = time.perf_counter() - t0
    st = latency_stats(res)
    # float64 service time per step of each grid class, one run alone
    per_step = {}
    for r in reqs:
        key = r.grid
        if key in per_step:
            continue
        nx, ny, nz, C, d = key
        plan = FCPlan(nx, ny, nz, C, d, precision=SPX_FP64)
        sol = Solver(plan, SolverParams(solver="HD", ord=r.ord, dt=r.dt, nu=r.nu))
        for name, arr in zip(("vx", "vy", "vz"), ic.spectrum_peaked_velocity(nx, ny, nz, C, 1)):
            sol.set_real_field(name, arr)
        sol.step(3)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        sol.step(20)
        torch.cuda.synchronize()
        per_step[key] = (time.perf_counter() - t1) / 20
        sol.close()
        plan.close()
    # FIFO single-run service of the same arrival stream
    clock, lat = 0.0, []
    for r in sorted(reqs, key=lambda q: q.arrival):
        clock = max(clock, r.arrival) + per_step[r.grid] * r.nsteps
        lat.append(clock - r.arrival)
    seq_busy = sum(per_step[r.grid] * r.nsteps for r in reqs)
    lat = np.array(lat)
    return {
        "runs": nruns, "workload": "Poisson 4/s, 300 RK2 steps, {32,64,128,192}^3, d in {5,8}, "
                                   "nu log-U[5e-4,5e-3], FP32-mixed, continuous batching",
        "p50_s": round(st["p50_s"], 4), "p99_s": round(st["p99_s"], 4),
        "makespan_s": round(wall, 3),
        "aggregate_Mpts_per_s": round(work / wall / 1e6, 2),
        "peak_classes": srv.peak_classes, "peak_hbm_GB": round(srv.peak_used / 1e9, 2),
        "sequential_fp64": {"p50_s": round(float(np.percentile(lat, 50)), 4),
                            "p99_s": round(float(np.percentile(lat, 99)), 4),
                            "makespan_s": round(clock, 3),
                            "aggregate_Mpts_per_s": round(work / clock / 1e6, 2),
                            "busy_s": round(seq_busy, 3)},
        "throughput_vs_sequential_fp64": round(clock / wall, 3),
    }


def _log(msg):
    print(f"[bench] {time.strftime('%H:%M:%S')} {msg}", file=sys.stderr, flush=True)


def _bench_large(args, w):
    """The 1024^3 HD and 512^3 Boussinesq workloads: the mixed run and the
    float64 run go one after the other from the same initial state (kept on
    the host; at 1024^3 the two do not fit together); each runs --err-steps
    full steps for the error comparison, then is timed."""
    import numpy as np
    import torch
    from specter_amd import ic
    from specter_amd.lib import SPX_FP32_MIXED, SPX_FP64, SpxError
    nx, ny, nz, C = w["nx"], w["ny"], w["nzp"] + w["C"], w["C"]
    npts = nx * ny * w["nzp"]
    bouss = w.get("solver") == "BOUSS"
    _log("initial fields")
    if bouss:
        names = ("th",)
        fields = [ic.thermal_perturbation(nx, ny, nz, C, 1001)]
    else:
        names = ("vx", "vy", "vz", "fx", "fy", "fz")
        fields = [*ic.spectrum_peaked_velocity_device(nx, ny, nz, C, 1000),
                  *ic.large_scale_forcing_device(nx, ny, nz, C)]
        torch.cuda.empty_cache()   # the generator's temporaries go back to the device
    out_names = ("vx", "vy", "vz", "th") if bouss else ("vx", "vy", "vz")
    res, host_ic, diag = {}, None, {}
    for mode in ("mixed", "fp64"):
        prec = SPX_FP32_MIXED if mode == "mixed" else SPX_FP64
        try:
            plan = _plan(w, prec)
            from specter_amd import Solver, SolverParams
            prm = dict(solver="BOUSS", kappa=w["kappa"], xmom=w["xmom"],
                       xtemp=w["xtemp"]) if bouss else dict(solver="HD")
            sol = Solver(plan, SolverParams(ord=w["ord"], dt=w["dt"], nu=w["nu"], **prm))
        except SpxError as e:
            res[mode] = {"error": str(e)}
            _log(f"{mode}: {e}")
            continue
        if host_ic is None:
            for i, name in enumerate(names):
                sol.set_real_field(name, fields[i])
                fields[i] = None
                torch.cuda.empty_cache()
            host_ic = {n: sol.spectral(n) for n in names}
        else:
            for n in names:
                sol.field(n).copy_(torch.from_numpy(host_ic[n]))
            plan.sync()
        _log(f"{mode}: {args.err_steps} steps")
        sol.step(args.err_steps)
        plan.sync()
        vx, vy, vz = sol.field("vx"), sol.field("vy"), sol.field("vz")
        diag[mode] = {"spectrum": np.array(plan.energy_spectrum(vx, vy, vz, nx // 2)),
                      "dissipation": w["nu"] * plan.energy(vx, vy, vz, 0),
                      "wall_shear": np.array(plan.wall_shear(vx, vy, w["nu"])),
                      "fields": {n: sol.spectral(n) for n in out_names}}
        if bouss:
            diag[mode]["nusselt"] = np.array(plan.nusselt(sol.field("th")))
        _log(f"{mode}: timing")
        ms, wall, pos = _time(sol, w, args.steps, args.warmup)
        r = {"ms_per_step": ms, "value": npts / (ms * 1e-3) / 1e6}
        if not args.no_profile:
            r["rows"], r["kernels"] = _kernels(plan, sol, w, args.steps, pos)
        res[mode] = r
        sol.close()
        plan.close()
        del sol, plan
        torch.cuda.empty_cache()
    hr = res["mixed"]
    if "value" not in hr:
        print(json.dumps({"metric": "RK substep Mpts/s", "value": None, "unit": "Mpts/s",
                          "workload": w["label"], "error": res}), flush=True)
        return
    line = {
        "metric": "RK substep Mpts/s (physical points per RK stage per second), FP32-mixed",
        "value": round(hr["value"], 3), "unit": "Mpts/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(hr["ms_per_step"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 nonlinear term / f64 state, continuation, pressure, walls",
        "precision_mode": "mixed",
        "data": "synthetic (seeded white-noise theta', v = 0)" if bouss else
                "synthetic (seeded spectrum-peaked solenoidal IC generated on the GPU "
                "x sin(pi z/Lz))",
        "config": {"workload": w["label"], "grid": [nx, ny, nz], "physical_points": npts,
                   "parallelism": "single GPU"},
    }
    if not args.no_profile:
        line["roofline"] = _roofline(hr["rows"], {})
        line["kernels"] = hr["kernels"]
    f64 = res.get("fp64", {})
    if "ms_per_step" in f64:
        line["fp64"] = {"value": round(f64["value"], 3), "ms_per_step": round(f64["ms_per_step"], 4)}
        if not args.no_profile:
            line["fp64"]["kernels"] = f64["kernels"]
        line["speedup_vs_fp64"] = round(f64["ms_per_step"] / hr["ms_per_step"], 3)
        a, b = diag["mixed"], diag["fp64"]

        def rel(x, y, norm):
            d = norm(y)
            return float(f"{norm(x - y) / d:.3e}") if d > 0 else None

        l2 = np.linalg.norm
        amax = lambda v: float(np.abs(v).max())
        err = {"steps": args.err_steps}
        for n in out_names:
            err[n] = rel(a["fields"][n], b["fields"][n], l2)
        err["energy_spectrum"] = rel(a["spectrum"], b["spectrum"], l2)
        err["dissipation"] = rel(a["dissipation"], b["dissipation"], abs)
        err["wall_shear"] = rel(a["wall_shear"], b["wall_shear"], amax)
        if bouss:
            err["nusselt"] = rel(a["nusselt"], b["nusselt"], amax)
            line["nusselt_fp64"] = [round(float(x), 10) for x in b["nusselt"]]
        if args.err_steps > 0:
            line["error_vs_fp64"] = err
    else:
        line["fp64"] = f64
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8, help="timed RK stages")
    ap.add_argument("--warmup", type=int, default=4, help="untimed RK stages")
    ap.add_argument("--workload", default="hd256", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="both", choices=("both", "mixed", "fp64"))
    ap.add_argument("--err-steps", type=int, default=100,
                    help="full RK steps of the mixed-vs-float64 field comparison (0 = skip)")
    ap.add_argument("--cpu-substeps", type=int, default=1,
                    help="RK stages of the CPU baseline sample (0 = skip)")
    ap.add_argument("--no-profile", action="store_true", help="skip the per-kernel pass")
    ap.add_argument("--ensemble-runs", type=int, default=96,
                    help="requests of the ensemble-serving measurement (0 = skip)")
    ap.add_argument("--ensemble-slots", type=int, default=16)
    ap.add_argument("--ensemble-chunk", type=int, default=5, help="RK steps of a class's first chunk")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); the "
              "multi-GPU slab decomposition is a design, not a measured path", file=sys.stderr)
        sys.exit(2)
    from specter_amd.lib import SPX_FP32_MIXED, SPX_FP64
    w = WORKLOADS[args.workload]
    if args.workload in ("hd1024", "bouss512"):
        _bench_large(args, w)
        return
    u, f = _fields(w)
    npts = w["nx"] * w["ny"] * w["nzp"]
    ntot = w["nx"] * w["ny"] * (w["nzp"] + w["C"])
    modes = ["mixed", "fp64"] if args.precision == "both" else [args.precision]
    plans, res = {}, {}
    for mode in modes:
        plan = _plan(w, SPX_FP32_MIXED if mode == "mixed" else SPX_FP64)
        plans[mode] = plan
        sol = _solver(plan, w, u, f)
        ms, wall, pos = _time(sol, w, args.steps, args.warmup)
        r = {"ms_per_step": ms, "wall_s": wall, "value": npts / (ms * 1e-3) / 1e6}
        if not args.no_profile:
            r["rows"], r["kernels"] = _kernels(plan, sol, w, args.steps, pos)
        sol.close()
        res[mode] = r
    head = modes[0]
    hr = res[head]
    line = {
        "metric": "RK substep Mpts/s (physical points per RK stage per second), "
                  + ("FP32-mixed" if head == "mixed" else "FP64"),
        "value": round(hr["value"], 3),
        "unit": "Mpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(hr["ms_per_step"], 4),
        "wall_s": round(hr["wall_s"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 nonlinear term / f64 state, continuation, pressure, walls"
                 if head == "mixed" else "f64",
        "precision_mode": head,
        "data": "synthetic (seeded spectrum-peaked solenoidal IC x sin(pi z/Lz))",
        "config": {"workload": w["label"], "grid": [w["nx"], w["ny"], w["nzp"] + w["C"]],
                   "z_dft": {0: "direct", 1: "bluestein", 2: "rader"}[
                       plans[head].z_algorithm] + f" (in-LDS FFT length {plans[head].z_fft_len})",
                   "physical_points": npts, "total_points": ntot,
                   "total_Mpts_per_s": round(ntot / (hr["ms_per_step"] * 1e-3) / 1e6, 3),
                   "parallelism": "single GPU"},
    }
    if "fp64" in res and head == "mixed":
        line["fp64"] = {"value": round(res["fp64"]["value"], 3),
                        "ms_per_step": round(res["fp64"]["ms_per_step"], 4)}
        line["speedup_vs_fp64"] = round(res["fp64"]["ms_per_step"] / hr["ms_per_step"], 3)
        if args.err_steps > 0:
            err = _field_error(w, u, f, args.err_steps, plans)
            line["error_vs_fp64"] = {"steps": args.err_steps,
                                     **{k: float(f"{v:.3e}") for k, v in err.items()}}
    if not args.no_profile:
        line["roofline"] = _roofline(hr["rows"], _measured_traffic(args.workload, head))
        line["kernels"] = hr["kernels"]
        if "fp64" in res and head == "mixed":
            line["fp64"]["roofline"] = _roofline(res["fp64"]["rows"],
                                                 _measured_traffic(args.workload, "fp64"))
            line["fp64"]["kernels"] = res["fp64"]["kernels"]
    if args.ensemble_runs > 0:
        line["ensemble"] = _ensemble(args.ensemble_runs, args.ensemble_slots,
                                     args.ensemble_chunk)
    if args.cpu_substeps > 0:
        t, cores = _cpu_baseline(w, u, f, args.cpu_substeps)
        line["cpu_baseline"] = {"value": round(npts / t / 1e6, 4), "unit": "Mpts/s",
                                "cores": cores, "kind": "port",
                                "sample": f"{args.cpu_substeps} RK stage(s) of the same "
                                          f"workload, float64 oracle (scipy.fft workers="
                                          f"{cores}, numpy elementwise single-threaded), "
                                          f"{t:.2f} s per stage"}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
