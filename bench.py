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

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--workload hd512|hd256|hd64|hd1024|bouss512]
                    [--precision both|mixed|fp64] [--err-steps N]

Timing runs whole RK steps through spx_step (one captured HIP graph per
step): the warmup is rounded up to whole steps (the stages actually run are
reported as ``warmup_stages``), and the timed stages are whole steps when K
is a multiple of the RK order.
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

_FORCING = "forcing 0.5(sin2y,sin2x,0)sin(pi z)"
WORKLOADS = {
    # HD 512^3 (z physical) d=5 C=25 RK4 nu=2.5e-4, forced as the 256^3 workload
    "hd512": dict(nx=512, ny=512, nzp=512, C=25, d=5, ord=4, nu=2.5e-4, dt=2.5e-4, forcing=True,
                  label="HD 512x512x(512+25) FC-Gram d=5 C=25, RK4, nu=2.5e-4, dt=2.5e-4, "
                        + _FORCING),
    # HD 256^3 (z physical) d=5 C=25 RK4 nu=5e-4, forced
    "hd256": dict(nx=256, ny=256, nzp=256, C=25, d=5, ord=4, nu=5e-4, dt=5e-4, forcing=True,
                  label="HD 256x256x(256+25) FC-Gram d=5 C=25, RK4, nu=5e-4, dt=5e-4, "
                        + _FORCING),
    # HD 1024^3 (one GPU's share of a production run) d=5 C=25 RK4 nu=1e-4, forced
    "hd1024": dict(nx=1024, ny=1024, nzp=1024, C=25, d=5, ord=4, nu=1e-4, dt=1e-4, forcing=True,
                   label="HD 1024x1024x(1024+25) FC-Gram d=5 C=25, RK4, nu=1e-4, dt=1e-4, "
                         + _FORCING),
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
DEFAULT_WORKLOAD = "hd512"


def _log(msg):
    print(f"[bench] {time.strftime('%H:%M:%S')} {msg}", file=sys.stderr, flush=True)


def _plan(w, precision):
    from specter_amd import FCPlan
    return FCPlan(w["nx"], w["ny"], w["nzp"] + w["C"], w["C"], w["d"], precision=precision)


def _params(w):
    from specter_amd import SolverParams
    if w.get("solver") == "BOUSS":
        return SolverParams(solver="BOUSS", ord=w["ord"], dt=w["dt"], nu=w["nu"],
                            kappa=w["kappa"], xmom=w["xmom"], xtemp=w["xtemp"])
    return SolverParams(solver="HD", ord=w["ord"], dt=w["dt"], nu=w["nu"])


def _initial_fields(w):
    """Real-space initial fields of the workload: {name: array or cuda tensor}.
    HD: the seeded spectrum-peaked velocity (and the forcing) generated on the
    GPU; Boussinesq: v = 0 and a seeded white-noise theta'."""
    from specter_amd import ic
    nx, ny, nz, C = w["nx"], w["ny"], w["nzp"] + w["C"], w["C"]
    if w.get("solver") == "BOUSS":
        return {"th": ic.thermal_perturbation(nx, ny, nz, C, 1001)}
    out = dict(zip(("vx", "vy", "vz"), ic.spectrum_peaked_velocity_device(nx, ny, nz, C, 1000)))
    if w["forcing"]:
        out.update(zip(("fx", "fy", "fz"), ic.large_scale_forcing_device(nx, ny, nz, C)))
    return out


def _stages(sol, ordr, nsub, start=0):
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
    _stages(sol, ordr, nsub, start)
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


def _time(sol, w, steps, warmup):
    """ms per RK stage over `steps` stages after the warmup rounded up to whole
    RK steps (HIP events on the solver's stream, which is torch's current
    stream).  Returns (ms, wall s, stage position, warmup stages run)."""
    import torch
    ordr = w["ord"]
    wst = -(-warmup // ordr) * ordr
    pos = _stages(sol, ordr, wst)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    pos = _stages(sol, ordr, steps, pos)
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps, time.perf_counter() - t0, pos, wst


def _kernels(plan, sol, w, steps, pos):
    rows = _profile(plan, sol, w["ord"], steps, pos)
    rows.sort(key=lambda r: -r[2])
    table = {r[0]: {"launches_per_step": round(r[1] / steps, 2),
                    "ms_per_step": round(r[2] / steps, 4),
                    "GBps": round(r[3] / (r[2] * 1e-3) / 1e9, 1)} for r in rows}
    return rows, table


def _measured_traffic(workload, mode):
    """HBM bytes per launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of the workload (read + write MiB columns, calibrated
    on a plain copy); {} when no summary is committed for it."""
    path = os.path.join(ROOT, "profiles", f"{workload}_hbm_traffic.md")
    if not os.path.exists(path):
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


def _diagnostics(w, plan, sol, bouss, out_names):
    import numpy as np
    vx, vy, vz = sol.field("vx"), sol.field("vy"), sol.field("vz")
    d = {"spectrum": np.array(plan.energy_spectrum(vx, vy, vz, w["nx"] // 2)),
         "dissipation": w["nu"] * plan.energy(vx, vy, vz, 0),
         "wall_shear": np.array(plan.wall_shear(vx, vy, w["nu"])),
         "fields": {n: sol.spectral(n) for n in out_names}}
    if bouss:
        d["nusselt"] = np.array(plan.nusselt(sol.field("th")))
    return d


def _errors(a, b, out_names, bouss, steps):
    import numpy as np

    def rel(x, y, norm):
        d = norm(y)
        return float(f"{norm(x - y) / d:.3e}") if d > 0 else None

    l2 = np.linalg.norm
    amax = lambda v: float(np.abs(v).max())
    err = {"steps": steps}
    for n in out_names:
        err[n] = rel(a["fields"][n], b["fields"][n], l2)
    err["energy_spectrum"] = rel(a["spectrum"], b["spectrum"], l2)
    err["dissipation"] = rel(a["dissipation"], b["dissipation"], abs)
    err["wall_shear"] = rel(a["wall_shear"], b["wall_shear"], amax)
    if bouss:
        err["nusselt"] = rel(a["nusselt"], b["nusselt"], amax)
    return err


def _cpu_baseline(w, spec, nsub):
    """The float64 oracle (numpy/scipy restatement) on the host cores, on a
    bounded sample of the same workload: nsub RK stages from the spectral
    initial state the GPU runs start from."""
    import numpy as np
    workers = int(os.environ.get("OMP_NUM_THREADS", str(os.cpu_count() or 1)))
    os.environ["SPX_ORACLE_WORKERS"] = str(workers)
    from oracle import specter_ref as O
    nx, ny, nz, C, d = w["nx"], w["ny"], w["nzp"] + w["C"], w["C"], w["d"]
    g = O.make_grid(nx, ny, nz, C, d, os.path.join(ROOT, "specter_amd", "tables"))
    zero = lambda: np.zeros(g.spec_shape, complex)
    get = lambda n: spec[n].copy() if n in spec else zero()
    bouss = w.get("solver") == "BOUSS"
    st = O.HDState(get("vx"), get("vy"), get("vz"), zero(), get("fx"), get("fy"), get("fz"),
                   get("th") if bouss else None, zero() if bouss else None)
    C1, C2, C3 = st.vx.copy(), st.vy.copy(), st.vz.copy()
    C7 = st.th.copy() if bouss else None
    t0 = time.perf_counter()
    for s in range(nsub):
        o = w["ord"] - (s % w["ord"])
        if bouss:
            O.bouss_substep(g, st, C1, C2, C3, C7, o, w["ord"], w["dt"], w["nu"], w["kappa"],
                            w["xmom"], w["xtemp"])
        else:
            O.hd_substep(g, st, C1, C2, C3, o, w["ord"], w["dt"], w["nu"])
    dt = (time.perf_counter() - t0) / nsub
    return dt, workers


def _ensemble(nruns, max_slots, chunk):
    """The ensemble workload (Poisson(4/s) arrivals of HD RK2 runs, 300 steps,
    grids {32,64,128,192}^3, d in {5,8}, nu log-uniform [5e-4,5e-3]) through
    the continuous-batching server (precision per class from its error
    budget, specter_amd.ensemble.default_precision), against the same stream
    served one run at a time in float64:
    every request is run alone in float64 (plans cached per grid), its
    service time measured, and FIFO latencies follow from the arrival times.
    The first run of each server class keeps its fields; its relative L2
    error against its float64 run is reported per class."""
    import numpy as np
    import torch
    from specter_amd.ensemble import EnsembleServer, latency_stats, poisson_stream
    from specter_amd import FCPlan, Solver, SolverParams, ic
    from specter_amd.lib import SPX_FP64
    reqs = poisson_stream(nruns=nruns)
    work = sum(r.points * r.nsteps * r.ord for r in reqs)
    srv = EnsembleServer(max_slots=max_slots, chunk=chunk, sample_per_class=True)
    t0 = time.perf_counter()
    res = srv.run(reqs, realtime=True)
    wall = time.perf_counter() - t0
    ok = [r for r in res if r.status == "ok"]
    st = latency_stats(res)
    by_rid = {r.rid: r for r in res}
    sampled = {v: k for k, v in srv.sampled.items()}      # rid -> class key
    # sequential float64 service, measured run by run
    _log("ensemble: sequential float64 baseline")
    cache, service, ref_fields = {}, {}, {}
    for r in sorted(reqs, key=lambda q: q.arrival):
        nx, ny, nz, C, d = r.grid
        if r.grid not in cache:
            plan = FCPlan(nx, ny, nz, C, d, precision=SPX_FP64)
            cache[r.grid] = (plan, Solver(plan, SolverParams(solver="HD", ord=r.ord, dt=r.dt,
                                                             nu=r.nu)))
        plan, sol = cache[r.grid]
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for name in ("pr", "fx", "fy", "fz"):
            sol.field(name).zero_()
        for name, arr in zip(("vx", "vy", "vz"),
                             ic.spectrum_peaked_velocity_device(nx, ny, nz, C, r.seed)):
            sol.set_real_field(name, arr)
        sol.set_member(0, r.nu)
        sol.step(r.nsteps)
        torch.cuda.synchronize()
        service[r.rid] = time.perf_counter() - t1
        if r.rid in sampled:
            ref_fields[r.rid] = {n: sol.spectral(n) for n in ("vx", "vy", "vz")}
    for plan, sol in cache.values():
        sol.close()
        plan.close()
    clock, lat = 0.0, []
    for r in sorted(reqs, key=lambda q: q.arrival):
        clock = max(clock, r.arrival) + service[r.rid]
        lat.append(clock - r.arrival)
    lat = np.array(lat)
    # per-class error of the sampled run against its float64 run
    classes = []
    for rid, key in sorted(sampled.items(), key=lambda kv: kv[1]):
        got = by_rid.get(rid)
        if got is None or got.status != "ok" or rid not in ref_fields:
            continue
        err = max(float(np.linalg.norm((got.fields[n] - ref_fields[rid][n]).ravel())
                        / np.linalg.norm(ref_fields[rid][n].ravel())) for n in ("vx", "vy", "vz"))
        nx, ny, nz, C, d = key[:5]
        classes.append({"grid": f"{nx}x{ny}x({nz - C}+{C})", "d": d,
                        "precision": "fp64" if key[-1] == SPX_FP64 else "mixed",
                        "sampled_run": rid, "field_rel_l2_vs_fp64": float(f"{err:.3e}")})
    seq_busy = sum(service.values())
    return {
        "runs": nruns, "workload": "Poisson 4/s, 300 RK2 steps, {32,64,128,192}^3, d in {5,8}, "
                                   "nu log-U[5e-4,5e-3]; precision per class for a 1e-5 error "
                                   "budget (FP32-mixed for d=5 at 128^3 and 192^3, FP64 "
                                   "otherwise); continuous batching",
        "completed": len(ok), "failed": len(res) - len(ok),
        "p50_s": round(st["p50_s"], 4), "p99_s": round(st["p99_s"], 4),
        "makespan_s": round(wall, 3),
        "aggregate_Mpts_per_s": round(work / wall / 1e6, 2),
        "peak_classes": srv.peak_classes, "peak_hbm_GB": round(srv.peak_used / 1e9, 2),
        "class_errors": classes,
        "sequential_fp64": {"p50_s": round(float(np.percentile(lat, 50)), 4),
                            "p99_s": round(float(np.percentile(lat, 99)), 4),
                            "makespan_s": round(clock, 3),
                            "aggregate_Mpts_per_s": round(work / clock / 1e6, 2),
                            "busy_s": round(seq_busy, 3),
                            "how": "every request run alone in FP64 on the GPU, service times "
                                   "measured; FIFO latencies from the stream's arrival times"},
        "throughput_vs_sequential_fp64": round(clock / wall, 3),
    }


def _bench(args, w):
    """Both precisions on the same workload, one after the other from the same
    initial state (its spectral fields are kept on the host, so the largest
    grids never hold two plans at once): --err-steps full steps for the error
    comparison, then the timed stages and the per-kernel pass."""
    import torch
    from specter_amd import Solver
    from specter_amd.lib import SPX_FP32_MIXED, SPX_FP64, SpxError
    nx, ny, nz = w["nx"], w["ny"], w["nzp"] + w["C"]
    npts = nx * ny * w["nzp"]
    ntot = nx * ny * nz
    bouss = w.get("solver") == "BOUSS"
    out_names = ("vx", "vy", "vz", "th") if bouss else ("vx", "vy", "vz")
    modes = ["mixed", "fp64"] if args.precision == "both" else [args.precision]
    _log(f"{args.workload}: initial fields")
    # generated on the GPU, parked on the host before the first plan exists
    # (six float64 fields are 53 GB at 1024^3, beside a 230 GB plan + state)
    real_ic = {k: (v.cpu().numpy() if hasattr(v, "cpu") else v)
               for k, v in _initial_fields(w).items()}
    torch.cuda.empty_cache()
    res, host_ic, diag, zinfo = {}, None, {}, None
    for mode in modes:
        prec = SPX_FP32_MIXED if mode == "mixed" else SPX_FP64
        try:
            plan = _plan(w, prec)
            sol = Solver(plan, _params(w))
        except SpxError as e:
            res[mode] = {"error": str(e)}
            _log(f"{mode}: {e}")
            continue
        if zinfo is None:
            zinfo = {0: "direct", 1: "bluestein", 2: "rader"}[plan.z_algorithm] + \
                    f" (in-LDS FFT length {plan.z_fft_len})"
        if host_ic is None:
            for name in list(real_ic):
                sol.set_real_field(name, real_ic.pop(name))
                torch.cuda.empty_cache()
            host_ic = {n: sol.spectral(n) for n in sol.fields if n != "pr"}
        else:
            for n, a in host_ic.items():
                sol.field(n).copy_(torch.from_numpy(a))
            plan.sync()
        if args.err_steps > 0 and len(modes) == 2:
            _log(f"{mode}: {args.err_steps} steps")
            sol.step(args.err_steps)
            plan.sync()
            diag[mode] = _diagnostics(w, plan, sol, bouss, out_names)
        _log(f"{mode}: timing")
        ms, wall, pos, wst = _time(sol, w, args.steps, args.warmup)
        r = {"ms_per_step": ms, "wall_s": wall, "value": npts / (ms * 1e-3) / 1e6,
             "warmup_stages": wst}
        if not args.no_profile:
            r["rows"], r["kernels"] = _kernels(plan, sol, w, args.steps, pos)
        res[mode] = r
        sol.close()
        plan.close()
        del sol, plan
        torch.cuda.empty_cache()
    head = modes[0]
    hr = res[head]
    if "value" not in hr:
        print(json.dumps({"metric": "RK substep Mpts/s", "value": None, "unit": "Mpts/s",
                          "config": {"workload": w["label"]}, "error": res}), flush=True)
        return
    line = {
        "metric": "RK substep Mpts/s (physical points per RK stage per second), "
                  + ("FP32-mixed" if head == "mixed" else "FP64"),
        "value": round(hr["value"], 3), "unit": "Mpts/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "warmup_stages": hr["warmup_stages"],
        "ms_per_step": round(hr["ms_per_step"], 4), "wall_s": round(hr["wall_s"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 nonlinear term / f64 state, continuation, pressure, walls"
                 if head == "mixed" else "f64",
        "precision_mode": head,
        "data": "synthetic (seeded white-noise theta', v = 0)" if bouss else
                "synthetic (seeded spectrum-peaked solenoidal IC generated on the GPU "
                "x sin(pi z/Lz))",
        "config": {"workload": w["label"], "grid": [nx, ny, nz], "z_dft": zinfo,
                   "physical_points": npts, "total_points": ntot,
                   "total_Mpts_per_s": round(ntot / (hr["ms_per_step"] * 1e-3) / 1e6, 3),
                   "parallelism": "single GPU"},
    }
    if not args.no_profile:
        line["roofline"] = _roofline(hr["rows"], _measured_traffic(args.workload, head))
        line["kernels"] = hr["kernels"]
    f64 = res.get("fp64", {})
    if head == "mixed" and "ms_per_step" in f64:
        line["fp64"] = {"value": round(f64["value"], 3),
                        "ms_per_step": round(f64["ms_per_step"], 4)}
        if not args.no_profile:
            line["fp64"]["roofline"] = _roofline(f64["rows"],
                                                 _measured_traffic(args.workload, "fp64"))
            line["fp64"]["kernels"] = f64["kernels"]
        line["speedup_vs_fp64"] = round(f64["ms_per_step"] / hr["ms_per_step"], 3)
        if diag:
            line["error_vs_fp64"] = _errors(diag["mixed"], diag["fp64"], out_names, bouss,
                                            args.err_steps)
            if bouss:
                line["nusselt_fp64"] = [round(float(x), 10) for x in diag["fp64"]["nusselt"]]
    elif head == "mixed" and f64:
        line["fp64"] = f64
    if args.ensemble_runs > 0:
        _log("ensemble")
        line["ensemble"] = _ensemble(args.ensemble_runs, args.ensemble_slots,
                                     args.ensemble_chunk)
    if args.cpu_substeps > 0 and host_ic is not None:
        _log("cpu baseline")
        t, cores = _cpu_baseline(w, host_ic, args.cpu_substeps)
        line["cpu_baseline"] = {"value": round(npts / t / 1e6, 4), "unit": "Mpts/s",
                                "cores": cores, "kind": "port",
                                "sample": f"{args.cpu_substeps} RK stage(s) of the same "
                                          f"workload from the same initial state, float64 "
                                          f"oracle (scipy.fft workers={cores}, numpy "
                                          f"elementwise single-threaded), {t:.2f} s per stage"}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8, help="timed RK stages")
    ap.add_argument("--warmup", type=int, default=4,
                    help="untimed RK stages (rounded up to whole RK steps)")
    ap.add_argument("--workload", default=DEFAULT_WORKLOAD, choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="both", choices=("both", "mixed", "fp64"))
    ap.add_argument("--err-steps", type=int, default=100,
                    help="full RK steps of the mixed-vs-float64 field comparison (0 = skip)")
    ap.add_argument("--cpu-substeps", type=int, default=1,
                    help="RK stages of the CPU baseline sample (0 = skip)")
    ap.add_argument("--no-profile", action="store_true", help="skip the per-kernel pass")
    ap.add_argument("--ensemble-runs", type=int, default=None,
                    help="requests of the ensemble-serving measurement (0 = skip; default "
                         "96 for the default workload, 0 otherwise)")
    ap.add_argument("--ensemble-slots", type=int, default=16)
    ap.add_argument("--ensemble-chunk", type=int, default=5, help="RK steps of a class's first chunk")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); the "
              "multi-GPU slab decomposition is a design, not a measured path", file=sys.stderr)
        sys.exit(2)
    if args.ensemble_runs is None:
        args.ensemble_runs = 96 if args.workload == DEFAULT_WORKLOAD else 0
    _bench(args, WORKLOADS[args.workload])


if __name__ == "__main__":
    main()
