#!/usr/bin/env python3
"""RK-substep throughput of the SPECTER hot path on one MI355X.

A "step" is one Runge-Kutta stage (one pass of the hot path: gradre with its
12 c2r + 3 r2c, filter + Laplacian + RK update, no-slip walls, pressure
projection) over the whole grid.  Inputs are resident in HBM before the timed
region.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload hd256|hd64]
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
    # HD 64^3 d=5 C=25 RK2 nu=2e-3 dt=1e-3, unforced
    "hd64": dict(nx=64, ny=64, nzp=64, C=25, d=5, ord=2, nu=2e-3, dt=1e-3, forcing=False,
                 label="HD 64x64x(64+25) FC-Gram d=5 C=25, RK2, nu=2e-3, dt=1e-3"),
}


def _setup(w, seed=1000):
    import numpy as np
    from specter_amd import FCPlan, Solver, SolverParams, ic
    nx, ny, nz, C, d = w["nx"], w["ny"], w["nzp"] + w["C"], w["C"], w["d"]
    plan = FCPlan(nx, ny, nz, C, d)
    sol = Solver(plan, SolverParams(solver="HD", ord=w["ord"], dt=w["dt"], nu=w["nu"]))
    u = ic.spectrum_peaked_velocity(nx, ny, nz, C, seed)
    zero = np.zeros((nz, ny, nx))
    f = ic.large_scale_forcing(nx, ny, nz, C) if w["forcing"] else (zero, zero, zero)
    for name, arr in zip(("vx", "vy", "vz", "fx", "fy", "fz"), (*u, *f)):
        sol.set_real_field(name, arr)
    return plan, sol, u, f


def _run_substeps(sol, ordr, nsub, start=0):
    """nsub RK stages cycling o = ord..1, with the rkstep1 copy before o = ord."""
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


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8, help="timed RK stages")
    ap.add_argument("--warmup", type=int, default=4, help="untimed RK stages")
    ap.add_argument("--workload", default="hd256", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-substeps", type=int, default=1,
                    help="RK stages of the CPU baseline sample (0 = skip)")
    ap.add_argument("--no-profile", action="store_true", help="skip the per-kernel pass")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); the "
              "multi-GPU slab decomposition is a design, not a measured path", file=sys.stderr)
        sys.exit(2)
    import torch
    w = WORKLOADS[args.workload]
    plan, sol, u, f = _setup(w)
    npts = w["nx"] * w["ny"] * w["nzp"]
    ntot = w["nx"] * w["ny"] * (w["nzp"] + w["C"])
    pos = _run_substeps(sol, w["ord"], args.warmup)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record()
    pos = _run_substeps(sol, w["ord"], args.steps, pos)
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1) / args.steps
    value = npts / (ms * 1e-3) / 1e6

    line = {
        "metric": "RK substep Mpts/s (physical points per RK stage per second)",
        "value": round(value, 3),
        "unit": "Mpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "wall_s": round(wall, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "precision_mode": "fp64",
        "field_rel_l2_vs_fp64": 0.0,
        "data": "synthetic (seeded spectrum-peaked solenoidal IC x sin(pi z/Lz))",
        "config": {"workload": w["label"], "grid": [w["nx"], w["ny"], w["nzp"] + w["C"]],
                   "z_dft": {0: "direct", 1: "bluestein", 2: "rader"}[plan.z_algorithm]
                            + f" (in-LDS FFT length {plan.z_fft_len})",
                   "physical_points": npts, "total_points": ntot,
                   "total_Mpts_per_s": round(ntot / (ms * 1e-3) / 1e6, 3),
                   "parallelism": "single GPU"},
    }
    if not args.no_profile:
        rows = _profile(plan, sol, w["ord"], args.steps, pos)
        rows.sort(key=lambda r: -r[2])
        name, n, tot_ms, tot_b = rows[0]
        avg_ms = tot_ms / n
        ach = (tot_b / n) / (avg_ms * 1e-3) / 1e9
        line["roofline"] = {"kernel": name, "bound": "hbm", "achieved": round(ach, 1),
                            "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                            "avg_launch_ms": round(avg_ms, 5),
                            "algorithmic_bytes_per_launch": tot_b / n}
        line["kernels"] = {r[0]: {"launches_per_step": round(r[1] / args.steps, 2),
                                  "ms_per_step": round(r[2] / args.steps, 4),
                                  "GBps": round(r[3] / (r[2] * 1e-3) / 1e9, 1)} for r in rows}
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
