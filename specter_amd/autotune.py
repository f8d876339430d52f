"""Autotuner for the launch shapes of libspx's z-line kernels and wall kernel.

For one grid (and precision and member count, i.e. one ensemble class) it
searches, coordinate by coordinate,

* the float64 z lines (state transforms, projection, harmonic correction,
  boundary planes) and, in mixed plans, the float32 z lines of the nonlinear
  term: threads per line, lines per workgroup, and whether the Rader /
  Bluestein tables sit in global memory or LDS;
* the wall kernel on the matrix cores: kappa blocks per wave and waves per
  SIMD of its register budget,

timing each candidate as whole RK steps of a batched solver on the GPU (HIP
events on the solver's stream) and keeping the fastest.  The choices are
persisted as JSON and applied (spx_tune_line / spx_tune_wall) before plans
are created; the overrides are keyed by line length, so they apply to every
plan whose lines have that length.

    python -m specter_amd.autotune --out specter_amd/tuning.json \\
        --grid 64,5,mixed,16 --grid 192,8,fp64,1 ...
"""
from __future__ import annotations

import argparse
import json
import os
import time

from . import lib as _l

DEFAULT_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")
FC_TABLE = {5: 25, 8: 27}


def apply_tuning(path: str = DEFAULT_PATH) -> int:
    """Register the persisted choices with libspx; returns how many entries
    were applied (0 when the file does not exist)."""
    if not path or not os.path.exists(path):
        return 0
    lib = _l.load_library()
    with open(path) as fh:
        data = json.load(fh)
    n = 0
    for e in data.get("line", []):
        _l.check(lib.spx_tune_line(int(e["n"]), int(e["csize"]), int(e["lines"]),
                                   int(e["threads"]), int(e["tables_global"])))
        n += 1
    for e in data.get("wall", []):
        _l.check(lib.spx_tune_wall(int(e["n"]), int(e["blocks_per_wave"]),
                                   int(e["waves_per_simd"])))
        n += 1
    return n


def _time_steps(grid, precision, members, nsteps=4, warmup=2, seed=7):
    """ms per RK2 step of a batched HD solver of the grid (plans are created
    with the overrides registered at call time)."""
    import torch
    from . import FCPlan, Solver, SolverParams, ic
    nx, ny, nz, C, d = grid
    plan = FCPlan(nx, ny, nz, C, d, precision=precision, members=members)
    sol = Solver(plan, SolverParams(solver="HD", ord=2, dt=1e-3, nu=2e-3))
    for b in range(members):
        for name, arr in zip(("vx", "vy", "vz"),
                             ic.spectrum_peaked_velocity_device(nx, ny, nz, C, seed + b)):
            sol.set_real_field(name, arr, member=b if members > 1 else None)
    sol.step(warmup)
    plan.sync()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    sol.step(nsteps)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / nsteps
    info = plan.info()
    shape = (info.z_threads_per_line, info.z_lines_per_block, info.z_tables_global)
    sol.close()
    plan.close()
    return ms, shape


def tune_grid(n: int, d: int, precision: int, members: int, log=print) -> dict:
    """Coordinate search for one class (n^3 physical points, FC order d).
    Returns {"line": [...], "wall": [...], "ms_default": .., "ms_tuned": ..}."""
    lib = _l.load_library()
    C = FC_TABLE[d]
    grid = (n, n, n + C, C, d)
    nz = n + C
    state = {16: [0, 0, -1], 8: [0, 0, -1]}   # lines, threads, tables_global per csize
    wall = [0, 0]

    def register():
        _l.check(lib.spx_tune_clear())
        for cs, (ln, th, gt) in state.items():
            _l.check(lib.spx_tune_line(nz, cs, ln, th, gt))
        _l.check(lib.spx_tune_wall(nz, wall[0], wall[1]))

    register()
    best, (t0, l0, g0) = _time_steps(grid, precision, members)
    ms_default = best
    log(f"  {grid} precision {precision} x{members}: default {best:.3f} ms/step "
        f"(z: {t0} threads/line, {l0} lines/block, global tables {g0})")
    csizes = [16, 8] if precision == _l.SPX_FP32_MIXED else [16]
    for cs in csizes:
        for field, values in ((1, (t0, 2 * t0, 4 * t0)), (0, (l0, max(1, l0 // 2), 2 * l0)),
                              (2, (0, 1))):
            for v in values:
                trial = list(state[cs])
                trial[field] = v
                if trial == state[cs]:
                    continue
                keep = state[cs]
                state[cs] = trial
                register()
                try:
                    ms, _ = _time_steps(grid, precision, members)
                except _l.SpxError as e:
                    log(f"    csize {cs} {trial}: {e}")
                    ms = float("inf")
                log(f"    csize {cs} lines/threads/gtab {trial}: {ms:.3f} ms/step")
                if ms < best * 0.99:
                    best = ms
                else:
                    state[cs] = keep
    for kbw, wpe in ((9, 3), (18, 2), (9, 4)):
        keep = list(wall)
        wall[:] = [kbw, wpe]
        register()
        ms, _ = _time_steps(grid, precision, members)
        log(f"    wall blocks/wave {kbw}, waves/SIMD {wpe}: {ms:.3f} ms/step")
        if ms < best * 0.99:
            best = ms
        else:
            wall[:] = keep
    _l.check(lib.spx_tune_clear())
    out = {"line": [], "wall": [], "grid": list(grid), "precision": precision,
           "members": members, "ms_default": round(ms_default, 4), "ms_tuned": round(best, 4)}
    for cs, (ln, th, gt) in state.items():
        if (ln, th, gt) != (0, 0, -1):
            out["line"].append({"n": nz, "csize": cs, "lines": ln, "threads": th,
                                "tables_global": gt})
    if wall != [0, 0]:
        out["wall"].append({"n": nz, "blocks_per_wave": wall[0], "waves_per_simd": wall[1]})
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--out", default=DEFAULT_PATH)
    ap.add_argument("--grid", action="append", required=True,
                    help="n,d,precision(mixed|fp64),members  e.g. 64,5,mixed,16")
    args = ap.parse_args()
    result = {"line": [], "wall": [], "classes": []}
    t0 = time.perf_counter()
    for g in args.grid:
        n, d, prec, members = g.split(",")
        precision = _l.SPX_FP32_MIXED if prec == "mixed" else _l.SPX_FP64
        print(f"tuning n={n} d={d} {prec} x{members}", flush=True)
        r = tune_grid(int(n), int(d), precision, int(members),
                      log=lambda s: print(s, flush=True))
        result["line"] += r["line"]
        result["wall"] += r["wall"]
        result["classes"].append({k: r[k] for k in ("grid", "precision", "members",
                                                    "ms_default", "ms_tuned")})
    # one entry per (n, csize) / n: the later class wins
    result["line"] = list({(e["n"], e["csize"]): e for e in result["line"]}.values())
    result["wall"] = list({e["n"]: e for e in result["wall"]}.values())
    result["seconds"] = round(time.perf_counter() - t0, 1)
    with open(args.out, "w") as fh:
        json.dump(result, fh, indent=1)
    print(f"wrote {args.out}", flush=True)


if __name__ == "__main__":
    main()
