"""Ensemble server: many independent HD runs on one GPU with continuous
batching.

Requests (grid, FC order, viscosity, seed, step count) arrive over time.
Runs of the same grid signature share one batched libspx plan ("class"); a
class owns a fixed number of member slots, its own HIP stream and one
captured HIP graph per active-member count.  The loop

1. admits every request that has arrived into a free slot of its class
   (creating the class when the HBM budget allows; otherwise the request
   waits until memory is released); a busy class takes new members at the
   end of its current chunk,
2. keeps every class stepping independently on its own stream: as soon as a
   class's chunk of RK steps has finished (polled with an event, the host
   never blocks on a busy stream), its finished runs are handed back (final
   fields and diagnostics), the class is compacted so the active members
   stay leading, waiting requests join, and the next chunk is enqueued.
   Chunks are sized to about `chunk_seconds` of GPU time from the class's
   measured rate, so small grids cycle fast and big ones do not hold them up.
   Classes without runs stay (with their captured graphs) until their memory
   is needed for another class.

A member of a batched plan is computed by exactly the per-line / per-row
code of a single run, so a result does not depend on what it was batched
with.  Time to result is measured from a request's arrival to the moment its
fields are on the host.

Policies:
* precision: a request without one gets the cheapest precision that meets a
  1e-5 field error budget against float64 (default_precision: float32-mixed
  for d = 5 on grids of 128^3 and more, float64 otherwise; the d = 8
  continuation multiplies float32 noise by |A Q^T| = 2.2e6);
* slots: a class gets at most max(1, fill_points / points per run) members
  (a 192^3 run already fills the GPU; batching more of them only slows each
  one down), and waiting requests join in arrival order;
* failures: after every chunk the active members' energies are checked; a
  member whose fields are no longer finite is returned with status
  "failed" and removed, without touching the other members.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import lib as _l
from .fftp import FCPlan
from .solver import Solver, SolverParams

FC_TABLE = {5: 25, 8: 27}      # matching order d -> continuation points C


@dataclass
class RunRequest:
    rid: int
    n: int                     # nx = ny = n physical points per direction
    d: int                     # FC matching order (5 or 8)
    nu: float
    seed: int
    nsteps: int
    arrival: float = 0.0       # seconds after the start of the stream
    ord: int = 2
    dt: float = 1e-3
    precision: int | None = None   # None: chosen by the server's policy

    @property
    def grid(self):
        C = FC_TABLE[self.d]
        return (self.n, self.n, self.n + C, C, self.d)

    @property
    def key(self):
        return (*self.grid, self.ord, self.dt, self.precision)

    @property
    def points(self) -> int:
        """physical points of one RK stage"""
        return self.n ** 3


@dataclass
class RunResult:
    rid: int
    arrival: float
    admitted: float
    finished: float
    fields: dict = field(default_factory=dict)     # spectral vx, vy, vz (float64)
    diagnostics: tuple = ()                          # energy, enstrophy, <div^2>
    status: str = "ok"                               # "ok" or "failed"
    error: str = ""
    precision: int = _l.SPX_FP32_MIXED
    steps_done: int = 0

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def state_bytes(plan: FCPlan, members: int | None = None) -> int:
    """Device bytes of a batched plan with its HD solver state: what the plan
    holds, the work arrays it allocates on first use and the state's fields
    (libspx counts each allocation; ``members`` must be the plan's own)."""
    info = plan.info()
    if members is not None and members != plan.members:
        raise ValueError("state_bytes: members must match the plan")
    return int(info.device_bytes + info.work_bytes + info.state_bytes)


def default_precision(req: RunRequest) -> int:
    """The cheapest precision that meets a 1e-5 field error budget against
    float64 after the ensemble's 300 RK2 steps, from measured errors of the
    mixed mode: d = 5 on 128^3 and 192^3 grids 8e-7 and 1.6e-6 (mixed);
    d = 5 on 32^3 and 64^3 grids 3.8e-5 and 1.7e-5, d = 8 5e-4 after only
    10 steps at 32^3 (float64)."""
    return _l.SPX_FP32_MIXED if (req.d <= 5 and req.n >= 128) else _l.SPX_FP64


def member_bytes(req: RunRequest) -> int:
    """Estimated device bytes per member of the request's class (used before
    the class exists)."""
    nx, ny, nz, _, _ = req.grid
    nkx = nx // 2 + 1
    nc, nr = nz * ny * nkx, nx * ny * nz
    c = 8 if req.precision == _l.SPX_FP32_MIXED else 16
    r = c // 2
    return 13 * nc * 16 + 12 * nc * c + 6 * nr * r + nr * 8 + 2 * nc * 16


class _Class:
    def __init__(self, req: RunRequest, slots: int):
        import torch
        self.key = req.key
        self.stream = torch.cuda.Stream()
        nx, ny, nz, C, d = req.grid
        with torch.cuda.stream(self.stream):
            self.plan = FCPlan(nx, ny, nz, C, d, precision=req.precision, members=slots,
                               stream=self.stream.cuda_stream)
            self.sol = Solver(self.plan, SolverParams(solver="HD", ord=req.ord, dt=req.dt,
                                                      nu=req.nu))
        self.slots = slots
        self.runs: list[RunRequest] = []          # active members, slot order
        self.done: list[int] = []                  # steps done per active member
        self.admitted: list[float] = []
        # the member plan for per-run I/O is part of the class's footprint
        self.bytes = state_bytes(self.plan) + int(self.sol._member_plan().info().device_bytes)
        self.stepping = 0          # steps of the chunk in flight (0: idle)
        self.stepping_members = 0  # members the chunk in flight advances
        self.event = torch.cuda.Event()
        self.t_launch = 0.0
        self.rate = None           # measured RK steps per second of this class
        self.sample_rids: set = set()   # runs whose final fields are returned

    @property
    def free(self) -> int:
        return self.slots - len(self.runs)

    def admit(self, req: RunRequest, now: float):
        import torch
        from . import ic
        b = len(self.runs)
        nx, ny, nz, C, _ = req.grid
        with torch.cuda.stream(self.stream):
            # generated on the class's stream: the host only enqueues work
            u = ic.spectrum_peaked_velocity_device(nx, ny, nz, C, req.seed)
            for name in ("pr", "fx", "fy", "fz"):
                self.sol.member_field(name, b).zero_()
            for name, arr in zip(("vx", "vy", "vz"), u):
                self.sol.set_real_field(name, arr, member=b)
            self.sol.set_member(b, req.nu)
        self.runs.append(req)
        self.done.append(0)
        self.admitted.append(now)
        self.plan.set_active(len(self.runs))

    def launch(self, chunk: int, now: float) -> int:
        """Enqueue min(chunk, steps left of any member) steps; returns it."""
        k = min(chunk, min(r.nsteps - s for r, s in zip(self.runs, self.done)))
        self.sol.step(k)
        self.event.record(self.stream)
        self.stepping = k
        self.stepping_members = len(self.runs)
        self.t_launch = now
        return k

    def ready(self) -> bool:
        return self.stepping > 0 and self.event.query()

    def finish_chunk(self, now_fn, keep_fields: bool):
        """Return the finished runs' results after a completed chunk and compact."""
        self.stream.synchronize()
        k, self.stepping = self.stepping, 0
        el = now_fn() - self.t_launch
        if el > 0:
            r = k / el
            self.rate = r if self.rate is None else 0.5 * (self.rate + r)
        m = self.stepping_members
        self.done = [s + k if b < m else s for b, s in enumerate(self.done)]
        out = []
        vx, vy, vz = (self.sol.field(n) for n in ("vx", "vy", "vz"))
        p = self.plan
        # per-member health: a run whose fields stopped being finite fails alone
        eng = np.atleast_1d(p.energy(vx, vy, vz, 1))
        b = 0
        while b < len(self.runs):
            r = self.runs[b]
            bad = not np.isfinite(eng[b])
            if self.done[b] < r.nsteps and not bad:
                b += 1
                continue
            res = RunResult(r.rid, r.arrival, self.admitted[b], 0.0, precision=r.precision,
                            steps_done=self.done[b])
            if bad:
                res.status = "failed"
                res.error = f"non-finite fields after {self.done[b]} steps"
            else:
                if keep_fields or r.rid in self.sample_rids:
                    res.fields = {n: self.sol.spectral(n, member=b) for n in ("vx", "vy", "vz")}
                res.diagnostics = (float(eng[b]),
                                   float(np.atleast_1d(p.energy(vx, vy, vz, 0))[b]),
                                   float(np.atleast_1d(p.divergence(vx, vy, vz))[b]))
            res.finished = now_fn()
            out.append(res)
            last = len(self.runs) - 1
            if b != last:
                eng[b] = eng[last]
                self.sol.copy_member(b, last)
                self.runs[b], self.done[b], self.admitted[b] = (self.runs[last], self.done[last],
                                                                self.admitted[last])
            eng = eng[:last]
            self.runs.pop()
            self.done.pop()
            self.admitted.pop()
            if self.runs:
                self.plan.set_active(len(self.runs))
        return out

    def close(self):
        self.stream.synchronize()
        self.sol.close()
        self.plan.close()


class EnsembleServer:
    """Continuous-batching server for HD runs (see the module docstring).

    hbm_budget: device bytes the server may hold; max_slots: members per
    class; chunk: RK steps enqueued per class between scheduling points."""

    def __init__(self, hbm_budget: int | None = None, max_slots: int = 16, chunk: int = 5,
                 keep_fields: bool = False, chunk_seconds: float = 0.05, max_chunk: int = 100,
                 fill_points: int = 1 << 23, precision_policy=default_precision,
                 sample_per_class: bool = False, tuning: str | None = "default"):
        import torch
        from . import autotune
        # launch shapes chosen by the autotuner (specter_amd/tuning.json when
        # present), registered before any class creates its plan
        self.tuned = autotune.apply_tuning(autotune.DEFAULT_PATH if tuning == "default"
                                           else tuning)
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info()
            hbm_budget = int(free * 0.85)
        self.budget = int(hbm_budget)
        self.max_slots = max_slots
        self.chunk = chunk                  # steps of a class's first chunk
        self.chunk_seconds = chunk_seconds
        self.max_chunk = max_chunk
        self.keep_fields = keep_fields
        self.fill_points = fill_points
        self.precision_policy = precision_policy
        self.sample_per_class = sample_per_class   # keep the first run's fields of each class
        self.sampled: dict[tuple, int] = {}
        self._class_bytes: dict[tuple, int] = {}   # measured footprint per class key
        self.classes: dict[tuple, _Class] = {}
        self.used = 0
        self.peak_used = 0
        self.peak_classes = 0

    def _slots_for(self, req: RunRequest) -> int:
        """Member slots of a new class: max_slots, at most the runs that fill
        the GPU (fill_points / points per run, at least 1), and what the budget
        has room for."""
        per = member_bytes(req)
        room = (self.budget - self.used) // max(per, 1)
        cap = max(1, self.fill_points // max(req.points, 1))
        return int(max(0, min(self.max_slots, cap, room)))

    def _evict_idle(self, keep=None):
        """Destroy classes without runs (their graphs and memory go back)."""
        for key in [k for k, c in self.classes.items() if not c.runs and k != keep]:
            c = self.classes.pop(key)
            self.used -= c.bytes
            c.close()

    def _try_admit(self, req: RunRequest, now: float, queued_same: int) -> bool:
        cls = self.classes.get(req.key)
        if cls is None:
            slots = self._slots_for(req)
            if slots < min(queued_same, self.max_slots):
                self._evict_idle()
                slots = self._slots_for(req)
            if slots < 1:
                return False
            need = self._class_bytes.get((req.key, slots))
            if need is not None and self.used + need > self.budget:
                self._evict_idle()
                if self.used + need > self.budget:
                    return False
            cls = _Class(req, slots)
            self._class_bytes[(req.key, slots)] = cls.bytes
            if self.used + cls.bytes > self.budget:
                # wait for memory (or, with nothing left to free, fail the stream)
                cls.close()
                return False
            self.classes[req.key] = cls
            self.used += cls.bytes
            self.peak_used = max(self.peak_used, self.used)
            self.peak_classes = max(self.peak_classes, len(self.classes))
        if cls.free < 1:
            return False
        if self.sample_per_class and req.key not in self.sampled:
            self.sampled[req.key] = req.rid
            cls.sample_rids.add(req.rid)
        cls.admit(req, now)
        return True

    def _chunk_for(self, c: _Class) -> int:
        if c.rate is None:
            return self.chunk
        return int(max(1, min(self.max_chunk, round(c.rate * self.chunk_seconds))))

    def run(self, requests: list[RunRequest], realtime: bool = True) -> list[RunResult]:
        """Serve a stream of requests (arrival times relative to the start);
        with realtime=False every request is available at once."""
        import torch
        torch.cuda.synchronize()
        for r in requests:
            if r.precision is None:
                r.precision = self.precision_policy(r)
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        waiting: list[RunRequest] = []
        results: list[RunResult] = []
        t0 = time.perf_counter()

        def now():
            return time.perf_counter() - t0

        while pending or waiting or any(c.runs for c in self.classes.values()):
            t = now()
            while pending and (not realtime or pending[0].arrival <= t):
                waiting.append(pending.pop(0))
            progressed = False
            # chunks that have finished: results, compaction
            for c in list(self.classes.values()):
                if c.ready():
                    results.extend(c.finish_chunk(now, self.keep_fields))
                    progressed = True
            # admissions into idle classes (or new classes)
            still = []
            for r in waiting:
                c = self.classes.get(r.key)
                if c is not None and c.stepping:
                    still.append(r)
                    continue
                same = sum(1 for q in waiting if q.key == r.key)
                if self._try_admit(r, now(), same):
                    progressed = True
                else:
                    still.append(r)
            waiting = still
            # idle classes with runs start their next chunk, oldest request first
            for c in sorted(self.classes.values(),
                            key=lambda c: min((r.arrival for r in c.runs), default=0.0)):
                if c.runs and not c.stepping:
                    c.launch(self._chunk_for(c), now())
                    progressed = True
            if not progressed:
                busy = any(c.stepping for c in self.classes.values())
                if not busy and not pending and waiting:
                    raise RuntimeError("requests cannot be admitted within the HBM budget")
                if busy:
                    time.sleep(0.0002)
                elif pending and realtime:
                    time.sleep(max(0.0, min(pending[0].arrival - now(), 0.01)))
        self._evict_idle()
        return results


def poisson_stream(nruns: int = 96, rate: float = 4.0, seed: int = 2026, nsteps: int = 300,
                   sizes=(32, 64, 128, 192), orders=(5, 8), nu_range=(5e-4, 5e-3),
                   precision: int | None = None) -> list[RunRequest]:
    """The ensemble workload: Poisson arrivals at `rate` per second, grids
    uniform over `sizes`, nu log-uniform over nu_range, FC order uniform over
    `orders`, seeds 1..nruns, `nsteps` RK2 steps each."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(nruns):
        t += rng.exponential(1.0 / rate)
        n = int(rng.choice(sizes))
        d = int(rng.choice(orders))
        nu = float(math.exp(rng.uniform(math.log(nu_range[0]), math.log(nu_range[1]))))
        out.append(RunRequest(rid=i, n=n, d=d, nu=nu, seed=i + 1, nsteps=nsteps, arrival=t,
                              precision=precision))
    return out


def latency_stats(results: list[RunResult]) -> dict:
    lat = np.array([r.latency for r in results])
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "max_s": float(lat.max())}
