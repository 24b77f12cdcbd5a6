"""Multi-simulation workloads on one GPU: the uneven batch and the served
job stream, each next to the plain path that runs one grid per launch
sequence (the reference's own usage: one time_integration call per grid).

- ``run_batched``: every simulation resident in HBM, all reset together,
  then advanced together: the batched steps replay captured HIP graphs of
  16 steps.
- ``run_plain``: the same simulations one after another through
  ``run_steps`` (the drop-in single-grid loop).
- ``Server``: continuous batching. Jobs arrive over time; a job is admitted
  when its two state buffers fit the HBM budget and a slot is free, and stays
  resident; every tick advances a batch of resident jobs (shortest remaining
  work first, up to the size that fills the device) by a few batched steps;
  finished jobs leave and free their memory; newly admitted jobs join the
  next tick while the others keep stepping, and the host prepares them while
  the device runs the current tick. ``serve_plain`` serves the same stream
  first come, first served, one job at a time.

Per-cell arithmetic does not depend on what else is in a launch, and the CFL
maximum is an exact max, so a simulation's result is bit-identical alone,
batched, or served in any mix.
"""

from __future__ import annotations

import math
import time
from collections import deque
from concurrent.futures import FIRST_COMPLETED, ThreadPoolExecutor, wait
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from .batch import Batch, Simulation
from .ics import SimSpec
from .time_integration import _TUNED, autotune, run_steps

STATE_BYTES_PER_CELL = 2 * 5 * 4  # two float32 state buffers


def state_bytes(spec: SimSpec) -> int:
    return STATE_BYTES_PER_CELL * spec.cells + _lib.TF_CTRL_BYTES


def upload(specs: Sequence[SimSpec], device) -> List[torch.Tensor]:
    return [torch.from_numpy(s.initial_state()).to(device) for s in specs]


@dataclass
class BatchResult:
    seconds: float
    states: List[torch.Tensor]
    statuses: list
    waves: int = 0


def tune_batch(specs: Sequence[SimSpec], initial: Sequence[torch.Tensor], candidates=(2, 4, 6, 8), steps: int = 4,
               tile_y_candidates=(0, 8, 16, 32)) -> Dict[str, float]:
    """Times `steps` batched steps of the real batch for each tile width and
    y-chunk candidate (0: the plan's default chunks) on the device (HIP
    events around the stencil launches) and returns the fastest as
    {"tile_x", "tile_y", "ms_per_step"} plus every candidate's time.

    Whole-line classes run tile width min(tile_x, 4) (2 above 8 waves); 6
    and 8 differ only for 62-cell z-tiles. On the 256-blast workload, width 2
    with 16-row chunks ran the 3-wave class (48- and 96-cell lines) at 107 G
    cell-updates/s against 92 G with the default width 4 and chunks."""
    device = initial[0].device
    sims = [Simulation(x, s.dx, s.gamma, device=device) for s, x in zip(specs, initial)]
    results = {}
    for tx in candidates:
        for ty in tile_y_candidates:
            b = Batch(sims, tile_x=tx, tile_y=ty)
            b.reset()
            b.step(1)  # warm-up
            results[(tx, ty)] = b.step_timed(steps) / steps
            b.close()
    best = min(results, key=results.get)
    return {"tile_x": best[0], "tile_y": best[1], "ms_per_step": results[best],
            "candidates": {f"{k[0]},{k[1]}": v for k, v in results.items()}}


def run_batched(specs: Sequence[SimSpec], initial: Sequence[torch.Tensor], *, graph=True, tile_x: int = 0,
                tile_y: int = 0, keep_states: bool = True, warm: bool = False) -> BatchResult:
    """All simulations advanced together for max(spec.steps) batched steps
    (each stops at its own step count). Timed from reset to completion.

    graph: True replays captured graphs of 16 steps; "whole" runs every step
    as one captured graph launched once; False launches each step eagerly.
    warm: a first run captures the graphs, the initial states are restored,
    and the second run is the one timed (capture excluded, as a compiled
    program's first call would be)."""
    device = initial[0].device
    sims = [Simulation(x, s.dx, s.gamma, max_steps=s.steps, device=device) for s, x in zip(specs, initial)]
    batch = Batch(sims, tile_x=tile_x, tile_y=tile_y)
    steps = max(s.steps for s in specs)

    def run():
        batch.reset()
        if graph == "whole":
            batch.run_graph(steps)
        elif graph:
            batch.step_graph(steps)
        else:
            batch.step(steps)

    if warm:
        run()
        torch.cuda.synchronize()
        for sim, x in zip(sims, initial):
            sim.buf[0].copy_(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    st = batch.status()
    states = [s.state(x) for s, x in zip(sims, st)] if keep_states else []
    batch.close()
    return BatchResult(el, states, st)


def tune_plain(specs: Sequence[SimSpec], device) -> None:
    """Measures the single-grid loop tiling of every distinct grid shape (the
    plain path does this on the first run of a shape), so that timed runs of
    the plain path do not include it."""
    for s in specs:
        if tuple(s.shape) not in _TUNED:
            autotune(torch.from_numpy(s.initial_state()).to(device), s.dx, s.gamma)


def run_plain(specs: Sequence[SimSpec], initial: Sequence[torch.Tensor], keep_states: bool = True) -> BatchResult:
    """One grid at a time through the single-grid loop (tilings measured
    beforehand, outside the timed region)."""
    tune_plain(specs, initial[0].device)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = []
    for s, x in zip(specs, initial):
        state, _, _ = run_steps(x, s.dx, s.gamma, s.steps)
        if keep_states:
            out.append(state)
    torch.cuda.synchronize()
    return BatchResult(time.perf_counter() - t0, out, [])


# ------------------------------------------------------------------ serving
@dataclass
class JobRecord:
    index: int
    spec: SimSpec
    arrival: float
    admitted: float = math.nan
    finished: float = math.nan
    steps_done: int = 0
    t: float = math.nan
    nonfinite: bool = False
    state: Optional[torch.Tensor] = None

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


@dataclass
class ServeReport:
    jobs: List[JobRecord]
    makespan: float
    ticks: int = 0
    max_resident: int = 0
    peak_bytes: int = 0
    extra: Dict[str, float] = field(default_factory=dict)

    def summary(self) -> Dict[str, float]:
        lat = np.array([j.latency for j in self.jobs])
        return {
            "jobs": len(self.jobs),
            "sims_per_s": len(self.jobs) / self.makespan,
            "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)),
            "mean_latency_s": float(lat.mean()),
            "makespan_s": self.makespan,
            "cell_updates_per_s": sum(j.spec.cells * j.spec.steps for j in self.jobs) / self.makespan,
            "ticks": self.ticks,
            "max_resident": self.max_resident,
            "peak_state_bytes": self.peak_bytes,
            **self.extra,
        }


class Server:
    """Continuous-batching server of simulation jobs on one GPU.

    Every admitted job stays resident in HBM. Jobs are grouped by grid size
    (< 2^20 cells, < 2^23, larger): all jobs of a batch advance one step per
    batched step, so only jobs of similar per-step cost share a batch, and
    each group runs its own batch on its own stream, concurrently with the
    others. Each tick, every group advances by as many steps as fit in about
    ``tick_ms`` at its own step cost (estimated from its cells), so small
    jobs progress at their own pace while large ones fill the device.

    Within a group, a batch always holds the oldest resident job (so every
    job progresses), then the others shortest remaining work first (cells x
    steps left; ``policy="fcfs"``: arrival order) until it holds
    ``fill_cells`` cells, about the size at which one launch saturates HBM.
    ``steps_per_tick`` > 0 fixes the steps per tick instead.

    Admission reserves a job's memory and slot in arrival order; its upload
    then runs on a pool of ``upload_threads`` host threads, and the job joins
    its group when the upload is done."""

    GROUP_BOUNDS = (1 << 20, 1 << 23)

    def __init__(self, hbm_budget_bytes: int, max_resident: int = 1024, steps_per_tick: int = 0, device=None,
                 keep_states: bool = False, tile_x: int = 0, fill_cells: int = 1 << 24, tick_ms: float = 4.0,
                 policy: str = "shortest", prep_threads: int = 4, upload_threads: int = 2, depth: int = 2):
        if policy not in ("shortest", "fcfs"):
            raise ValueError(f"policy must be 'shortest' or 'fcfs', got {policy!r}")
        self.policy = policy
        self.budget = int(hbm_budget_bytes)
        self.max_resident = int(max_resident)
        self.steps_per_tick = int(steps_per_tick)
        self.device = device or torch.device("cuda:0")
        self.keep_states = keep_states
        self.tile_x = tile_x
        self.fill_cells = int(fill_cells)
        self.tick_ms = float(tick_ms)
        self.prep_threads = max(1, int(prep_threads))
        self.upload_threads = max(1, int(upload_threads))
        self.depth = max(1, int(depth))

    @classmethod
    def group_of(cls, spec: SimSpec) -> int:
        return sum(spec.cells >= b for b in cls.GROUP_BOUNDS)

    def _select(self, members: List[int], jobs: List[JobRecord]) -> List[int]:
        """Jobs of a group's next tick (see the class description)."""
        oldest = min(members, key=lambda i: (jobs[i].arrival, i))
        chosen, cells = [oldest], jobs[oldest].spec.cells
        if self.policy == "fcfs":
            key = lambda i: (jobs[i].arrival, i)  # noqa: E731
        else:
            key = lambda i: (jobs[i].spec.cells * (jobs[i].spec.steps - jobs[i].steps_done), i)  # noqa: E731
        for i in sorted((i for i in members if i != oldest), key=key):
            if cells >= self.fill_cells:
                break
            chosen.append(i)
            cells += jobs[i].spec.cells
        return sorted(chosen)

    def _tick_steps(self, ids: List[int], jobs: List[JobRecord], queued=()) -> int:
        """Steps of a group's next tick; 0 when the ticks already queued
        (``queued``: (plan, ids, steps, ...) tuples) cover every job's steps."""
        ahead: Dict[int, int] = {}
        for t in queued:
            for i in t[1]:
                ahead[i] = ahead.get(i, 0) + t[2]
        left = max(jobs[i].spec.steps - jobs[i].steps_done - ahead.get(i, 0) for i in ids)
        if left <= 0:
            return 0
        if self.steps_per_tick > 0:
            return self.steps_per_tick
        cells = sum(jobs[i].spec.cells for i in ids)
        step_us = 12.0 + cells * 40.0 / 4.5e6  # launch overhead + HBM traffic at ~4.5 TB/s
        return int(max(1, min(left, 1024, self.tick_ms * 1e3 / step_us)))

    def serve(self, specs: Sequence[SimSpec], time_scale: float = 1.0,
              initial: Optional[Sequence[np.ndarray]] = None) -> ServeReport:
        """Serves jobs arriving at spec.arrival * time_scale seconds after the
        start (wall clock). Returns per-job records. The tilings of the
        stream's grid shapes are measured beforehand, as serve_plain does; a
        batch of one job then runs exactly like the plain path.

        ``initial``: each job's initial state (host float32 arrays), as a
        request would carry it. Without it, a job's initial state is built
        from its spec by a pool of host threads from the moment it arrives.
        Either way, an admitted job's state is uploaded by a second pool of
        host threads on a stream of its own: the serving loop never blocks on
        a host-to-device copy, so the ticks it has queued keep the device
        busy, and a group stream waits only for its own job's upload."""
        jobs = [JobRecord(i, s, s.arrival * time_scale) for i, s in enumerate(specs)]
        if initial is not None and len(initial) != len(jobs):
            raise ValueError(f"{len(initial)} initial states for {len(jobs)} jobs")
        for j in jobs:
            if state_bytes(j.spec) > self.budget:
                raise ValueError(f"job {j.index} ({j.spec.shape}) does not fit the HBM budget")
        tune_plain(specs, self.device)
        ngroups = len(self.GROUP_BOUNDS) + 1
        streams = [torch.cuda.Stream(self.device) for _ in range(ngroups)]  # default priority
        upload = torch.cuda.Stream(self.device)
        # two pools: an upload waits for its job's host state, so uploads
        # must not occupy the threads that build those states
        with ThreadPoolExecutor(max_workers=self.prep_threads) as pool, \
                ThreadPoolExecutor(max_workers=self.upload_threads) as uploader:
            return self._serve(jobs, initial, streams, upload, pool, uploader)

    def _upload(self, spec: SimSpec, host_state, upload) -> tuple:
        """Runs on an upload thread: the job's buffers and control block,
        allocated and filled on the upload stream (the copy from pageable host
        memory blocks this thread only). Returns the simulation and an event
        after its last upload-stream work."""
        x = host_state.result()
        with torch.cuda.device(self.device), torch.cuda.stream(upload):
            sim = Simulation(x, spec.dx, spec.gamma, max_steps=spec.steps, device=self.device)
            ev = torch.cuda.Event()
            ev.record(upload)
        return sim, ev

    def _serve(self, jobs: List[JobRecord], initial, streams, upload, pool, uploader) -> ServeReport:
        ngroups = len(streams)
        prepared = {}  # job index -> future of its initial state (host array)
        uploading = {}  # job index -> future of (Simulation, event) on the upload pool
        pending = list(jobs)  # in arrival order
        queue: List[JobRecord] = []
        active: Dict[int, Simulation] = {}  # resident jobs, until their memory is released
        finished = set()  # finished jobs whose buffers a queued tick still holds
        holds: Dict[int, int] = {}  # job index -> queued ticks whose plan includes it
        batches: List[Optional[Batch]] = [None] * ngroups  # each group's current plan
        ids: List[List[int]] = [[] for _ in range(ngroups)]
        queued: List[deque] = [deque() for _ in range(ngroups)]  # (plan, ids, steps, StatusCopy)
        used = peak = ticks = max_res = done = 0
        t_wait = t_plan = 0.0  # host seconds waiting for tick outcomes / rebuilding plans
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        while done < len(jobs):
            # 1. every group keeps `depth` ticks queued on its current plan, each
            #    followed by an asynchronous copy of its statuses
            for g in range(ngroups):
                b = batches[g]
                while b is not None and len(queued[g]) < self.depth:
                    n = self._tick_steps(ids[g], jobs, queued[g])
                    if n <= 0:
                        break
                    with torch.cuda.stream(streams[g]):
                        b.step(n)
                        queued[g].append((b, ids[g], n, b.status_async()))
                    for i in ids[g]:
                        holds[i] = holds.get(i, 0) + 1
                    ticks += 1
            changed = False
            # 2. arrivals; admission in arrival order while memory and slots
            #    allow (memory and a slot are reserved now, the upload runs on
            #    the upload pool); uploaded jobs join their group
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                j = pending.pop(0)
                queue.append(j)
                if initial is not None:
                    prepared[j.index] = pool.submit(lambda a=initial[j.index]: a)
                else:
                    prepared[j.index] = pool.submit(j.spec.initial_state)
            while (queue and len(active) + len(uploading) < self.max_resident
                   and used + state_bytes(queue[0].spec) <= self.budget):
                j = queue.pop(0)
                uploading[j.index] = uploader.submit(self._upload, j.spec, prepared.pop(j.index), upload)
                used += state_bytes(j.spec)
                j.admitted = time.perf_counter() - t0
            for i in [i for i, f in uploading.items() if f.done()]:
                sim, ev = uploading.pop(i).result()
                g = self.group_of(jobs[i].spec)
                streams[g].wait_event(ev)
                # The buffers come from the upload stream's pool but are used
                # on the group's stream: without this, freeing them after the
                # group's last use (a kept state's clone still queued) would
                # let the next upload reuse them before that use has run.
                for t in (*sim.buf, sim.ctrl):
                    t.record_stream(streams[g])
                with torch.cuda.stream(streams[g]):
                    sim.reset()
                active[i] = sim
                changed = True
            peak = max(peak, used)
            max_res = max(max_res, len(active))
            # 3. completed ticks, oldest first in each group: finished jobs
            #    leave their group's plan, and their memory is released once no
            #    queued tick holds it; a replaced plan is closed after its last
            #    tick
            for g in range(ngroups):
                while queued[g] and queued[g][0][3].ready():
                    b, tids, _, copy = queued[g].popleft()
                    tw = time.perf_counter()
                    statuses = copy.result()
                    t_wait += time.perf_counter() - tw
                    changed = True
                    now = time.perf_counter() - t0
                    for i, st in zip(tids, statuses):
                        holds[i] -= 1
                        j = jobs[i]
                        j.steps_done = max(j.steps_done, int(st.n_iter))
                        if not st.active and i not in finished and math.isnan(j.finished):
                            j.finished = now
                            j.t = float(st.t)
                            j.nonfinite = bool(st.nonfinite)
                            if self.keep_states:
                                with torch.cuda.stream(streams[g]):
                                    j.state = active[i].state(st).clone()
                            finished.add(i)
                            done += 1
                    for i in [i for i in finished if holds[i] == 0]:
                        finished.discard(i)
                        del active[i]
                        used -= state_bytes(jobs[i].spec)
                    if b is not batches[g] and all(t[0] is not b for t in queued[g]):
                        b.close()
            # 4. a group's plan is rebuilt when its chosen jobs change (built
            #    without waiting: the group's queued ticks keep running)
            if changed:
                tp = time.perf_counter()
                members = [[] for _ in range(ngroups)]
                for i in active:
                    if i not in finished:
                        members[self.group_of(jobs[i].spec)].append(i)
                for g in range(ngroups):
                    chosen = self._select(members[g], jobs) if members[g] else []
                    if chosen == ids[g] and (batches[g] is not None or not chosen):
                        continue
                    old = batches[g]
                    if old is not None and all(t[0] is not old for t in queued[g]):
                        old.close()
                    ids[g] = chosen
                    with torch.cuda.stream(streams[g]):
                        batches[g] = Batch([active[i] for i in chosen], tile_x=self.tile_x,
                                           asynchronous=True) if chosen else None
                t_plan += time.perf_counter() - tp
            elif any(queued):
                time.sleep(2e-4)  # the device is busy: poll its ticks
            elif uploading:
                wait(list(uploading.values()), timeout=0.01, return_when=FIRST_COMPLETED)
            elif pending and not queue:
                # idle: wait for the next arrival
                time.sleep(max(0.0, min(0.01, pending[0].arrival - (time.perf_counter() - t0))))
        torch.cuda.synchronize()
        for g in range(ngroups):
            plans = {id(t[0]): t[0] for t in queued[g]}
            if batches[g] is not None:
                plans[id(batches[g])] = batches[g]
            queued[g].clear()
            for b in plans.values():
                b.close()
        makespan = time.perf_counter() - t0
        return ServeReport(jobs, makespan, ticks, max_res, peak,
                           extra={"status_wait_s": t_wait, "plan_rebuild_s": t_plan})


def serve_plain(specs: Sequence[SimSpec], time_scale: float = 1.0, device=None,
                keep_states: bool = False, initial: Optional[Sequence[np.ndarray]] = None) -> ServeReport:
    """First come, first served, one job at a time through run_steps (loop
    tilings of the stream's grid shapes measured beforehand). ``initial`` as
    for Server.serve; without it each job's state is built when it starts."""
    device = device or torch.device("cuda:0")
    if initial is not None and len(initial) != len(specs):
        raise ValueError(f"{len(initial)} initial states for {len(specs)} jobs")
    tune_plain(specs, device)
    jobs = [JobRecord(i, s, s.arrival * time_scale) for i, s in enumerate(specs)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    peak = 0
    for j in jobs:
        now = time.perf_counter() - t0
        if j.arrival > now:
            time.sleep(j.arrival - now)
        j.admitted = time.perf_counter() - t0
        x0 = initial[j.index] if initial is not None else j.spec.initial_state()
        x = torch.from_numpy(x0).to(device)
        state, t, nonfinite = run_steps(x, j.spec.dx, j.spec.gamma, j.spec.steps)
        torch.cuda.synchronize()
        j.finished = time.perf_counter() - t0
        j.steps_done = j.spec.steps
        j.t = t
        j.nonfinite = nonfinite
        if keep_states:
            j.state = state
        peak = max(peak, state_bytes(j.spec))
    return ServeReport(jobs, time.perf_counter() - t0, len(jobs), 1, peak)
