"""Trajectory serving: a continuous-batching request loop over the C ABI.

Jobs (a parameter set and a step count) arrive over time.  The loop admits a job as soon as it
has arrived and its device memory (fhd_bytes_needed) fits in the budget, keeps its fields,
particles and noise counter resident in its own context (own stream, graph-replayed steps), and
advances every admitted job by a few steps per round, so jobs join and leave at step boundaries
and their kernels overlap on the GPU (or, merged, runs them as one merged-stage group).  A job's result does not depend on what runs beside it:
every kernel is deterministic and the noise is keyed by the job's own seed and step.

Reports per-job latency (arrival to completion observed by the loop), p50/p99 and the
aggregate cell-updates/s.  No fluam counterpart: fluam runs one simulation per process.
"""
import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import FhdError, Group, Simulation, bytes_needed, mem_info
from ._abi import FhdParams, SCHEME_QUASI_NEUTRALLY_BUOYANT


@dataclass
class Job:
    name: str
    params: FhdParams
    steps: int
    arrival: float = 0.0                 # seconds after the loop starts
    # filled by the loop
    admitted: Optional[float] = None
    finished: Optional[float] = None
    velocity: Optional[np.ndarray] = None
    particles: Optional[np.ndarray] = None
    error: Optional[str] = None          # set when the job was evicted on a device error
    _sim: Optional[Simulation] = field(default=None, repr=False)
    _issued: int = 0
    _mark: int = 0
    _bytes: int = 0

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival

    @property
    def cell_updates(self):
        return self.params.mx * self.params.my * self.params.mz * self.steps


class Server:
    """Continuous-batching loop.  budget: device bytes the jobs may hold together (default 80%
    of the free device memory).  round_steps: steps each running job issues per round (streams
    mode), so that every stream holds several replays at once.  max_inflight: steps a job may
    run ahead of its completed work, which bounds the host's lead over the GPU.  max_active: most
    jobs admitted at once (None: as many as the memory budget holds); jobs large enough to fill
    the GPU alone run faster one after another than side by side.  keep_results: copy each job's
    final velocity and particles back before its memory is released."""

    def __init__(self, budget: Optional[int] = None, max_inflight: int = 8, keep_results: bool = False,
                 graphs: bool = True, merged: bool = False, chunk: int = 8, round_steps: int = 4,
                 max_active: Optional[int] = None):
        if budget is None:
            free, _ = mem_info()
            budget = int(0.8 * free)
        self.budget = budget
        self.max_inflight = max_inflight
        self.keep_results = keep_results
        self.graphs = graphs
        self.merged = merged
        self.chunk = chunk
        self.round_steps = max(1, round_steps)
        self.max_active = max_active

    def _room(self, active, used, j: Job) -> bool:
        if self.max_active is not None and len(active) >= self.max_active:
            return False
        return used + j._bytes <= self.budget

    def _admit(self, j: Job, t0: float):
        j._sim = Simulation(j.params)
        j._sim.set_graph(self.graphs)
        j.admitted = time.perf_counter() - t0

    def _retire(self, j: Job, t0: float):
        j.finished = time.perf_counter() - t0
        if self.keep_results and not j.error:
            j.velocity = j._sim.velocity()
            if j.params.np > 0:
                j.particles = j._sim.particles()
        j._sim.close()
        j._sim = None

    def _report(self, done: List[Job], wall: float) -> dict:
        ok = [j for j in done if not j.error]
        lat = np.array([j.latency for j in ok]) if ok else np.array([0.0])
        total = sum(j.cell_updates for j in ok)
        return {"jobs": len(ok), "evicted": len(done) - len(ok), "wall_s": wall, "cell_updates_per_s": total / wall,
                "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
                "mean_latency_s": float(lat.mean())}

    def run(self, jobs: List[Job]) -> dict:
        if self.merged:
            return self._run_merged(jobs)
        pending = sorted(jobs, key=lambda j: j.arrival)
        for j in pending:
            j._bytes = bytes_needed(j.params)
            if j._bytes > self.budget:
                raise ValueError(f"job {j.name} needs {j._bytes} bytes, more than the budget {self.budget}")
        active: List[Job] = []
        used = 0
        t0 = time.perf_counter()
        done: List[Job] = []
        while pending or active:
            now = time.perf_counter() - t0
            # admission: arrived jobs, in arrival order, while memory allows
            while pending and pending[0].arrival <= now and self._room(active, used, pending[0]):
                j = pending.pop(0)
                self._admit(j, t0)
                used += j._bytes
                active.append(j)
            # one round: every admitted job advances up to round_steps steps.  Run-ahead is
            # bounded: once a job has issued max_inflight steps past its last marker, it waits
            # (without blocking the others) until the GPU has reached that marker.
            progressed = False
            for j in active:
                if j._issued >= j.steps or j.error:
                    continue
                k = min(self.round_steps, j.steps - j._issued)
                try:
                    if j._issued - j._mark >= self.max_inflight:
                        if not j._sim.done():
                            continue
                        j._sim.record()
                        j._mark = j._issued
                    j._sim.step(k)
                except FhdError as e:
                    # a device error flagged by this job's steps (fluam's errorKernel): the job is
                    # evicted at once; the others keep running
                    j.error = str(e)
                    continue
                j._issued += k
                progressed = True
                if j._issued == j.steps:
                    j._sim.record()
                    j._mark = j._issued
            # retire jobs whose last step has completed, and evicted ones
            still = []
            for j in active:
                finished = j.error is not None
                if not finished and j._issued == j.steps:
                    try:
                        finished = j._sim.done()
                        if finished:
                            j._sim.sync()
                    except FhdError as e:
                        j.error = str(e)
                        finished = True
                if finished:
                    self._retire(j, t0)
                    used -= j._bytes
                    done.append(j)
                    progressed = True
                else:
                    still.append(j)
            active = still
            if active and not progressed:
                time.sleep(0.0002)
            if not active and pending:
                time.sleep(max(0.0, pending[0].arrival - (time.perf_counter() - t0)))
        return self._report(done, time.perf_counter() - t0)

    def _run_merged(self, jobs: List[Job]) -> dict:
        """The running jobs as one merged-stage batch (fhd_batch_adopt, mode 2: one dispatch per
        stage and kernel instance for all of them), regrouped whenever a job joins or leaves.
        The group advances in chunks of up to `chunk` steps, cut at the next job's end; a job
        leaves once its last chunk has completed.  Inertial-coupling jobs take their two
        bootstrap steps alone before they join (a merged program needs steady-state steps)."""
        pending = sorted(jobs, key=lambda j: j.arrival)
        for j in pending:
            j._bytes = bytes_needed(j.params)
            if j._bytes > self.budget:
                raise ValueError(f"job {j.name} needs {j._bytes} bytes, more than the budget {self.budget}")
        active: List[Job] = []
        done: List[Job] = []
        used = 0
        group = None
        t0 = time.perf_counter()
        while pending or active:
            now = time.perf_counter() - t0
            changed = False
            while pending and pending[0].arrival <= now and self._room(active, used, pending[0]):
                j = pending.pop(0)
                self._admit(j, t0)
                if j.params.scheme == SCHEME_QUASI_NEUTRALLY_BUOYANT and j.params.np > 0:
                    k = min(2, j.steps)
                    j._sim.step(k)
                    j._issued = k
                used += j._bytes
                active.append(j)
                changed = True
            running = [j for j in active if j._issued < j.steps]
            members = [j._sim for j in running]
            if group is not None and (changed or group.replicas != members):
                group.close()
                group = None
            if group is None and running:
                group = Group(members)
                group.set_mode(2)
            if running:
                k = min(self.chunk, min(j.steps - j._issued for j in running))
                try:
                    group.step(k)
                    group.sync()
                except FhdError:
                    # a device error flagged by some job's steps: evict the flagged ones
                    for j in running:
                        try:
                            j._sim.sync()
                        except FhdError as e:
                            j.error = str(e)
                for j in running:
                    j._issued += k
            # retire finished and evicted jobs (their work has completed: the group was synced)
            still = []
            for j in active:
                if j.error is not None or j._issued >= j.steps:
                    if j.error is None:
                        try:
                            j._sim.sync()
                        except FhdError as e:
                            j.error = str(e)
                    if group is not None and j._sim in group.replicas:
                        group.close()
                        group = None
                    self._retire(j, t0)
                    used -= j._bytes
                    done.append(j)
                else:
                    still.append(j)
            active = still
            if not active and pending:
                time.sleep(max(0.0, pending[0].arrival - (time.perf_counter() - t0)))
        if group is not None:
            group.close()
        return self._report(done, time.perf_counter() - t0)


def poisson_jobs(rate: float, n: int, seed: int, make) -> List[Job]:
    """n jobs with exponential inter-arrival times (rate per second); make(i, rng) -> (params, steps)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        params, steps = make(i, rng)
        jobs.append(Job(name=f"job{i}", params=params, steps=steps, arrival=t))
    return jobs
