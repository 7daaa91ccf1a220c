"""A serving loop for many simulation jobs on one GPU.

Jobs (a parameter set, an initial state and a step count) queue up; the loop
admits them while their device footprint fits the HBM budget, advances every
admitted run in lockstep as one ensemble (one graph launch per steady step),
retires runs as they finish and admits queued jobs in their place (continuous
batching).  The reference has no counterpart: it runs one simulation per
Julia process (FluidFlowSimulation.jl:23-65); here each job's results are
identical to running it alone.
"""
import ctypes
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .ensemble import Ensemble
from .simulation import Simulation, footprint

__all__ = ["Job", "Server", "plan_admission", "device_memory"]


def device_memory():
    """(free, total) bytes of the current device."""
    lib = _lib.load()
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    _lib.check(lib.ffs_device_memory(ctypes.byref(f), ctypes.byref(t)))
    return f.value, t.value


@dataclass
class Job:
    """One run: `init(sim)` loads its initial state (set_velocity / set_density)."""

    params: object
    nsteps: int
    init: object
    name: str = ""
    on_output: object = None          # on_output(sim) at the run's output iterations
    keep_fields: bool = True          # fetch the final Fourier state when the run retires
    # filled by the server
    footprint: int = 0
    submitted: float = 0.0
    started: float = 0.0
    finished: float = 0.0
    stats: list = field(default_factory=list)
    fields: list = None
    iteration: int = 0
    time: float = 0.0


def plan_admission(queue_bytes, used, budget, active, max_runs):
    """How many jobs at the head of the queue to admit, first come first served.

    queue_bytes: footprints of the queued jobs in order.  A job is admitted
    while used + its footprint <= budget and fewer than max_runs are active;
    admission stops at the first job that does not fit (no overtaking, so a
    large job is not starved by small ones behind it).  A job larger than the
    whole budget can never run: ValueError."""
    n = 0
    for b in queue_bytes:
        if b > budget:
            raise ValueError(f"a job needs {b} device bytes, more than the budget of {budget}")
        if used + b > budget or active + n >= max_runs:
            break
        used += b
        n += 1
    return n


class Server:
    """Continuous-batching loop over one device.

    budget: device bytes the runs may hold (default: 90% of the free memory at
    construction).  max_runs: cap on concurrently admitted runs.  chunk: most
    lockstep steps per ensemble call between admission points."""

    def __init__(self, budget=None, max_runs=64, chunk=50):
        if budget is None:
            free, _ = device_memory()
            budget = int(0.9 * free)
        self.budget = int(budget)
        self.max_runs = int(max_runs)
        self.chunk = int(chunk)
        self.queue = deque()
        self.active = []          # (job, sim, steps left)
        self.done = []
        self.used = 0
        self.step_ms = []         # per lockstep step: wall ms of the chunk / its steps
        self.step_weights = []    # request-steps at that latency (steps x active runs)
        self.max_active = 0
        self.peak_used = 0
        self._ens = None

    def submit(self, job):
        job.footprint = footprint(job.params)
        if job.footprint > self.budget:
            raise ValueError(f"job {job.name!r} needs {job.footprint} device bytes, budget is {self.budget}")
        job.submitted = time.perf_counter()
        self.queue.append(job)
        return job

    # ---------------------------------------------------------------- loop
    def _admit(self):
        n = plan_admission([j.footprint for j in self.queue], self.used, self.budget, len(self.active),
                           self.max_runs)
        for _ in range(n):
            job = self.queue.popleft()
            sim = Simulation(job.params)
            job.init(sim)
            sim.initialize()
            job.started = time.perf_counter()
            self.used += job.footprint
            self.active.append([job, sim, job.nsteps])
        self.max_active = max(self.max_active, len(self.active))
        self.peak_used = max(self.peak_used, self.used)
        return n

    def _retire(self):
        keep = []
        for entry in self.active:
            job, sim, left = entry
            if left > 0:
                keep.append(entry)
                continue
            sim.synchronize()
            job.stats.extend(sim.pop_stats())
            if job.keep_fields:
                job.fields = [sim.get_field(c, _lib.FFS_SPACE_FOURIER) for c in range(3)]
                if job.params.density:
                    job.fields.append(sim.density_hat())
            job.iteration = sim.iteration
            job.time = sim.time
            sim.close()
            self.used -= job.footprint
            job.finished = time.perf_counter()
            self.done.append(job)
        changed = len(keep) != len(self.active)
        self.active = keep
        return changed

    def run(self, arrivals=None):
        """Serves until the queue is empty and every run has finished; returns the finished jobs.

        arrivals: optional list of (seconds after the start, Job); each job is submitted once its
        time has come, between lockstep chunks (or by waiting when nothing is running)."""
        pending = deque(sorted(arrivals or [], key=lambda a: a[0]))
        t_start = time.perf_counter()

        def arrive():
            now = time.perf_counter() - t_start
            while pending and pending[0][0] <= now:
                self.submit(pending.popleft()[1])

        while self.queue or self.active or pending:
            arrive()
            if not self.queue and not self.active:
                time.sleep(max(0.0, pending[0][0] - (time.perf_counter() - t_start)))
                continue
            admitted = self._admit()
            if not self.active:
                raise RuntimeError("no job could be admitted")
            if admitted or self._ens is None:
                if self._ens is not None:
                    self._ens.close()
                self._ens = Ensemble([s for _, s, _ in self.active])
            k = min(min(left for _, _, left in self.active), self.chunk)
            t0 = time.perf_counter()
            done, ev, _ = self._ens.advance(k)
            if done:
                self._ens.synchronize()
                dt = (time.perf_counter() - t0) * 1e3
                self.step_ms.append(dt / done)
                self.step_weights.append(done * len(self.active))
            for entry in self.active:
                entry[2] -= done
            if ev == _lib.FFS_EVENT_OUTPUT:
                for job, sim, _ in self.active:
                    if job.on_output is not None and sim.iteration % job.params.write_time == 0:
                        job.on_output(sim)
            for job, sim, _ in self.active:   # drain stats rows as they complete
                job.stats.extend(sim.pop_stats())
            if self._retire():
                self._ens.close()
                self._ens = None
        if self._ens is not None:
            self._ens.close()
            self._ens = None
        return self.done

    # ---------------------------------------------------------------- metrics
    def latency(self):
        """Per-step wall latency percentiles (ms) over request-steps (a lockstep step's latency counts once
        per active run), and job turnaround percentiles (s, submission to retirement)."""
        out = {}
        if self.step_ms:
            ms = np.repeat(np.array(self.step_ms), np.array(self.step_weights))
            out["step_ms_p50"] = float(np.percentile(ms, 50))
            out["step_ms_p99"] = float(np.percentile(ms, 99))
        if self.done:
            tt = np.array([j.finished - j.submitted for j in self.done])
            out["job_s_p50"] = float(np.percentile(tt, 50))
            out["job_s_p99"] = float(np.percentile(tt, 99))
        return out
