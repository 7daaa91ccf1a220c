"""A serving loop for many simulation jobs on one GPU.

Jobs (a parameter set, an initial state and a step count) queue up; the loop
admits them first come first served while their device footprint fits the HBM
budget.  Admitted runs are grouped into lanes by grid and precision; each lane
is an ensemble advanced in lockstep (one graph launch per steady step, runs of
the same launch shapes merged into one launch per pass) on its own streams, so
lanes run side by side on the device and a small run is not held to a large
run's pace.  The host never waits for the device as a whole: it polls each
lane, and a lane whose chunk of steps has finished retires its finished runs,
takes in newly admitted ones and is given its next chunk (continuous
batching).  Step latency is measured on the device (timestamps around each
chunk).  The reference has no counterpart: it runs one simulation per Julia
process (FluidFlowSimulation.jl:23-65); here each job's results are identical
to running it alone.
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


def lane_key(params):
    """The lane a job runs in: runs of one grid and precision step at about the same pace."""
    return (params.nx, params.ny, params.nz, params.precision)


class _Lane:
    """Runs of one lane key: an ensemble over them, and the chunk it has in flight."""

    def __init__(self, key):
        self.key = key
        self.entries = []     # [job, sim, steps left]
        self.joining = []     # admitted, joining at the lane's next chunk boundary
        self.ens = None
        self.inflight = False
        self.done = 0         # steps of the chunk in flight
        self.event = _lib.FFS_EVENT_NONE


class Server:
    """Continuous-batching loop over one device.

    budget: device bytes the runs may hold (default: 90% of the free memory at
    construction).  max_runs: cap on concurrently admitted runs.  chunk: most
    lockstep steps a lane is given at a time (its admission and retirement
    points).  plan: every run's kernel plan (see Simulation).  A retired run's
    device memory is released when a queued job needs it or when the loop ends
    (a release waits for the whole device)."""

    def __init__(self, budget=None, max_runs=64, chunk=50, poll_s=5e-5, plan=None):
        if budget is None:
            free, _ = device_memory()
            budget = int(0.9 * free)
        self.budget = int(budget)
        self.max_runs = int(max_runs)
        self.chunk = int(chunk)
        self.poll_s = float(poll_s)
        self.plan = plan  # passed to every run's Simulation ("auto": the tuned plan of its parameters)
        self.queue = deque()
        self.lanes = {}
        self.done = []
        self.used = 0             # bytes held by admitted runs (retired ones included until released)
        self.retired = []         # (sim, bytes) not yet released
        self.step_ms = []         # device ms per lockstep step of each finished chunk
        self.step_weights = []    # request-steps at that latency (steps x runs in the lane)
        self.lane_step_ms = {}    # lane key -> list of (ms per step, request-steps)
        self.max_active = 0
        self.peak_used = 0
        self.releases = 0

    @property
    def active(self):
        return [e for ln in self.lanes.values() for e in ln.entries + ln.joining]

    def submit(self, job):
        job.footprint = footprint(job.params)
        if job.footprint > self.budget:
            raise ValueError(f"job {job.name!r} needs {job.footprint} device bytes, budget is {self.budget}")
        job.submitted = time.perf_counter()
        self.queue.append(job)
        return job

    # ---------------------------------------------------------------- loop
    def _release(self):
        for sim, nbytes in self.retired:
            sim.close()
            self.used -= nbytes
        self.releases += 1 if self.retired else 0
        self.retired = []

    def _admit(self):
        n_active = len(self.active)
        fps = [j.footprint for j in self.queue]
        n = plan_admission(fps, self.used, self.budget, n_active, self.max_runs)
        if n == 0 and self.queue and self.retired and n_active < self.max_runs:
            self._release()  # the head job waits for memory that retired runs still hold
            n = plan_admission(fps, self.used, self.budget, n_active, self.max_runs)
        for _ in range(n):
            job = self.queue.popleft()
            sim = Simulation(job.params, plan=self.plan)
            job.init(sim)
            sim.initialize()
            job.started = time.perf_counter()
            self.used += job.footprint
            key = lane_key(job.params)
            if key not in self.lanes:
                self.lanes[key] = _Lane(key)
            self.lanes[key].joining.append([job, sim, job.nsteps])
        self.max_active = max(self.max_active, len(self.active))
        self.peak_used = max(self.peak_used, self.used)
        return n

    def _finish_chunk(self, ln):
        """The lane's chunk has finished on the device: account its latency, serve outputs, retire."""
        ms, steps = ln.ens.last_advance_ms() if ln.done else (0.0, 0)
        ln.inflight = False
        if steps:
            w = steps * len(ln.entries)
            self.step_ms.append(ms / steps)
            self.step_weights.append(w)
            self.lane_step_ms.setdefault(ln.key, []).append((ms / steps, w))
        for e in ln.entries:
            e[2] -= ln.done
        if ln.event == _lib.FFS_EVENT_OUTPUT:
            for job, sim, _ in ln.entries:
                if job.on_output is not None and sim.iteration % job.params.write_time == 0:
                    job.on_output(sim)
        keep = []
        for entry in ln.entries:
            job, sim, left = entry
            job.stats.extend(sim.pop_stats())
            if left > 0:
                keep.append(entry)
                continue
            if job.keep_fields:
                job.fields = [sim.get_field(c, _lib.FFS_SPACE_FOURIER) for c in range(3)]
                if job.params.density:
                    job.fields.append(sim.density_hat())
            job.iteration = sim.iteration
            job.time = sim.time
            job.finished = time.perf_counter()
            self.retired.append((sim, job.footprint))
            self.done.append(job)
        if len(keep) != len(ln.entries):
            ln.entries = keep
            self._drop_ensemble(ln)

    def _drop_ensemble(self, ln):
        if ln.ens is not None:
            ln.ens.close()
            ln.ens = None

    def _launch(self, ln):
        """Gives an idle lane its next chunk (after taking in the runs admitted to it)."""
        if ln.joining:
            ln.entries += ln.joining
            ln.joining = []
            self._drop_ensemble(ln)
        if not ln.entries:
            return False
        if ln.ens is None:
            ln.ens = Ensemble([s for _, s, _ in ln.entries])
        k = min(min(left for _, _, left in ln.entries), self.chunk)
        ln.done, ln.event, _ = ln.ens.advance(k)
        ln.inflight = True
        return True

    def run(self, arrivals=None):
        """Serves until the queue is empty and every run has finished; returns the finished jobs.

        arrivals: optional list of (seconds after the start, Job); each job is submitted once its
        time has come."""
        pending = deque(sorted(arrivals or [], key=lambda a: a[0]))
        t_start = time.perf_counter()

        def arrive():
            now = time.perf_counter() - t_start
            while pending and pending[0][0] <= now:
                self.submit(pending.popleft()[1])

        while self.queue or pending or any(ln.entries or ln.joining for ln in self.lanes.values()):
            arrive()
            busy = 0
            for ln in list(self.lanes.values()):
                if ln.inflight:
                    if not ln.ens.idle():
                        busy += 1
                        continue
                    self._finish_chunk(ln)
            if self.queue:
                self._admit()
            for key, ln in list(self.lanes.items()):
                if not ln.inflight:
                    if self._launch(ln):
                        busy += 1
                    elif not ln.joining:
                        self._drop_ensemble(ln)
                        del self.lanes[key]
            if not busy:
                if self.queue and not self.lanes:
                    raise RuntimeError("no job could be admitted")
                if pending and not self.queue:
                    time.sleep(max(0.0, pending[0][0] - (time.perf_counter() - t_start)))
            else:
                time.sleep(self.poll_s)
        for ln in self.lanes.values():
            self._drop_ensemble(ln)
        self.lanes = {}
        self._release()
        return self.done

    # ---------------------------------------------------------------- metrics
    def latency(self):
        """Per-step device latency percentiles (ms) over request-steps (a lane's lockstep step counts once per
        run in the lane), per lane as well, and job turnaround percentiles (s, submission to retirement)."""
        out = {}
        if self.step_ms:
            ms = np.repeat(np.array(self.step_ms), np.array(self.step_weights))
            out["step_ms_p50"] = float(np.percentile(ms, 50))
            out["step_ms_p99"] = float(np.percentile(ms, 99))
            lanes = {}
            for key, rows in sorted(self.lane_step_ms.items()):
                a = np.repeat(np.array([r[0] for r in rows]), np.array([r[1] for r in rows]))
                lanes["%dx%dx%d %s" % key] = {"step_ms_p50": float(np.percentile(a, 50)),
                                              "step_ms_p99": float(np.percentile(a, 99)),
                                              "request_steps": int(sum(r[1] for r in rows))}
            out["lanes"] = lanes
        if self.done:
            tt = np.array([j.finished - j.submitted for j in self.done])
            out["job_s_p50"] = float(np.percentile(tt, 50))
            out["job_s_p99"] = float(np.percentile(tt, 99))
        return out
