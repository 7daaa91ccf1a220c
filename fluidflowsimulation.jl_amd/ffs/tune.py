"""Run-time autotuner for a run's kernel plan, with a persisted cache.

A plan (ffs_plan, include/ffs/ffs.h) picks among kernel variants that compute
the same step: the mirror-paired or the plain register x pass for
power-of-two rows, one or two wavefronts per x-pass block, the density /
scalar / LES products as register launches of their own or in one general
LDS pass, the forward z pass fused with the update or not, and the inverse
y, x and forward y passes run slab by slab along z or whole.  Which is
fastest depends on the grid, the precision, the model and the device, so the
tuner times every candidate the run admits on a twin of the run (same
parameters, seeded field, no output) and keeps the fastest.  Results are
cached in a JSON file ($FFS_PLAN_CACHE, default ~/.cache/ffs_hip/plans.json)
per device, library build and plan key, so a process tunes a key once and
later processes reuse it.

The reference has no counterpart: FFTW chooses its own plans inside
plan_rfft (spec.jl); the planning flags are the analogue of this cache.
"""
import ctypes
import dataclasses
import hashlib
import itertools
import json
import os
import time

from . import _lib
from .config import Params

__all__ = ["plan_key", "candidates", "tune", "tuned_plan", "cache_path", "device_name", "library_id"]

CACHE_ENV = "FFS_PLAN_CACHE"
FIELDS = ("x_pair", "x_waves", "x_split", "fusion", "z_slab")
_LIB_ID = None


def cache_path():
    return os.environ.get(CACHE_ENV) or os.path.join(os.path.expanduser("~"), ".cache", "ffs_hip", "plans.json")


def device_name():
    """The current device as the library reports it (architecture, compute units, memory)."""
    lib = _lib.load()
    buf = ctypes.create_string_buffer(256)
    _lib.check(lib.ffs_device_name(buf, len(buf)))
    return buf.value.decode()


def library_id():
    """A digest of the loaded libffs_hip.so: plans timed with one build are not reused with another."""
    global _LIB_ID
    if _LIB_ID is None:
        path = _lib.LIB_PATH
        h = hashlib.sha256()
        with open(path, "rb") as f:
            for chunk in iter(lambda: f.read(1 << 20), b""):
                h.update(chunk)
        _LIB_ID = h.hexdigest()[:16]
    return _LIB_ID


def plan_key(params: Params):
    """What decides the kernels a step runs and their cost: the grid, the precision and the model."""
    p = params
    parts = [f"{p.nx}x{p.ny}x{p.nz}", p.precision, f"cut{p.cutoff:g}", str(p.integrator),
             f"les={p.les_model}", f"dens={int(bool(p.density))}", f"scalar={int(bool(p.passive_scalar))}",
             f"eq={p.equation}", f"hv={p.hyperviscosity}", f"forcing={p.forcing}"]
    return "|".join(parts)


def candidates(options, base):
    """The plans worth timing: every combination of the fields that change the run's kernels
    (options: field -> bool, as Simulation.plan_options; z_slab -> the run's default slab planes, 0 when
    slabs do not apply), the other fields as in `base`."""
    out = []
    zs = int(options.get("z_slab", 0) or 0)
    # the default slab, half as thick, and none: never thicker than the default, whose buffers ffs_footprint
    # counts, so a tuned run holds no more device memory than its admission assumed
    slabs = [zs, zs // 2, 0] if zs >= 16 else ([zs, 0] if zs > 0 else None)
    for split in ([1, 0] if options["x_split"] else [base["x_split"]]):
        reg = split == 1 or not options["x_split"]
        pairs = list(range(int(options["x_pair"]), -1, -1)) if (options["x_pair"] and reg) else [base["x_pair"]]
        waves = [2, 1] if (options["x_waves"] and reg) else [base["x_waves"]]
        fusion = [1, 0] if options["fusion"] else [base["fusion"]]
        for pr, w, f in itertools.product(pairs, waves, fusion):
            if pr == 2 and int(options["x_pair"]) == 2 and w != waves[0]:
                continue  # the looping form has blocks of its own shape: x_waves does not change it
            c = {"x_pair": pr, "x_waves": w, "x_split": split, "fusion": f}
            if slabs is None:
                if "z_slab" in base:
                    c["z_slab"] = base["z_slab"]
                out.append(c)
            else:
                out.extend({**c, "z_slab": z} for z in slabs)
    return out


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(path, data):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    tmp = f"{path}.{os.getpid()}.tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def _twin(params):
    """A twin of the run for timing: same kernels, seeded field, no stats, spectra or output steps."""
    from .simulation import Simulation

    q = dataclasses.replace(params, dt_stat=1 << 30, dt_spec=1 << 30, write_time=1 << 30)
    s = Simulation(q)
    s.seed_velocity(12345)
    if q.density:
        s.seed_density(23456)
    s.initialize()
    s.advance(4)  # past the start-up steps, into the graph-replayed steady steps
    return s


def _measure_on(sim, plan, steps):
    sim.set_plan(plan)
    sim.advance(3)  # captures the plan's graphs
    sim.synchronize()
    t0 = time.perf_counter()
    left = steps
    while left > 0:
        done, _ = sim.advance(left)
        left -= done
    sim.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def tune(params, cache=None, rounds=2, min_ms=25.0, refresh=False, device=None, lib=None, measure=None,
         options=None, base=None):
    """Returns the fastest plan for `params` (a dict of ffs_plan fields), timing the candidates on a twin
    unless the cache already holds the key (refresh=True re-times it).  The entry written to the cache
    also holds every candidate's ms per step.  measure(plan, steps) -> ms per step, options and base
    replace the twin (for tests)."""
    path = cache or cache_path()
    device = device or device_name()
    lib = lib or library_id()
    key = plan_key(params)
    data = _load(path)
    entry = data.get(device, {}).get(lib, {}).get(key)
    if entry is not None and not refresh:
        return dict(entry["plan"])
    twin = None
    try:
        if measure is None:
            twin = _twin(params)
            options, base = twin.plan_options(), twin.plan
            measure = lambda plan, steps: _measure_on(twin, plan, steps)  # noqa: E731
        cands = candidates(options, base)
        ms = {}
        if len(cands) > 1:
            probe = measure(cands[0], 4)
            steps = int(min(400, max(8, min_ms / max(probe, 1e-3))))
            for _ in range(rounds):  # interleaved rounds: drift hits every candidate alike
                for c in cands:
                    k = json.dumps(c, sort_keys=True)
                    ms[k] = min(ms.get(k, float("inf")), measure(c, steps))
            best = json.loads(min(ms, key=ms.get))
        else:
            best = cands[0]
    finally:
        if twin is not None:
            twin.close()
    data = _load(path)  # another process may have written meanwhile
    data.setdefault(device, {}).setdefault(lib, {})[key] = {
        "plan": best, "ms_per_step": ms, "tuned_at": time.strftime("%Y-%m-%dT%H:%M:%S")}
    _store(path, data)
    return dict(best)


def tuned_plan(params, **kw):
    """The cached plan for `params`, tuned first if the cache lacks it."""
    return tune(params, **kw)
