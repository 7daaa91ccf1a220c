"""Ensembles: many runs on one GPU advanced in lockstep (ffs_batch_* of the C ABI).

The reference runs one simulation per Julia process (FluidFlowSimulation.jl:23);
an ensemble of small runs on one MI355X would leave most of its 256 CUs idle if
each step were launched run by run.  A steady step of the whole ensemble is one
HIP graph launch whose independent branches are the runs' steps, so the device
overlaps them; each run's results are identical to running it alone.
"""
import ctypes

from . import _lib

__all__ = ["Ensemble"]


class Ensemble:
    """Lockstep advance of initialized `Simulation`s (which it does not own)."""

    def __init__(self, sims, merge=True):
        self.sims = list(sims)
        self._lib = _lib.load()
        arr = (ctypes.c_void_p * len(self.sims))(*[s._h.value for s in self.sims])
        h = ctypes.c_void_p()
        rc = self._lib.ffs_batch_create(arr, len(self.sims), ctypes.byref(h))
        self._check(rc, None)
        self._h = h
        if not merge:
            self.set_merge(False)

    def set_merge(self, on):
        """Merge the launches of runs with the same launch shapes into one launch per pass (default on)."""
        self._check(self._lib.ffs_batch_set_merge(self._h, 1 if on else 0), self._h)

    def launch_info(self):
        """(kernel launches per step, merged groups, runs in them) of the last steady-step graph, or None."""
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        rc = self._lib.ffs_batch_launch_info(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        if rc < 0:
            return None
        return a.value, b.value, c.value

    def _check(self, rc, h=None):
        if rc is not None and rc < 0:
            msg = self._lib.ffs_batch_last_error(h if h is not None else getattr(self, "_h", None))
            raise _lib.FfsError(rc, msg.decode() if msg else "")
        return rc

    def advance(self, nsteps=1):
        """Up to nsteps lockstep steps; returns (steps done, event, index of the run with an output due)."""
        done = ctypes.c_int64()
        ev = ctypes.c_int()
        who = ctypes.c_int()
        self._check(self._lib.ffs_batch_advance(self._h, nsteps, ctypes.byref(done), ctypes.byref(ev),
                                                ctypes.byref(who)), self._h)
        return done.value, ev.value, who.value

    def advance_in_time(self, nsteps, on_output=None):
        """nsteps steps of every run; on_output(sim) for each run at its output iterations."""
        left = nsteps
        while left > 0:
            done, ev, _ = self.advance(left)
            left -= done
            if ev == _lib.FFS_EVENT_OUTPUT and on_output is not None:
                for s in self.sims:
                    if s.iteration % s.params.write_time == 0:
                        on_output(s)
        return self

    def synchronize(self):
        self._check(self._lib.ffs_batch_synchronize(self._h), self._h)

    def idle(self):
        """True when every run's queued work has finished on the device (does not block)."""
        return self._check(self._lib.ffs_batch_query(self._h), self._h) == 1

    def last_advance_ms(self):
        """(device ms, steps) of the last advance call that took a step, or None while it still runs."""
        ms, n = ctypes.c_double(), ctypes.c_int64()
        rc = self._check(self._lib.ffs_batch_last_advance_ms(self._h, ctypes.byref(ms), ctypes.byref(n)), self._h)
        return None if rc == 1 else (ms.value, n.value)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.ffs_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
