"""The serving loop on the GPU: jobs of mixed sizes under an HBM budget that
holds only some of them at once (so runs retire and queued jobs are admitted
mid-stream) must each give results bitwise identical to the same run alone."""
import numpy as np
import pytest

import ffs
from ffs import _lib
from ffs.initial import random_spectrum
from ffs.serve import Job, Server

from conftest import print_case

pytestmark = pytest.mark.gpu

SIZES = [32, 16, 48, 24, 32, 64]
STEPS = [30, 45, 20, 60, 25, 35]


def job_params(i, n):
    return ffs.Params(nx=n, ny=n, nz=n, nu=1.5e-3 + 1e-4 * i, integrator="Adams_Bashforth3rdO", variable_dt=True,
                      cfl=0.5, dt_stat=7, les_model="Smagorinsky", smagorinsky_c=0.16, precision="fp32",
                      dt_spec=1 << 30, write_time=1 << 30)


def init_for(i, n):
    u0 = random_spectrum(n, seed=100 + i)
    return lambda s: s.set_velocity(*u0)


def test_server_results_match_runs_alone():
    print_case(f"serving loop sizes={SIZES} steps={STEPS}")
    fps = [ffs.footprint(job_params(i, n)) for i, n in enumerate(SIZES)]
    budget = sorted(fps)[-1] + sorted(fps)[-2]  # at most about two or three runs at a time
    srv = Server(budget=budget, max_runs=8, chunk=8)
    for i, n in enumerate(SIZES):
        srv.submit(Job(job_params(i, n), STEPS[i], init_for(i, n), name=f"job{i}"))
    done = srv.run()
    assert len(done) == len(SIZES)
    assert srv.used == 0
    lat = srv.latency()
    assert lat["step_ms_p50"] > 0 and lat["step_ms_p99"] >= lat["step_ms_p50"]
    by_name = {j.name: j for j in done}
    for i, n in enumerate(SIZES):
        j = by_name[f"job{i}"]
        assert j.iteration == STEPS[i]
        with ffs.Simulation(job_params(i, n)) as s:
            init_for(i, n)(s)
            s.initialize()
            s.set_graph(False)
            s.advance(STEPS[i])
            for c in range(3):
                assert np.array_equal(j.fields[c], s.get_field(c, _lib.FFS_SPACE_FOURIER)), (i, c)
            rows = s.pop_stats()
            assert len(rows) == len(j.stats), (i, len(rows), len(j.stats))
            for a, b in zip(rows, j.stats):
                for x, y in zip(a, b):
                    assert np.array_equal(x, y, equal_nan=True), i
