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


# (n, precision, steps, model): FP64 and FP32 jobs up to 128^3, plain, Smagorinsky and Boussinesq, so the loop
# runs several lanes (one per grid and precision) side by side and merges runs within a lane
MIX = [(128, "fp64", 24, "plain"), (64, "fp32", 40, "les"), (32, "fp64", 36, "les"), (128, "fp32", 20, "les"),
       (48, "fp64", 30, "density"), (64, "fp32", 28, "les"), (32, "fp64", 44, "plain"), (128, "fp64", 16, "plain")]
DENS = dict(density=True, pr=0.7, density_gradient=1.0, z_acceleration=-1.0, reference_density=1.0)


def mix_params(i, n, prec, model):
    kw = dict(nx=n, ny=n, nz=n, nu=1.2e-3 + 2e-4 * i, integrator="Adams_Bashforth3rdO", variable_dt=True, cfl=0.5,
              dt_stat=6, precision=prec, dt_spec=1 << 30, write_time=1 << 30, forcing="shellForcing",
              forcing_kf=2.5, forcing_power=0.1)
    if model == "les":
        kw.update(les_model="Smagorinsky", smagorinsky_c=0.16)
    if model == "density":
        kw.update(DENS)
    return ffs.Params(**kw)


def mix_init(i, model):
    def init(s):
        s.seed_velocity(300 + i)
        if model == "density":
            s.seed_density(400 + i)
    return init


def test_server_fp64_and_large_jobs_match_runs_alone():
    print_case(f"serving loop, lanes of mixed size and precision: {MIX}")
    fps = [ffs.footprint(mix_params(i, n, p, m)) for i, (n, p, _, m) in enumerate(MIX)]
    budget = sorted(fps)[-1] + sorted(fps)[-2] + sorted(fps)[-3]  # a few runs at a time: admissions mid-stream
    srv = Server(budget=budget, max_runs=6, chunk=6)
    arrivals = [(0.002 * i, Job(mix_params(i, n, p, m), steps, mix_init(i, m), name=f"job{i}"))
                for i, (n, p, steps, m) in enumerate(MIX)]
    done = srv.run(arrivals=arrivals)
    assert len(done) == len(MIX)
    assert srv.used == 0
    lat = srv.latency()
    print_case(f"latency {lat}, max active {srv.max_active}, releases {srv.releases}")
    assert lat["step_ms_p50"] > 0 and lat["step_ms_p99"] >= lat["step_ms_p50"]
    assert len(lat["lanes"]) >= 4
    by_name = {j.name: j for j in done}
    for i, (n, prec, steps, model) in enumerate(MIX):
        j = by_name[f"job{i}"]
        assert j.iteration == steps
        with ffs.Simulation(mix_params(i, n, prec, model)) as s:
            mix_init(i, model)(s)
            s.initialize()
            s.set_graph(False)
            s.advance(steps)
            for c in range(3):
                assert np.array_equal(j.fields[c], s.get_field(c, _lib.FFS_SPACE_FOURIER)), (i, c)
            if model == "density":
                assert np.array_equal(j.fields[3], s.density_hat()), i
            rows = s.pop_stats()
            assert len(rows) == len(j.stats), (i, len(rows), len(j.stats))
            for a, b in zip(rows, j.stats):
                for x, y in zip(a, b):
                    assert np.array_equal(x, y, equal_nan=True), i
