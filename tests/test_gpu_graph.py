"""HIP-graph replay and lockstep ensembles against the eager path.

A steady step replayed from a graph runs the same kernels with the same
arguments as the eager path, so every field and every statistics row must be
bitwise identical, after every replay (eager work -- stats steps, field reads,
synchronizations -- sits between the replays).  An ensemble run must be
bitwise identical to the same run advanced alone."""
import numpy as np
import pytest

import ffs
from ffs import _lib
from ffs.initial import random_scalar, random_spectrum

from conftest import print_case

pytestmark = pytest.mark.gpu


def params(n, **kw):
    base = dict(nx=n, ny=n, nz=n, nu=2e-3, integrator="Adams_Bashforth3rdO", variable_dt=True, cfl=0.5,
                dt_stat=5, dt_spec=1 << 30, write_time=1 << 30)
    base.update(kw)
    return ffs.Params(**base)


def start(p, seed, dens=False):
    s = ffs.Simulation(p)
    s.set_velocity(*random_spectrum(p.nx, p.ny, p.nz, seed=seed))
    if dens:
        s.set_density(random_scalar(p.nx, p.ny, p.nz, seed=seed + 1))
    s.initialize()
    return s


def state(s, dens=False):
    f = [s.get_field(c, _lib.FFS_SPACE_FOURIER) for c in range(3)]
    if dens:
        f.append(s.density_hat())
    return f


def assert_same(a, b, what):
    for i, (x, y) in enumerate(zip(a, b)):
        assert np.array_equal(x, y), (what, i, np.abs(x - y).max())


def assert_same_rows(ra, rb, what):
    assert len(ra) == len(rb), what
    for x, y in zip(ra, rb):
        for u, v in zip(x, y):
            assert np.array_equal(u, v, equal_nan=True), what


CASES = [
    ("ab3 variable dt, shell forcing, fp64 48^3", dict(forcing="shellForcing", forcing_kf=2.5), 48, False),
    ("etd3 fixed dt, fp64 32^3", dict(integrator="ETD3rdO", variable_dt=False, dt=2e-3), 32, False),
    ("ab3 smagorinsky + boussinesq, fp64 32^3",
     dict(les_model="Smagorinsky", smagorinsky_c=0.16, density=True, pr=0.7, density_gradient=1.0,
          z_acceleration=-1.0, reference_density=1.0), 32, True),
    ("etd3 variable dt smagorinsky, fp32 32^3",
     dict(integrator="ETD3rdO", les_model="Smagorinsky", smagorinsky_c=0.16, precision="fp32"), 32, False),
]


@pytest.mark.parametrize("name,kw,n,dens", CASES, ids=[c[0] for c in CASES])
def test_graph_replay_matches_eager_every_step(name, kw, n, dens):
    print_case(f"graph vs eager: {name}")
    p = params(n, **kw)
    with start(p, 11, dens) as eager, start(p, 11, dens) as graph:
        eager.set_graph(False)
        for step in range(1, 19):
            eager.advance(1)
            graph.advance(1)
            assert eager.iteration == graph.iteration == step
            assert eager.dt == graph.dt, (step, eager.dt, graph.dt)
            assert_same(state(graph, dens), state(eager, dens), f"{name} step {step}")
        assert_same_rows(graph.pop_stats(), eager.pop_stats(), name)
        # a long stretch of replays without host reads in between
        eager.advance(40)
        graph.advance(40)
        assert_same(state(graph, dens), state(eager, dens), f"{name} after 58 steps")
        assert_same_rows(graph.pop_stats(), eager.pop_stats(), name)


def test_graph_toggle_mid_run():
    print_case("graph toggled off and on mid-run")
    p = params(32, forcing="shellForcing")
    with start(p, 5) as a, start(p, 5) as b:
        a.set_graph(False)
        b.advance(10)
        b.set_graph(False)
        b.advance(7)
        b.set_graph(True)
        b.advance(9)
        a.advance(26)
        assert_same(state(b), state(a), "toggle")


ENSEMBLE_SIZES = [16, 24, 32, 48, 64]


def ensemble_params(n, i):
    return params(n, nu=1.5e-3 + 1e-4 * i, les_model="Smagorinsky", smagorinsky_c=0.16, precision="fp32",
                  dt_stat=4 + i, write_time=9 + 3 * i)


def test_ensemble_matches_runs_alone():
    print_case(f"ensemble of {ENSEMBLE_SIZES} fp32 smagorinsky vs each run alone")
    runs = [start(ensemble_params(n, i), 20 + i) for i, n in enumerate(ENSEMBLE_SIZES)]
    alone = [start(ensemble_params(n, i), 20 + i) for i, n in enumerate(ENSEMBLE_SIZES)]
    for s in alone:
        s.set_graph(False)
    outputs = {i: [] for i in range(len(runs))}
    try:
        with ffs.Ensemble(runs) as ens:
            total = 0
            while total < 45:
                done, ev, who = ens.advance(45 - total)
                total += done
                if ev == _lib.FFS_EVENT_OUTPUT:
                    assert 0 <= who < len(runs)
                    for i, s in enumerate(runs):
                        if s.iteration % s.params.write_time == 0:
                            outputs[i].append(s.iteration)
            ens.synchronize()
        for i, (s, a) in enumerate(zip(runs, alone)):
            left = 45
            while left > 0:
                done, _ = a.advance(left)
                left -= done
            assert s.iteration == a.iteration == 45
            assert_same(state(s), state(a), f"ensemble run {i} n={ENSEMBLE_SIZES[i]}")
            assert_same_rows(s.pop_stats(), a.pop_stats(), f"ensemble run {i}")
            w = s.params.write_time
            assert outputs[i] == [k for k in range(1, 45) if k % w == 0], (i, outputs[i])
    finally:
        for s in runs + alone:
            s.close()


def test_ensemble_then_alone_then_ensemble():
    print_case("ensemble, one run advanced alone, ensemble again")
    a = [start(params(32, nu=2e-3 + 1e-4 * i), 40 + i) for i in range(3)]
    b = [start(params(32, nu=2e-3 + 1e-4 * i), 40 + i) for i in range(3)]
    try:
        with ffs.Ensemble(a) as ens:
            ens.advance(12)
            a[1].advance(2)  # shifts run 1's history phase against the others
            ens.advance(12)
        for i, s in enumerate(b):
            s.advance(24 + (2 if i == 1 else 0))
        for x, y in zip(a, b):
            assert x.iteration == y.iteration
            assert_same(state(x), state(y), "ensemble phases")
    finally:
        for s in a + b:
            s.close()


def test_footprint_matches_allocation():
    for kw in (dict(), dict(precision="fp32"), dict(les_model="Smagorinsky"),
               dict(density=True, pr=1.0, density_gradient=1.0, z_acceleration=-1.0, reference_density=1.0)):
        p = params(48, **kw)
        print_case(f"footprint {kw}")
        with ffs.Simulation(p) as s:
            padded = 2 * (p.nx // 2 + 1) * p.ny * p.nz
            transient = padded * (8 + (4 if p.precision == "fp32" else 0))
            assert ffs.footprint(p) == s.device_bytes() + transient
