"""The forward z pass fused with the update (plain steps) against the two kernels.

On a plain step (rotational form, no LES, density, passive scalar, plane forcing
or hyperviscosity, no stats or spectra at that iteration) the forward z
transform of the nonlinear term and the update run as one kernel.  It computes
the same operations on the same values as the forward z kernel followed by the
update kernel, so every field, the time step and every stats row must be
bitwise identical to the unfused run, with and without graph replay.  (FP64
fuses by default; the FP32 cases ask for it.)"""
import numpy as np
import pytest

import ffs
from ffs import _lib
from ffs.initial import random_spectrum

from conftest import print_case

pytestmark = pytest.mark.gpu


def params(shape, **kw):
    nx, ny, nz = shape
    base = dict(nx=nx, ny=ny, nz=nz, nu=2e-3, integrator="Adams_Bashforth3rdO", variable_dt=True, cfl=0.5,
                dt_stat=7, dt_spec=1 << 30, write_time=1 << 30)
    base.update(kw)
    return ffs.Params(**base)


def start(p, seed):
    s = ffs.Simulation(p)
    s.set_velocity(*random_spectrum(p.nx, p.ny, p.nz, seed=seed))
    s.initialize()
    return s


def fields(s):
    return [s.get_field(c, _lib.FFS_SPACE_FOURIER) for c in range(3)]


CASES = [
    ("ab3 variable dt shell forcing fp64 48^3", (48, 48, 48), dict(forcing="shellForcing", forcing_kf=2.5)),
    ("ab3 fixed dt fp64 32x48x64", (32, 48, 64),
     dict(variable_dt=False, dt=1e-3, x_domain=1.0, y_domain=1.5, z_domain=2.0)),
    ("etd3 variable dt fp64 64^3", (64, 64, 64), dict(integrator="ETD3rdO")),
    ("etd3 fixed dt shell forcing fp32 96^3", (96, 96, 96),
     dict(integrator="ETD3rdO", variable_dt=False, dt=2e-3, precision="fp32", forcing="shellForcing",
          forcing_kf=2.5)),
    ("ab3 variable dt fp32 24x32x16", (24, 32, 16), dict(precision="fp32")),
]


@pytest.mark.parametrize("name,shape,kw", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("graph", [False, True], ids=["eager", "graph"])
def test_fused_update_matches_two_kernels(name, shape, kw, graph):
    print_case(f"fused forward z + update vs two kernels: {name}, graph {graph}")
    p = params(shape, **kw)
    with start(p, 3) as fused, start(p, 3) as split:
        split.set_fusion(False)
        fused.set_fusion(True)  # FP32 runs the two kernels unless asked
        fused.set_graph(graph)
        split.set_graph(graph)
        for step in range(1, 13):
            fused.advance(1)
            split.advance(1)
            assert fused.dt == split.dt, (step, fused.dt, split.dt)
            assert fused.time == split.time, (step, fused.time, split.time)
            for c, (a, b) in enumerate(zip(fields(fused), fields(split))):
                assert np.array_equal(a, b), (name, step, c, np.abs(a - b).max())
        fused.advance(30)
        split.advance(30)
        for a, b in zip(fields(fused), fields(split)):
            assert np.array_equal(a, b), name
        ra, rb = fused.pop_stats(), split.pop_stats()
        assert len(ra) == len(rb) == 6
        for x, y in zip(ra, rb):
            for u, v in zip(x, y):
                assert np.array_equal(u, v, equal_nan=True), name


def test_fusion_toggle_mid_run():
    print_case("fusion toggled off and on mid-run, graphs on")
    p = params((32, 32, 32), forcing="shellForcing", forcing_kf=2.5)
    with start(p, 9) as a, start(p, 9) as b:
        a.set_fusion(False)
        b.set_fusion(True)
        b.advance(10)
        b.set_fusion(False)
        b.advance(7)
        b.set_fusion(True)
        b.advance(9)
        a.advance(26)
        for x, y in zip(fields(a), fields(b)):
            assert np.array_equal(x, y)
