"""Exact resume: a run saved after k steps and restored into a fresh handle
continues bitwise as the uninterrupted run (fields, dt, time, statistics),
for AB3 with variable dt and forcing, ETD3 with LES, and Boussinesq FP32; a
checkpoint from a run with other parameters is refused."""
import numpy as np
import pytest

import ffs
from ffs import _lib

from conftest import print_case

pytestmark = pytest.mark.gpu

CASES = [
    ("ab3 variable dt forcing fp64 32^3", dict(forcing="shellForcing")),
    ("etd3 smagorinsky fp64 32^3", dict(integrator="ETD3rdO", les_model="Smagorinsky", smagorinsky_c=0.16)),
    ("ab3 boussinesq fp32 32^3", dict(precision="fp32", density=True, pr=1.0, density_gradient=1.0,
                                      z_acceleration=-1.0, reference_density=1.0)),
]


def params(**kw):
    base = dict(nx=32, ny=32, nz=32, nu=2e-3, integrator="Adams_Bashforth3rdO", variable_dt=True, cfl=0.4,
                dt_stat=4, dt_spec=1 << 30, write_time=1 << 30)
    base.update(kw)
    return ffs.Params(**base)


def fresh(p):
    s = ffs.Simulation(p)
    s.seed_velocity(7)
    if p.density:
        s.seed_density(8)
    s.initialize()
    return s


def state(s):
    f = [s.get_field(c, _lib.FFS_SPACE_FOURIER) for c in range(3)]
    if s.params.density:
        f.append(s.density_hat())
    return f


@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
def test_restored_run_continues_bitwise(name, kw):
    print_case(f"checkpoint {name}")
    p = params(**kw)
    with fresh(p) as a, fresh(p) as b, ffs.Simulation(p) as c:
        a.advance(26)
        b.advance(13)
        b.pop_stats()
        blob = b.checkpoint()
        c.restore(blob)
        assert c.iteration == 13 and c.time == b.time and c.dt == b.dt
        c.advance(13)
        assert c.iteration == a.iteration == 26
        assert c.dt == a.dt and c.time == a.time
        for x, y in zip(state(c), state(a)):
            assert np.array_equal(x, y)
        ra = [r for r in a.pop_stats() if r[0][0] > 13]
        rc = c.pop_stats()
        assert len(ra) == len(rc) > 0
        for x, y in zip(ra, rc):
            for u, v in zip(x, y):
                assert np.array_equal(u, v, equal_nan=True)


def test_checkpoint_of_other_parameters_is_refused():
    print_case("checkpoint parameter mismatch")
    with fresh(params()) as a, ffs.Simulation(params(nx=48, ny=48, nz=48)) as b:
        a.advance(3)
        with pytest.raises(_lib.FfsError):
            b.restore(a.checkpoint())
