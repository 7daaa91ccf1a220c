"""The Euller integrator (time_step_functions/Euller.jl: u = muladd(dt, rhs, u), the
viscosity explicit in the rhs, dt set without AB3's first-step halving) on the
GPU against the oracle: the velocity, with the shell forcing, and with the density
and the passive scalar (no diffusion on the explicit path, as AB3).  FP64 fields
1e-12 relative, stats 1e-10."""
import numpy as np
import pytest

import ffs
from ffs import _lib
from ffs.initial import random_scalar, random_spectrum
from oracle.ffs_oracle import Grid, OracleSim

from conftest import print_case

pytestmark = pytest.mark.gpu

N = 32


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


CASES = [("fixed dt", dict(variable_dt=False, dt=1e-3), {}),
         ("variable dt, shell forcing", dict(variable_dt=True, dt=0.0, cfl=0.3), dict(forcing_kf=2.5)),
         ("variable dt, density + passive scalar", dict(variable_dt=True, dt=0.0, cfl=0.3), dict(dens=True))]


@pytest.mark.parametrize("name,kw,extra", CASES, ids=[c[0] for c in CASES])
def test_euler_matches_oracle(name, kw, extra):
    print_case(f"Euller: {name}")
    g = Grid(N, N, N)
    u0 = random_spectrum(N, seed=41, k0=4.0)
    dens = extra.get("dens", False)
    dkw = {"Pr": 0.7, "gradient": 1.0, "g": -1.0} if dens else None
    skw = {"Pr": 0.9, "gradient": 0.5, "dir": 1} if dens else None
    ref = OracleSim(g, 5e-3, dt_stat=3, integrator="Euler", density=dkw, scalar=skw,
                    forcing_kf=extra.get("forcing_kf"), **kw)
    ref.set_velocity(u0)
    p = ffs.Params(nx=N, ny=N, nz=N, nu=5e-3, integrator="Euller", dt_stat=3, dt_spec=1 << 30, write_time=1 << 30,
                   forcing="shellForcing" if extra.get("forcing_kf") else None, forcing_kf=2.5, **kw)
    if dens:
        p.density, p.pr, p.density_gradient, p.z_acceleration, p.reference_density = True, 0.7, 1.0, -1.0, 1.0
        p.passive_scalar, p.scalar_pr, p.scalar_gradient, p.scalar_direction = True, 0.9, 0.5, "y"
        r0, f0 = random_scalar(N, N, N, seed=42), random_scalar(N, N, N, seed=43)
        ref.set_density(r0)
        ref.set_scalar(f0)
    with ffs.Simulation(p) as s:
        s.set_velocity(*u0)
        if dens:
            s.set_density(r0)
            s.set_field(_lib.FFS_FIELD_SCALAR, f0)
        ref.initialize()
        s.initialize()
        for step in range(8):
            ref.advance_in_time(1)
            s.advance(1)
            err = max(rel(s.velocity_hat()[c], ref.uh[c]) for c in range(3))
            assert err < 1e-12, (step, err)
            assert abs(s.dt - ref.dt) <= 1e-13 * ref.dt
        if dens:
            assert rel(s.density_hat(), ref.rhoh) < 1e-12
        rows = s.pop_stats()
    assert len(rows) == len(ref.records)
    for (st, _, _), (ost, _, _) in zip(rows, ref.records):
        assert np.abs(st - ost).max() <= 1e-10 * np.abs(ost).max()
