"""Device-generated seeded fields against their host twin (oracle/seeded.py):
FP64 within 1e-14 of the max (the transcendentals of Box-Muller may differ by
an ulp between the device and the host libm), FP32 within 1e-6; and a run
started from a seeded field steps like the oracle started from the twin."""
import numpy as np
import pytest

import ffs
from ffs import _lib
from oracle.ffs_oracle import Grid, OracleSim
from oracle.seeded import seeded_scalar_hat, seeded_velocity_hat

from conftest import print_case

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("shape,dom,prec", [((32, 32, 32), (1.0, 1.0, 1.0), "fp64"),
                                            ((24, 32, 48), (1.0, 1.5, 0.75), "fp64"),
                                            ((48, 48, 48), (1.0, 1.0, 1.0), "fp32")])
def test_seeded_velocity_matches_host_twin(shape, dom, prec):
    print_case(f"seeded velocity {shape} {dom} {prec}")
    nx, ny, nz = shape
    g = Grid(nx, ny, nz, *dom)
    ref = seeded_velocity_hat(g, 2024, k0=3.5, energy=0.4)
    p = ffs.Params(nx=nx, ny=ny, nz=nz, x_domain=dom[0], y_domain=dom[1], z_domain=dom[2], nu=0.01, dt=0.01,
                   variable_dt=False, precision=prec)
    tol = 1e-14 if prec == "fp64" else 1e-6
    with ffs.Simulation(p) as s:
        s.seed_velocity(2024, k0=3.5, energy=0.4)
        for c in range(3):
            got = s.get_field(c, _lib.FFS_SPACE_FOURIER)
            assert rel(got, ref[c]) < tol, (c, rel(got, ref[c]))


def test_seeded_density_matches_host_twin():
    print_case("seeded density 32^3")
    g = Grid(32, 32, 32)
    ref = seeded_scalar_hat(g, 77, k0=4.0, rms=0.01)
    p = ffs.Params(nx=32, ny=32, nz=32, nu=0.01, dt=0.01, variable_dt=False, density=True, pr=1.0,
                   density_gradient=1.0, z_acceleration=1.0, reference_density=1.0)
    with ffs.Simulation(p) as s:
        s.seed_density(77, k0=4.0, rms=0.01)
        assert rel(s.density_hat(), ref) < 1e-14


def test_seeded_run_steps_like_the_oracle():
    print_case("seeded 32^3 run, 5 AB3 steps vs oracle from the twin field")
    n = 32
    g = Grid(n, n, n)
    uh = seeded_velocity_hat(g, 3)
    ref = OracleSim(g, 0.01, dt=0.01)
    ref.set_velocity(np.array([g.irfft(x) for x in uh]))
    p = ffs.Params(nx=n, ny=n, nz=n, nu=0.01, dt=0.01, integrator="Adams_Bashforth3rdO", variable_dt=False)
    with ffs.Simulation(p) as s:
        s.seed_velocity(3)
        s.initialize()
        ref.initialize()
        s.advance(5)
        ref.advance_in_time(5)
        got = s.velocity_hat()
        err = max(rel(got[c], ref.uh[c]) for c in range(3))
        assert err < 1e-12, err
