"""gravityDirection (density_stratification_types.jl:152-162): g along x or y
moves the buoyancy term of the rhs (models.jl:334-336), the density's mean
gradient term (models.jl:358-360), the buoyancy flux and the pressure statistics
to that axis, while calculate_N and spectrum_buoyancy keep reading g.z (0 off
the z axis).  GPU against the oracle: fields 1e-12 relative, the stats row 1e-10
of its largest entry, the spectra 1e-11 of each array's largest value."""
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


@pytest.mark.parametrize("axis,integrator", [("x", "AB3"), ("y", "ETD3"), ("z", "ETD3")])
def test_gravity_direction(axis, integrator):
    print_case(f"gravityDirection {axis}, {integrator}")
    d = {"x": 0, "y": 1, "z": 2}[axis]
    g = Grid(N, N, N)
    u0 = random_spectrum(N, seed=51, k0=4.0)
    r0 = random_scalar(N, N, N, seed=52, rms=0.1)
    # variable dt with AB3; a fixed dt with ETD3 (with a variable dt its closed forms magnify 1-ulp dt
    # differences, the plain runs' 1e-8 bound)
    vdt = integrator == "AB3"
    ref = OracleSim(g, 5e-3, dt=5e-3, variable_dt=vdt, cfl=0.3, dt_stat=2, dt_spec=3, integrator=integrator,
                    density={"Pr": 1.0, "gradient": 1.0, "g": -1.5, "dir": d})
    ref.set_velocity(u0)
    ref.set_density(r0)
    p = ffs.Params(nx=N, ny=N, nz=N, nu=5e-3, dt=5e-3, variable_dt=vdt, cfl=0.3, dt_stat=2, dt_spec=3,
                   write_time=1 << 30, integrator="Adams_Bashforth3rdO" if integrator == "AB3" else "ETD3rdO",
                   density=True, pr=1.0, density_gradient=1.0, z_acceleration=-1.5, reference_density=1.0,
                   gravity_direction=axis)
    with ffs.Simulation(p) as s:
        s.set_velocity(*u0)
        s.set_density(r0)
        ref.initialize()
        s.initialize()
        for step in range(6):
            ref.advance_in_time(1)
            s.advance(1)
            err = max(rel(s.velocity_hat()[c], ref.uh[c]) for c in range(3))
            assert err < 1e-12 and rel(s.density_hat(), ref.rhoh) < 1e-12, (step, err)
        rows = s.pop_stats()
        recs = s.pop_spectra()
    assert len(rows) == len(ref.records) == 3
    for (st, _, _), (ost, _, _) in zip(rows, ref.records):
        assert np.abs(st - ost).max() <= 1e-8 * np.abs(ost).max()
        assert abs(st[29] - ost[29]) <= 1e-8 * abs(ost[29]) and ost[29] != 0  # bflux along the axis
    assert [r[0] for r in recs] == [r[0] for r in ref.spectra]
    for (it, got), (_, want) in zip(recs, ref.spectra):
        for q in ("press_h", "press_v", "buoyancy_v"):
            for b in (0, 1):
                scale = max(np.abs(want[q][b]).max(), 1e-300)
                assert np.abs(got[q][b] - want[q][b]).max() <= 1e-8 * scale or scale < 1e-30, (it, q, b)
        if axis != "z":
            assert np.all(got["buoyancy_v"][0] == 0)  # spectrum_buoyancy reads g.z
