"""Admission control of the serving loop (host logic, no GPU): the first-come
first-served admission plan and the device footprint it is fed, which the
library computes without touching the device."""
import pytest

import ffs
from ffs.serve import plan_admission


def test_admits_in_order_while_it_fits():
    assert plan_admission([3, 3, 3], used=0, budget=10, active=0, max_runs=8) == 3
    assert plan_admission([3, 3, 3], used=2, budget=10, active=0, max_runs=8) == 2
    assert plan_admission([3, 3, 3], used=0, budget=10, active=6, max_runs=8) == 2
    assert plan_admission([], used=0, budget=10, active=0, max_runs=8) == 0


def test_no_overtaking():
    # the 8-byte job does not fit next to the 4 in use; the 1-byte job behind it waits too
    assert plan_admission([8, 1], used=4, budget=10, active=1, max_runs=8) == 0


def test_job_larger_than_budget_is_rejected():
    with pytest.raises(ValueError):
        plan_admission([11], used=0, budget=10, active=0, max_runs=8)


def params(n, **kw):
    return ffs.Params(nx=n, ny=n, nz=n, nu=1e-3, dt_stat=10, dt_spec=100, write_time=1000, **kw)


def test_footprint_scales_with_grid_and_options():
    f = {n: ffs.footprint(params(n)) for n in (32, 64, 128)}
    assert f[32] < f[64] < f[128]
    # O(N^3): doubling the grid multiplies the bulk by about 8
    assert 6.0 < f[128] / f[64] < 9.0
    assert ffs.footprint(params(64, precision="fp32")) < f[64]
    assert ffs.footprint(params(64, les_model="Smagorinsky")) > f[64]
    dens = dict(density=True, pr=1.0, density_gradient=1.0, z_acceleration=-1.0, reference_density=1.0)
    assert ffs.footprint(params(64, **dens)) > f[64]


def test_footprint_of_the_fp64_256_cube_fits_many_times_in_hbm():
    b = ffs.footprint(params(256))
    # u, rhs, fm1, fm2 (3 components each) and 1/2|u|^2 in compact storage plus two pencil buffers:
    # well under 2 GiB, so over a hundred such runs fit in 288 GB
    assert 1.0e9 < b < 2.0e9


def test_footprint_rejects_bad_parameters():
    with pytest.raises(ffs._lib.FfsError):
        ffs.footprint(params(0))
