"""Merged ensemble launches against each run advanced alone.

Runs of an ensemble whose steps launch the same kernels with the same shapes
(same transform lengths, precision and model) are merged into one launch per
pass over all of them, each run reading its own call record (its buffers,
dt history, viscosity, integrator coefficients).  A merged launch runs the
same kernel body on the same arguments as the run's own launch, so every
field and statistics row must be bitwise identical to the run advanced alone,
whatever the runs' history phases and parameters."""
import numpy as np
import pytest

import ffs

from conftest import print_case
from test_gpu_graph import assert_same, assert_same_rows, params, start, state

pytestmark = pytest.mark.gpu

LES32 = dict(les_model="Smagorinsky", smagorinsky_c=0.16, precision="fp32")
DENS = dict(density=True, pr=0.7, density_gradient=1.0, z_acceleration=-1.0, reference_density=1.0)

# (n, params, density): three FP32 LES runs of 32^3 and two of 48^3 (merged groups), one of 64^3 (alone),
# two plain FP64 32^3 runs with shell forcing (the fused forward-z + update step), two FP64 Boussinesq
# 24^3 runs with LES (the general x pass), and an ETD3 fixed-dt FP32 LES run of 32^3
MIX = [
    (32, dict(LES32, nu=1.5e-3), False),
    (48, dict(LES32, nu=1.1e-3), False),
    (32, dict(LES32, nu=2.5e-3), False),
    (64, dict(LES32, nu=1.2e-3), False),
    (48, dict(LES32, nu=3.0e-3), False),
    (32, dict(LES32, nu=1.9e-3), False),
    (32, dict(forcing="shellForcing", forcing_kf=2.5, nu=2e-3), False),
    (32, dict(forcing="shellForcing", forcing_kf=2.5, nu=2.7e-3), False),
    (24, dict(DENS, les_model="Smagorinsky", smagorinsky_c=0.16, nu=2e-3), True),
    (24, dict(DENS, les_model="Smagorinsky", smagorinsky_c=0.16, nu=3e-3), True),
    (32, dict(LES32, integrator="ETD3rdO", variable_dt=False, dt=1e-3, nu=2e-3), False),
]


def make(i, n, kw, dens):
    return start(params(n, dt_stat=7, **kw), 60 + i, dens)


def advance_alone(s, k):
    while k > 0:
        done, _ = s.advance(k)
        k -= done


@pytest.mark.parametrize("merge", [True, False], ids=["merged", "branches"])
def test_mixed_ensemble_matches_runs_alone(merge):
    print_case(f"ensemble of {len(MIX)} mixed runs, merge={merge}, vs each run alone")
    runs = [make(i, *m) for i, m in enumerate(MIX)]
    alone = [make(i, *m) for i, m in enumerate(MIX)]
    for s in alone:
        s.set_graph(False)
    try:
        # run 2 starts two steps ahead, so its history buffers rotate out of phase with its group
        runs[2].advance(2)
        with ffs.Ensemble(runs, merge=merge) as ens:
            ens.advance_in_time(16)
            info = ens.launch_info()
            runs[4].advance(1)  # and run 4 goes one step ahead mid-way
            ens.advance_in_time(15)
            ens.synchronize()
            info2 = ens.launch_info()
        assert info is not None and info2 is not None
        launches, groups, merged = info
        print_case(f"launches per step {launches}, groups {groups}, runs merged {merged}; after the shift {info2}")
        if merge:
            # 32^3 FP32 LES (AB3: runs 0, 2, 5), 48^3 FP32 LES (1, 4), plain FP64 32^3 (6, 7), FP64 Boussinesq
            # LES 24^3 (8, 9); whether the ETD3 run joins the AB3 group depends on its update kernel
            assert groups >= 4 and merged >= 9, info
        else:
            assert groups == 0 and merged == 0, info
        for i, (s, a) in enumerate(zip(runs, alone)):
            advance_alone(a, s.iteration)
            assert s.iteration == a.iteration
            assert s.dt == a.dt, (i, s.dt, a.dt)
            dens = MIX[i][2]
            assert_same(state(s, dens), state(a, dens), f"run {i} n={MIX[i][0]}")
            assert_same_rows(s.pop_stats(), a.pop_stats(), f"run {i}")
    finally:
        for s in runs + alone:
            s.close()


def test_merged_group_counts_launches_once():
    """Eight 32^3 FP32 LES runs merged cost the launches of one run's step."""
    print_case("eight 32^3 fp32 LES runs: launches per merged step")
    runs = [make(i, 32, dict(LES32, nu=1e-3 * (1 + i)), False) for i in range(8)]
    try:
        with ffs.Ensemble(runs, merge=False) as ens:
            ens.advance_in_time(6)
            branch_launches = ens.launch_info()[0]
            ens.set_merge(True)
            ens.advance_in_time(6)
            launches, groups, merged = ens.launch_info()
        print_case(f"branches {branch_launches} launches, merged {launches} ({groups} group, {merged} runs)")
        assert groups == 1 and merged == 8
        assert launches * 8 == branch_launches, (launches, branch_launches)
    finally:
        for s in runs:
            s.close()
