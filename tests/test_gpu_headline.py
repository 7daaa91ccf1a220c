"""Parity at the benchmark's own sizes: the 256^3 FP64 forced-turbulence run that bench.py times
(seeded random-phase field, AB3 with CFL 0.5 and variable dt, constant-power shell forcing on
|k| <= 2.5 at power 0.1, nu from k_max eta = 1.5) and a 128^3 FP32 Smagorinsky run of the LES
ensemble's largest size, each against the oracle started from the host twin of the same seeded
field.  The 256^3 run goes through the iteration-0 stats/spectra step, the fused forward-z +
update steps and the graph-replayed steady steps (iteration >= 4), so every kernel of the timed
step is compared at its timed length.  Tolerances: FP64 fields 1e-12 per step relative to the
component's max, dt and time 1e-13; FP32 fields 2e-5, dt 1e-5 (dt = cfl dx / max|u_i| inherits
the FP32 maximum)."""
import numpy as np
import pytest

import ffs
from oracle.ffs_oracle import Grid, OracleSim
from oracle.seeded import seeded_velocity_hat

from conftest import print_case

pytestmark = pytest.mark.gpu


def nu_for_kmax_eta(n, eps=0.1, kmax_eta=1.5):
    eta = kmax_eta / (n / 3.0)
    return (eta**4 * eps) ** (1.0 / 3.0)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def run_pair(n, prec, steps, les_c=None, seed=11, check_every=1):
    nu = nu_for_kmax_eta(n)
    g = Grid(n, n, n)
    uh = seeded_velocity_hat(g, seed, k0=4.0, energy=0.5)
    ref = OracleSim(g, nu, variable_dt=True, cfl=0.5, dt_stat=100, forcing_kf=2.5, forcing_power=0.1, les_c=les_c)
    ref.set_velocity(np.array([g.irfft(c) for c in uh]))
    del uh
    p = ffs.Params(nx=n, ny=n, nz=n, nu=nu, dt=0.0, integrator="Adams_Bashforth3rdO", variable_dt=True, cfl=0.5,
                   dt_stat=100, dt_spec=1 << 30, write_time=1 << 30, precision=prec, forcing="shellForcing",
                   forcing_kf=2.5, forcing_power=0.1)
    if les_c is not None:
        p.les_model, p.smagorinsky_c = "Smagorinsky", les_c
    out = []
    with ffs.Simulation(p) as s:
        s.seed_velocity(seed, k0=4.0, energy=0.5)
        ref.initialize()
        s.initialize()
        for step in range(steps):
            ref.advance_in_time(1)
            s.advance_in_time(1)
            if (step + 1) % check_every == 0 or step + 1 == steps:
                got = s.velocity_hat()
                err = max(rel(got[c], ref.uh[c]) for c in range(3))
                out.append((step, err, s.dt, ref.dt, s.time, ref.time))
                print_case(f"{n}^3 {prec} step {step}: field err {err:.3e}, dt {s.dt:.17g} vs {ref.dt:.17g}")
    return out


def test_forced_256_fp64_matches_oracle():
    print_case("forced isotropic turbulence 256^3 FP64, AB3 CFL 0.5, shell forcing, 6 steps vs oracle")
    for step, err, dt, rdt, t, rt in run_pair(256, "fp64", 6, check_every=2):
        assert abs(dt - rdt) <= 1e-13 * rdt, (step, dt, rdt)
        assert abs(t - rt) <= 1e-13 * rt, (step, t, rt)
        assert err < 1e-12 * (step + 1), (step, err)


def test_les_128_fp32_matches_oracle():
    print_case("Smagorinsky LES 128^3 FP32 (c = 0.16), AB3 CFL 0.5, shell forcing, 3 steps vs FP64 oracle")
    for step, err, dt, rdt, t, rt in run_pair(128, "fp32", 3, les_c=0.16):
        assert abs(dt - rdt) <= 1e-5 * rdt, (step, dt, rdt)
        assert err < 2e-5, (step, err)


def test_forced_256_fp32_matches_oracle():
    """The FP32 leg's own run: 256^3 with the library's default plan (mirror-paired x pass, 128-plane z
    slabs), against the FP64 oracle."""
    print_case("forced isotropic turbulence 256^3 FP32, AB3 CFL 0.5, shell forcing, 3 steps vs FP64 oracle")
    for step, err, dt, rdt, t, rt in run_pair(256, "fp32", 3):
        assert abs(dt - rdt) <= 1e-5 * rdt, (step, dt, rdt)
        assert err < 2e-5, (step, err)
