"""The conservation form (equation NS) on the GPU against the oracle's: fields,
stats rows and spectra after 10 steps (FP64 1e-12, ETD3 with variable dt 1e-8,
FP32 2e-5), with LES, density and the passive scalar riding on it; graph
replay bitwise."""
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


CASES = [
    ("ab3 fixed dt", dict(integrator="Adams_Bashforth3rdO", variable_dt=False, dt=2e-3), {}, "fp64", 1e-12),
    ("ab3 variable dt, smagorinsky", dict(integrator="Adams_Bashforth3rdO", variable_dt=True, cfl=0.3,
                                          les_model="Smagorinsky", smagorinsky_c=0.16), dict(les_c=0.16), "fp64", 1e-12),
    ("etd3 fixed dt, density + passive scalar",
     dict(integrator="ETD3rdO", variable_dt=False, dt=2e-3, density=True, pr=0.7, density_gradient=1.0,
          z_acceleration=-1.0, reference_density=1.0, passive_scalar=True, scalar_pr=1.0, scalar_gradient=0.5),
     dict(density={"Pr": 0.7, "gradient": 1.0, "g": -1.0}, scalar=dict(Pr=1.0, gradient=0.5, dir=2)), "fp64", 1e-12),
    ("ab3 fp32", dict(integrator="Adams_Bashforth3rdO", variable_dt=True, cfl=0.3, precision="fp32"), {}, "fp32", 2e-5),
    # the scalars' eddy-diffusion LES fluxes (les_real_calc.jl:120-136) need no vorticity
    ("ab3 fixed dt, smagorinsky with the density's and the scalar's eddy fluxes",
     dict(integrator="Adams_Bashforth3rdO", variable_dt=False, dt=2e-3, les_model="Smagorinsky", smagorinsky_c=0.16,
          density=True, pr=0.7, density_gradient=1.0, z_acceleration=-1.0, reference_density=1.0,
          density_les="eddyDiffusion", passive_scalar=True, scalar_pr=1.0, scalar_gradient=0.5),
     dict(les_c=0.16, density={"Pr": 0.7, "gradient": 1.0, "g": -1.0}, density_les="eddy",
          scalar=dict(Pr=1.0, gradient=0.5, dir=2), scalar_les="eddy"), "fp64", 1e-12),
]


def test_ns_form_vorticity_diffusion_is_refused():
    p = ffs.Params(nx=16, ny=16, nz=16, nu=5e-3, dt=1e-3, variable_dt=False, dt_stat=10, dt_spec=10, write_time=10,
                   equation="NS", les_model="Smagorinsky", passive_scalar=True, scalar_pr=1.0,
                   scalar_les="vorticityDiffusion")
    with pytest.raises(_lib.FfsError, match="vorticityDiffusion with equation = NS is not implemented"):
        ffs.Simulation(p)


@pytest.mark.parametrize("name,kw,okw,prec,tol", CASES, ids=[c[0] for c in CASES])
def test_ns_form_matches_oracle(name, kw, okw, prec, tol):
    print_case(f"NS form {name}")
    g = Grid(N, N, N)
    u0 = random_spectrum(N, seed=51, k0=3.0)
    integ = "ETD3" if kw["integrator"] == "ETD3rdO" else "AB3"
    vdt = kw["variable_dt"]
    ref = OracleSim(g, 5e-3, dt=kw.get("dt", 0.0), variable_dt=vdt, cfl=kw.get("cfl", 0.18), dt_stat=3, dt_spec=4,
                    integrator=integ, equation="NS", **okw)
    ref.set_velocity(u0)
    p = ffs.Params(nx=N, ny=N, nz=N, nu=5e-3, dt_stat=3, dt_spec=4, write_time=1 << 30, equation="NS", **kw)
    with ffs.Simulation(p) as s:
        s.set_velocity(*u0)
        if "density" in okw:
            r0 = random_scalar(N, N, N, seed=52)
            ref.set_density(r0)
            s.set_density(r0)
        if "scalar" in okw:
            q0 = random_scalar(N, N, N, seed=53)
            ref.set_scalar(q0)
            s.set_field(_lib.FFS_FIELD_SCALAR, q0)
        ref.initialize()
        s.initialize()
        ref.advance_in_time(10)
        s.advance(10)
        err = max(rel(s.velocity_hat()[c], ref.uh[c]) for c in range(3))
        assert err < tol, err
        if "density" in okw:
            assert rel(s.density_hat(), ref.rhoh) < tol
        if "scalar" in okw:
            assert rel(s.get_field(_lib.FFS_FIELD_SCALAR, _lib.FFS_SPACE_FOURIER), ref.phih) < tol
        rows = s.pop_stats()
        assert len(rows) == len(ref.records) == 4
        stol = 1e-10 if prec == "fp64" else 1e-4
        for (st, _, _), (ost, _, _) in zip(rows, ref.records):
            assert len(st) == len(ost)
            for c in (11, 12, 13, 14, 15, 16, 17, 18, 21):   # k, dissh, dissv, diss, nlh, nlv, pressh, pressv, tdiss
                scale = max(abs(ost[c]), abs(ost[14]) * 1e-3)
                assert abs(st[c] - ost[c]) <= stol * scale, (c, st[c], ost[c])
            # the columns after tdiss: the scalar's, the density's (rho, rho2, its derivatives, rhodiss,
            # trhodiss, bflux, rhopr) and the LES ones, in the reference's order
            for c in range(22, len(st)):
                assert abs(st[c] - ost[c]) <= stol * abs(ost[c]) + 1e-14, (c, st[c], ost[c])
        recs = s.pop_spectra()
        assert [r[0] for r in recs] == [r[0] for r in ref.spectra] == [0, 4, 8]
        for (it, got), (_, want) in zip(recs, ref.spectra):
            for q in ("u_h", "nl_h", "nl_v", "press_h", "press_v"):
                for b in (0, 1):
                    assert rel(got[q][b], want[q][b]) < (1e-11 if prec == "fp64" else 1e-4), (it, q, b)


def test_ns_form_graph_replay_bitwise():
    print_case("NS form graph replay")
    p = ffs.Params(nx=N, ny=N, nz=N, nu=5e-3, integrator="Adams_Bashforth3rdO", variable_dt=True, cfl=0.3,
                   dt_stat=7, dt_spec=1 << 30, write_time=1 << 30, equation="NS")
    u0 = random_spectrum(N, seed=54, k0=3.0)
    with ffs.Simulation(p) as a, ffs.Simulation(p) as b:
        for s in (a, b):
            s.set_velocity(*u0)
            s.initialize()
        a.set_graph(False)
        a.advance_in_time(30)
        b.advance_in_time(30)
        for x, y in zip(a.velocity_hat(), b.velocity_hat()):
            assert np.array_equal(x, y)
