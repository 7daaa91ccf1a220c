"""The scalars' LES fluxes (les_real_calc.jl:120-136; simulationtypes.jl:334-346,
density_stratification_types.jl:176-184): with an LES model the passive scalar's
flux is -phi u + nu_t grad phi (EddyDiffusion, its default) or also
(c Delta^2 / 2) w x grad phi (vorticityDiffusion), nu_t being the velocity model's
eddy viscosity; densityLesModel does the same for the density, whose stats gain
rhopr (scalar_les_stats of the total flux, counted in trhodiss).  GPU against the
oracle: fields 1e-12 relative, the stats row 1e-10 of its largest entry."""
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
    ("scalar + Smagorinsky, eddy diffusion, AB3", "AB3", "Smagorinsky", None, ("eddy", None), None),
    ("scalar + Vreman, vorticity diffusion c=1.3, ETD3", "ETD3", "Vreman", None, (("vorticity", 1.3), 1.3), None),
    ("density + Smagorinsky, eddy diffusion, ETD3", "ETD3", "Smagorinsky", None, None, ("eddy", None)),
    ("density vorticity + scalar eddy, Silvis, AB3", "AB3", "Silvis", None, ("eddy", None), (("vorticity", 0.1), None)),
    ("density eddy, SandP over Vreman, AB3", "AB3", "SandP", "Vreman", None, ("eddy", None)),
]


@pytest.mark.parametrize("name,integ,model,smodel,sles,dles", CASES, ids=[c[0] for c in CASES])
def test_scalar_les_flux(name, integ, model, smodel, sles, dles):
    print_case(f"scalar LES: {name}")
    from oracle import les_models
    g = Grid(N, N, N)
    u0 = random_spectrum(N, seed=61, k0=4.0)
    c, cp = les_models.default_constants(model, smodel)
    skw = {"Pr": 0.8, "gradient": 0.5, "dir": 2} if sles else None
    dkw = {"Pr": 0.9, "gradient": 1.0, "g": -1.0} if dles else None
    ref = OracleSim(g, 3e-3, dt=2e-3, variable_dt=False, dt_stat=3, integrator=integ, les_c=c, les_model=model,
                    les_cp=cp, les_smodel=smodel, scalar=skw, density=dkw,
                    scalar_les=sles[0] if sles else None, density_les=dles[0] if dles else None)
    ref.set_velocity(u0)
    p = ffs.Params(nx=N, ny=N, nz=N, nu=3e-3, dt=2e-3, variable_dt=False, dt_stat=3, dt_spec=1 << 30,
                   write_time=1 << 30, integrator="ETD3rdO" if integ == "ETD3" else "Adams_Bashforth3rdO",
                   les_model=model, les_s_model=smodel)
    if sles:
        phi0 = random_scalar(N, N, N, seed=62)
        ref.set_scalar(phi0)
        p.passive_scalar, p.scalar_pr, p.scalar_gradient, p.scalar_direction = True, 0.8, 0.5, "z"
        if sles[0] != "eddy":
            p.scalar_les, p.scalar_les_c = "vorticityDiffusion", sles[1]
    if dles:
        rho0 = random_scalar(N, N, N, seed=63)
        ref.set_density(rho0)
        p.density, p.pr, p.density_gradient, p.z_acceleration, p.reference_density = True, 0.9, 1.0, -1.0, 1.0
        p.density_les = "eddyDiffusion" if dles[0] == "eddy" else "vorticityDiffusion"
    with ffs.Simulation(p) as s:
        s.set_velocity(*u0)
        if sles:
            s.set_field(_lib.FFS_FIELD_SCALAR, phi0)
        if dles:
            s.set_density(rho0)
        ref.initialize()
        s.initialize()
        for step in range(6):
            ref.advance_in_time(1)
            s.advance(1)
            err = max(rel(s.velocity_hat()[k], ref.uh[k]) for k in range(3))
            assert err < 1e-12, (step, err)
            if sles:
                assert rel(s.get_field(_lib.FFS_FIELD_SCALAR, _lib.FFS_SPACE_FOURIER), ref.phih) < 1e-12
            if dles:
                assert rel(s.density_hat(), ref.rhoh) < 1e-12
        head = s.header(_lib.FFS_TABLE_STATS).split(",")
        rows = s.pop_stats()
    if dles:
        assert "rhopr" in head and head[head.index("rhopr") - 1] == "bflux"
    assert len(rows) == len(ref.records) == 2
    for (st, _, _), (ost, _, _) in zip(rows, ref.records):
        assert len(st) == len(ost) == len(head)
        j = int(np.argmax(np.abs(st - ost)))
        assert np.abs(st - ost).max() <= 1e-10 * np.abs(ost).max(), (head[j], st[j], ost[j])
        if dles:
            i = head.index("rhopr")
            assert abs(st[i - 2] - (st[i - 3] + st[i])) <= 1e-12 * abs(st[i - 2])  # trhodiss = rhodiss + rhopr


def test_scalar_les_refusals():
    print_case("scalar LES refusals: densityLesModel without lesModel, dynamic models, vorticity diffusion in NS")
    base = dict(nx=16, ny=16, nz=16, nu=1e-2, dt=1e-3, variable_dt=False)
    dens = dict(density=True, pr=1.0, density_gradient=1.0, z_acceleration=-1.0, reference_density=1.0)
    for kw in (dict(dens, density_les="eddyDiffusion"),
               dict(passive_scalar=True, les_model="dynamicSmagorinsky"),
               dict(passive_scalar=True, les_model="Smagorinsky", equation="NS", scalar_les="vorticityDiffusion")):
        with pytest.raises(_lib.FfsError):
            ffs.Simulation(ffs.Params(**base, **kw))
