"""writeoutput's LES read-outs (output.jl:22-41) on the GPU against the oracle:
dflux1..3, the density's real-space LES flux nu_t grad rho (+ (c Delta^2 / 2)
w x grad rho), and cs / cp, the dynamic models' averaged coefficients (with
Piomelli's extrapolation state left as it was: reading twice gives the same
field, and the run goes on as if nothing was read).  Fields 1e-10 of their
largest entry after two steps at 32^3."""
import os

import numpy as np
import pytest

import ffs
from ffs import _lib, fields
from ffs.initial import random_scalar, random_spectrum
from oracle.ffs_oracle import Grid, OracleSim

from conftest import print_case
from test_gpu_les_models import dyn_pair

pytestmark = pytest.mark.gpu

N = 32


def close(got, want, tol=1e-10):
    return np.abs(got - want).max() <= tol * max(np.abs(want).max(), 1e-300)


DYN = [("dynamicSmagorinsky", None), ("piomelliSmagorinsky", None), ("dynamicSandP", "dynamicSmagorinsky"),
       ("piomelliSandP", "piomelliSmagorinsky")]


@pytest.mark.parametrize("model,smodel", DYN, ids=[f"{m} {s or ''}".strip() for m, s in DYN])
def test_dynamic_coefficient_read_out(model, smodel):
    print_case(f"cs/cp read-out: {model} {smodel or ''}")
    _, ref, s = dyn_pair((N, N, N), model=model, smodel=smodel, variable_dt=False, seed=91)
    _, _, twin = dyn_pair((N, N, N), model=model, smodel=smodel, variable_dt=False, seed=91)
    with s, twin:
        ref.initialize()
        s.initialize()
        twin.initialize()
        ref.advance_in_time(2)
        s.advance(2)
        twin.advance(2)
        ref.calculate_rhs(collect=False)
        cs = s.get_field(_lib.FFS_FIELD_CS, _lib.FFS_SPACE_REAL)
        assert close(cs[:N], ref.les_coef), np.abs(cs[:N] - ref.les_coef).max()
        assert np.array_equal(cs, s.get_field(_lib.FFS_FIELD_CS, _lib.FFS_SPACE_REAL))
        if smodel is not None:
            cp = s.get_field(_lib.FFS_FIELD_CP, _lib.FFS_SPACE_REAL)
            assert close(cp[:N], ref.les_pcoef), np.abs(cp[:N] - ref.les_pcoef).max()
        else:
            with pytest.raises(_lib.FfsError, match="no dynamic P coefficient"):
                s.get_field(_lib.FFS_FIELD_CP, _lib.FFS_SPACE_REAL)
        # the read-out leaves the extrapolation state alone: the run matches a twin that read nothing
        s.advance(2)
        twin.advance(2)
        for a, b in zip(s.velocity_hat(), twin.velocity_hat()):
            assert np.array_equal(a, b)


FLUX = [("Smagorinsky, eddy diffusion", "Smagorinsky", "eddyDiffusion", None, "rotational"),
        ("Vreman, vorticity diffusion c=0.3", "Vreman", "vorticityDiffusion", 0.3, "rotational"),
        ("Smagorinsky, eddy diffusion, conservation form", "Smagorinsky", "eddyDiffusion", None, "NS")]


@pytest.mark.parametrize("name,model,dles,cv,eq", FLUX, ids=[f[0] for f in FLUX])
def test_density_les_flux_read_out(tmp_path, name, model, dles, cv, eq):
    print_case(f"dflux read-out: {name}")
    from oracle import les_models
    g = Grid(N, N, N)
    u0 = random_spectrum(N, seed=92, k0=4.0)
    rho0 = random_scalar(N, N, N, seed=93)
    c, cp = les_models.default_constants(model)
    spec = "eddy" if dles == "eddyDiffusion" else ("vorticity", cv)
    ref = OracleSim(g, 3e-3, dt=2e-3, variable_dt=False, dt_stat=3, les_c=c, les_model=model, les_cp=cp,
                    density={"Pr": 0.9, "gradient": 1.0, "g": -1.0}, density_les=spec, equation=eq)
    ref.set_velocity(u0)
    ref.set_density(rho0)
    p = ffs.Params(nx=N, ny=N, nz=N, nu=3e-3, dt=2e-3, variable_dt=False, dt_stat=3, dt_spec=1 << 30,
                   write_time=1 << 30, integrator="Adams_Bashforth3rdO", les_model=model, density=True, pr=0.9,
                   density_gradient=1.0, z_acceleration=-1.0, reference_density=1.0, density_les=dles,
                   density_les_c=cv, equation=eq)
    with ffs.Simulation(p) as s:
        s.set_velocity(*u0)
        s.set_density(rho0)
        ref.initialize()
        s.initialize()
        ref.advance_in_time(2)
        s.advance(2)
        assert close(s.density_hat(), ref.rhoh, 1e-12)
        ref.calculate_rhs(collect=False)
        for k in range(3):
            got = s.get_field(_lib.FFS_FIELD_DFLUX1 + k, _lib.FFS_SPACE_REAL)[:N]
            assert close(got, ref.dfluxr[k]), (k, np.abs(got - ref.dfluxr[k]).max())
        with pytest.raises(_lib.FfsError, match="real-space"):
            s.get_field(_lib.FFS_FIELD_DFLUX1, _lib.FFS_SPACE_FOURIER)
        with pytest.raises(_lib.FfsError, match="no dynamic LES model"):
            s.get_field(_lib.FFS_FIELD_CS, _lib.FFS_SPACE_REAL)
        s.writeoutput(str(tmp_path))
        it = s.iteration
        for k in range(3):
            path = os.path.join(str(tmp_path), f"dflux{k + 1}.{it}")
            assert close(fields.read_field(path, N, N, N)[:N], ref.dfluxr[k])
        assert os.path.exists(os.path.join(str(tmp_path), f"T11.{it}"))
