"""GPU parity of the LES models other than Smagorinsky (Vreman, Silvis,
productionViscosity, stableNonLinear, SandP with each S model; calc_les!,
les_real_calc.jl:74-114; and dynamicSmagorinsky, dynamic_smagorinsky.jl) against
the oracle, through the C ABI: steps of the rotational form with AB3 and ETD3, the
LES stats columns, the real-space stress fields T11 ... T23, and the `global` keys
each model reads.

The oracle's tensor conventions (AntiSymTen, Lie) are pinned by the gradient
model identity in tests/test_les_models.py; the package that defines them in the
reference is not vendored, so these models' parity is unpinned against the
reference itself.  Tolerances: fields 1e-11 relative in FP64 (the models' square
roots and clips at zero are evaluated pointwise on both sides from fields that
agree to roundoff), 2e-4 in FP32 after 3 steps."""
import numpy as np
import pytest

import ffs
from ffs import _lib
from ffs.initial import random_spectrum
from oracle import les_models
from oracle.ffs_oracle import Grid, OracleSim

from conftest import print_case

pytestmark = pytest.mark.gpu

N = 32
CASES = [
    ("Vreman", None, 0.05, None),
    ("Silvis", None, 0.4, 5.0),
    ("productionViscosity", None, 0.4, None),
    ("stableNonLinear", None, 0.1, None),
    ("SandP", "Smagorinsky", 0.16, 0.09),
    ("SandP", "Vreman", 0.05, 0.09),
    ("SandP", "productionViscosity", 0.1, 0.09),
    ("SandP", "fakeSmagorinsky", 0.0, 0.09),
]


def rel(a, b):
    return max(np.abs(a[c] - b[c]).max() / max(np.abs(b[c]).max(), 1e-300) for c in range(len(b)))


def pair(model, smodel, c, cp, integrator="AB3", prec="fp64", variable_dt=True, seed=71):
    g = Grid(N, N, N)
    u0 = random_spectrum(N, seed=seed, k0=3.0)
    ref = OracleSim(g, 2e-3, dt=1e-3, variable_dt=variable_dt, cfl=0.5, dt_stat=2, integrator=integrator,
                    les_c=c, les_model=model, les_cp=cp, les_smodel=smodel)
    ref.set_velocity(u0)
    p = ffs.Params(nx=N, ny=N, nz=N, nu=2e-3, dt=1e-3, variable_dt=variable_dt, cfl=0.5, dt_stat=2,
                   dt_spec=1 << 30, write_time=1 << 30, precision=prec,
                   integrator="Adams_Bashforth3rdO" if integrator == "AB3" else "ETD3rdO",
                   les_model=model, smagorinsky_c=c, les_cp=cp, les_s_model=smodel)
    s = ffs.Simulation(p)
    s.set_velocity(*u0)
    return g, ref, s


@pytest.mark.parametrize("model,smodel,c,cp", CASES, ids=[f"{m}-{sm}" if sm else m for m, sm, _, _ in CASES])
def test_model_steps_fp64(model, smodel, c, cp):
    print_case(f"LES {model} {smodel or ''} fp64 AB3 variable dt {N}^3")
    g, ref, s = pair(model, smodel, c, cp)
    with s:
        ref.initialize()
        s.initialize()
        for step in range(5):
            ref.advance_in_time(1)
            s.advance(1)
            err = rel(s.velocity_hat(), ref.uh)
            assert err < 1e-11, (step, err)
            assert abs(s.dt - ref.dt) <= 1e-12 * ref.dt
        rows = s.pop_stats()
    assert len(rows) == len(ref.records)
    for (st, _, _), (ost, _, _) in zip(rows, ref.records):
        scale = np.abs(ost).max()
        assert np.abs(st - ost).max() <= 1e-10 * scale, np.abs(st - ost).max() / scale
        # the LES columns (prh, prv, praniso, pr) are the last four of the row here
        assert np.allclose(st[-4:], ost[-4:], rtol=1e-9, atol=1e-12 * np.abs(ost[-4:]).max())


@pytest.mark.parametrize("model,smodel", [("Vreman", None), ("stableNonLinear", None), ("SandP", "Smagorinsky")])
def test_model_etd3_fixed_dt(model, smodel):
    c, cp = les_models.default_constants(model, smodel)
    print_case(f"LES {model} {smodel or ''} fp64 ETD3 fixed dt")
    g, ref, s = pair(model, smodel, c, cp, integrator="ETD3", variable_dt=False, seed=72)
    with s:
        ref.initialize()
        s.initialize()
        ref.advance_in_time(4)
        s.advance(4)
        assert rel(s.velocity_hat(), ref.uh) < 1e-11


@pytest.mark.parametrize("model,smodel", [("Silvis", None), ("SandP", "productionViscosity")])
def test_model_fp32(model, smodel):
    c, cp = les_models.default_constants(model, smodel)
    print_case(f"LES {model} {smodel or ''} fp32")
    g, ref, s = pair(model, smodel, c, cp, prec="fp32", variable_dt=False, seed=73)
    with s:
        ref.initialize()
        s.initialize()
        ref.advance_in_time(3)
        s.advance(3)
        assert rel(s.velocity_hat(), ref.uh) < 2e-4


@pytest.mark.parametrize("model,smodel", [("Vreman", None), ("Silvis", None), ("stableNonLinear", None),
                                          ("SandP", "Vreman")])
def test_model_stress_fields(model, smodel):
    """T11 ... T23 in real space: the model's stress of the current velocity."""
    c, cp = les_models.default_constants(model, smodel)
    print_case(f"LES {model} {smodel or ''} stress fields")
    g, ref, s = pair(model, smodel, c, cp, variable_dt=False, seed=74)
    with s:
        ref.initialize()
        s.initialize()
        ref.advance_in_time(2)
        s.advance(2)
        ref.calculate_rhs(collect=False)
        names = (_lib.FFS_FIELD_T11, _lib.FFS_FIELD_T12, _lib.FFS_FIELD_T13, _lib.FFS_FIELD_T22,
                 _lib.FFS_FIELD_T23)
        for k, f in enumerate(names):
            got = s.get_field(f, _lib.FFS_SPACE_REAL)[:N]
            want = ref.taur[k]
            assert np.abs(got - want).max() <= 1e-11 * np.abs(want).max(), (k, np.abs(got - want).max())


def test_default_constants_from_the_library():
    """Unset constants (NaN) take each model's default in the library: a run given the defaults
    explicitly is bitwise the same."""
    print_case("LES model constants' defaults")
    for model, smodel in [("Vreman", None), ("Silvis", None), ("SandP", "productionViscosity")]:
        c, cp = les_models.default_constants(model, smodel)
        _, _, a = pair(model, smodel, None, None, variable_dt=False)
        _, _, b = pair(model, smodel, c, cp, variable_dt=False)
        with a, b:
            a.initialize()
            b.initialize()
            a.advance(2)
            b.advance(2)
            for x, y in zip(a.velocity_hat(), b.velocity_hat()):
                assert np.array_equal(x, y), model


def test_refused_combinations():
    print_case("LES models refused: piomelliSandP over Vreman, standalone fakeSmagorinsky, w-models in the NS form")
    base = dict(nx=16, ny=16, nz=16, nu=1e-2, dt=1e-3, variable_dt=False)
    for kw in (dict(les_model="piomelliSandP", les_s_model="Vreman"), dict(les_model="fakeSmagorinsky"),
               dict(les_model="Vreman", equation="NS"), dict(les_model="SandP", les_s_model="Silvis")):
        with pytest.raises(_lib.FfsError):
            ffs.Simulation(ffs.Params(**base, **kw))


# ---- dynamicSmagorinsky ------------------------------------------------------------------------------
def dyn_pair(shape, integrator="AB3", prec="fp64", variable_dt=True, seed=81, test_width=None, cmin=0.0,
             model="dynamicSmagorinsky", smodel=None):
    nx, ny, nz = shape
    g = Grid(nx, ny, nz)
    u0 = random_spectrum(nx, ny, nz, seed=seed, k0=3.0)
    ref = OracleSim(g, 2e-3, dt=1e-3, variable_dt=variable_dt, cfl=0.5, dt_stat=2, integrator=integrator,
                    les_c=0.0, les_model=model, les_test_width=test_width, les_cmin=cmin, les_smodel=smodel)
    ref.set_velocity(u0)
    p = ffs.Params(nx=nx, ny=ny, nz=nz, nu=2e-3, dt=1e-3, variable_dt=variable_dt, cfl=0.5, dt_stat=2,
                   dt_spec=1 << 30, write_time=1 << 30, precision=prec,
                   integrator="Adams_Bashforth3rdO" if integrator == "AB3" else "ETD3rdO",
                   les_model=model, les_test_width=test_width or 0.0, les_cmin=cmin, les_s_model=smodel)
    s = ffs.Simulation(p)
    s.set_velocity(*u0)
    return g, ref, s


DYN = [("AB3 variable dt 32^3", (32, 32, 32), dict()),
       # nx >= ny: the spectra of iteration 0 bin ky into the kx-spaced shells (calculate_specHV)
       ("ETD3 fixed dt 48x32x24", (48, 32, 24), dict(integrator="ETD3", variable_dt=False)),
       ("AB3 test width 0.5, cmin -0.1, 48^3", (48, 48, 48), dict(test_width=0.5, cmin=-0.1))]


@pytest.mark.parametrize("name,shape,kw", DYN, ids=[d[0] for d in DYN])
def test_dynamic_smagorinsky_steps(name, shape, kw):
    print_case(f"dynamicSmagorinsky {name}")
    g, ref, s = dyn_pair(shape, **kw)
    with s:
        ref.initialize()
        s.initialize()
        for step in range(4):
            ref.advance_in_time(1)
            s.advance(1)
            err = rel(s.velocity_hat(), ref.uh)
            print(f"step {step} err {err:.2e}")
            assert err < 1e-10, (step, err)
        rows = s.pop_stats()
    for (st, _, _), (ost, _, _) in zip(rows, ref.records):
        assert np.allclose(st[-4:], ost[-4:], rtol=1e-9, atol=1e-12 * np.abs(ost[-4:]).max())


def test_dynamic_smagorinsky_fp32_and_stress():
    print_case("dynamicSmagorinsky fp32 steps; fp64 stress fields")
    g, ref, s = dyn_pair((32, 32, 32), prec="fp32", variable_dt=False, seed=82)
    with s:
        ref.initialize()
        s.initialize()
        ref.advance_in_time(3)
        s.advance(3)
        assert rel(s.velocity_hat(), ref.uh) < 2e-4
    g, ref, s = dyn_pair((32, 32, 32), variable_dt=False, seed=83)
    with s:
        ref.initialize()
        s.initialize()
        ref.advance_in_time(2)
        s.advance(2)
        ref.calculate_rhs(collect=False)
        names = (_lib.FFS_FIELD_T11, _lib.FFS_FIELD_T12, _lib.FFS_FIELD_T13, _lib.FFS_FIELD_T22,
                 _lib.FFS_FIELD_T23)
        for k, f in enumerate(names):
            got = s.get_field(f, _lib.FFS_SPACE_REAL)[:32]
            want = ref.taur[k]
            assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max(), (k, np.abs(got - want).max())


PIO = [("AB3 variable dt 32^3", (32, 32, 32), dict()),
       ("ETD3 fixed dt 48x32x24, cmin -0.05", (48, 32, 24), dict(integrator="ETD3", variable_dt=False, cmin=-0.05))]


@pytest.mark.parametrize("name,shape,kw", PIO, ids=[d[0] for d in PIO])
def test_piomelli_smagorinsky_steps(name, shape, kw):
    """The extrapolated coefficient carries over the steps (dt ratio from the device dt), so parity
    over several steps checks the state as well as the filter chain."""
    print_case(f"piomelliSmagorinsky {name}")
    g, ref, s = dyn_pair(shape, model="piomelliSmagorinsky", **kw)
    with s:
        ref.initialize()
        s.initialize()
        for step in range(5):
            ref.advance_in_time(1)
            s.advance(1)
            err = rel(s.velocity_hat(), ref.uh)
            print(f"step {step} err {err:.2e}")
            assert err < 1e-10, (step, err)


def test_piomelli_stress_readout_and_resume():
    """Reading T11..T23 in real space does not move the coefficient state, and a checkpoint carries it:
    a run that reads the stress and a run restored mid-way both stay bitwise with a plain run."""
    print_case("piomelliSmagorinsky stress read-out and checkpoint resume")
    _, ref, a = dyn_pair((32, 32, 32), model="piomelliSmagorinsky", variable_dt=True, seed=84)
    _, _, b = dyn_pair((32, 32, 32), model="piomelliSmagorinsky", variable_dt=True, seed=84)
    with a, b:
        a.initialize()
        b.initialize()
        a.advance(3)
        b.advance(3)
        ref.initialize()
        ref.advance_in_time(3)
        ref.calculate_rhs(collect=False)
        got = b.get_field(_lib.FFS_FIELD_T12, _lib.FFS_SPACE_REAL)[:32]
        assert np.abs(got - ref.taur[1]).max() <= 1e-10 * np.abs(ref.taur[1]).max()
        blob = b.checkpoint()
        a.advance(3)
        b.advance(3)
        for x, y in zip(a.velocity_hat(), b.velocity_hat()):
            assert np.array_equal(x, y)
        _, _, c = dyn_pair((32, 32, 32), model="piomelliSmagorinsky", variable_dt=True, seed=84)
        with c:
            c.restore(blob)
            c.advance(3)
            for x, y in zip(a.velocity_hat(), c.velocity_hat()):
                assert np.array_equal(x, y)


SANDP = [("dynamicSandP", "dynamicSmagorinsky"), ("dynamicSandP", "piomelliSmagorinsky"),
         ("piomelliSandP", "dynamicSmagorinsky"), ("piomelliSandP", "piomelliSmagorinsky")]


@pytest.mark.parametrize("model,smodel", SANDP, ids=[f"{m}-{s}" for m, s in SANDP])
def test_dynamic_sandp_steps(model, smodel):
    """The P part's coefficient (per step, or extrapolated for piomelliSandP) over several steps."""
    print_case(f"{model} over {smodel}")
    g, ref, s = dyn_pair((32, 32, 32), model=model, smodel=smodel, seed=85)
    with s:
        ref.initialize()
        s.initialize()
        for step in range(5):
            ref.advance_in_time(1)
            s.advance(1)
            err = rel(s.velocity_hat(), ref.uh)
            print(f"step {step} err {err:.2e}")
            assert err < 1e-10, (step, err)


def test_piomelli_sandp_readout_and_resume():
    print_case("piomelliSandP over piomelliSmagorinsky: stress read-out and checkpoint resume")
    kw = dict(model="piomelliSandP", smodel="piomelliSmagorinsky", seed=86)
    _, ref, a = dyn_pair((32, 32, 32), **kw)
    _, _, b = dyn_pair((32, 32, 32), **kw)
    with a, b:
        a.initialize()
        b.initialize()
        a.advance(3)
        b.advance(3)
        ref.initialize()
        ref.advance_in_time(3)
        ref.calculate_rhs(collect=False)
        got = b.get_field(_lib.FFS_FIELD_T13, _lib.FFS_SPACE_REAL)[:32]
        assert np.abs(got - ref.taur[2]).max() <= 1e-10 * np.abs(ref.taur[2]).max()
        blob = b.checkpoint()
        a.advance(3)
        b.advance(3)
        for x, y in zip(a.velocity_hat(), b.velocity_hat()):
            assert np.array_equal(x, y)
        _, _, c = dyn_pair((32, 32, 32), **kw)
        with c:
            c.restore(blob)
            c.advance(3)
            for x, y in zip(a.velocity_hat(), c.velocity_hat()):
                assert np.array_equal(x, y)
