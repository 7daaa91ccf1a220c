"""Every kernel plan a run admits (ffs_plan: mirror-paired or plain x pass, waves per x block, split or
joint product launches, fused or separate forward z + update) against the oracle, and the tuner end to
end on the device: it picks one of the candidates, persists it, and Simulation(plan="auto") applies it.
Tolerances as the parity tests: FP64 fields 1e-11 relative to the component's max after three steps,
FP32 2e-5."""
import numpy as np
import pytest

import ffs
from ffs import tune
from ffs.initial import random_scalar, random_spectrum
from oracle.ffs_oracle import Grid, OracleSim

from conftest import print_case

pytestmark = pytest.mark.gpu

DENS = {"Pr": 0.7, "gradient": 1.0, "g": -1.0}
# (name, n, precision, les, density, shell forcing)
CASES = [
    ("plain fp64 64^3, shell forcing", 64, "fp64", False, False, True),
    ("smagorinsky fp32 64^3", 64, "fp32", True, False, False),
    ("smagorinsky + boussinesq fp64 32^3", 32, "fp64", True, True, False),
    ("smagorinsky fp64 32^3", 32, "fp64", True, False, False),
    ("plain fp32 32^3", 32, "fp32", False, False, False),
]


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def make_params(n, prec, les, dens, forcing):
    p = ffs.Params(nx=n, ny=n, nz=n, nu=2e-3, integrator="Adams_Bashforth3rdO", variable_dt=False, dt=1e-3,
                   dt_stat=1 << 30, dt_spec=1 << 30, write_time=1 << 30, precision=prec)
    if les:
        p.les_model, p.smagorinsky_c = "Smagorinsky", 0.16
    if dens:
        p.density, p.pr, p.density_gradient, p.z_acceleration, p.reference_density = True, 0.7, 1.0, -1.0, 1.0
    if forcing:
        p.forcing, p.forcing_kf, p.forcing_power = "shellForcing", 2.5, 0.1
    return p


def run(p, plan, u0, rho0, steps):
    with ffs.Simulation(p, plan=plan) as s:
        assert all(s.plan[k] == v for k, v in plan.items()), (s.plan, plan)
        s.set_velocity(*u0)
        if rho0 is not None:
            s.set_density(rho0)
        s.initialize()
        s.advance(steps)
        out = [s.velocity_hat()[c] for c in range(3)]
        if rho0 is not None:
            out.append(s.density_hat())
        return out


@pytest.mark.parametrize("name,n,prec,les,dens,forcing", CASES, ids=[c[0] for c in CASES])
def test_every_candidate_plan_matches_oracle(name, n, prec, les, dens, forcing):
    p = make_params(n, prec, les, dens, forcing)
    with ffs.Simulation(p) as probe:
        options, base = probe.plan_options(), probe.plan
    cands = tune.candidates(options, base)
    print_case(f"{name}: plan options {options}, {len(cands)} candidates")
    g = Grid(n, n, n)
    u0 = random_spectrum(n, n, n, seed=21)
    rho0 = random_scalar(n, n, n, seed=22) if dens else None
    ref = OracleSim(g, 2e-3, dt=1e-3, variable_dt=False, dt_stat=1 << 30, les_c=0.16 if les else None,
                    density=DENS if dens else None, forcing_kf=2.5 if forcing else None, forcing_power=0.1)
    ref.set_velocity(u0)
    if dens:
        ref.set_density(rho0)
    ref.initialize()
    ref.advance_in_time(3)
    want = [ref.uh[c] for c in range(3)] + ([ref.rhoh] if dens else [])
    tol = 1e-11 if prec == "fp64" else 2e-5
    first = None
    for c in cands:
        print_case(f"{name}: plan {c}")
        got = run(p, c, u0, rho0, 3)
        for i, (a, b) in enumerate(zip(got, want)):
            assert rel(a, b) < tol, (c, i, rel(a, b))
        if first is None:
            first = (c, got)
    # a plan's results are reproducible bit for bit
    again = run(p, first[0], u0, rho0, 3)
    for a, b in zip(again, first[1]):
        assert np.array_equal(a, b)


def test_tuner_on_the_device(tmp_path, monkeypatch):
    path = str(tmp_path / "plans.json")
    monkeypatch.setenv(tune.CACHE_ENV, path)
    p = make_params(32, "fp32", True, False, False)
    print_case("tuning 32^3 fp32 smagorinsky into a fresh cache")
    best = tune.tune(p, min_ms=5.0)
    with ffs.Simulation(p) as probe:
        cands = tune.candidates(probe.plan_options(), probe.plan)
    assert best in cands
    with ffs.Simulation(p, plan="auto") as s:  # read back from the cache
        assert s.plan == best
    import json

    data = json.load(open(path))
    entry = data[tune.device_name()][tune.library_id()][tune.plan_key(p)]
    assert entry["plan"] == best and len(entry["ms_per_step"]) == len(cands)


# (name, (nx, ny, nz), precision, options): runs whose inverse y, x and forward y passes go slab by slab along z
SLAB_CASES = [
    ("plain fp64 32^3, fused update, stats steps", (32, 32, 32), "fp64", {}),
    ("smagorinsky + boussinesq fp32 32^3", (32, 32, 32), "fp32", {"les": True, "dens": True}),
    ("conservation form + passive scalar fp64 24x32x48", (24, 32, 48), "fp64", {"ns": True, "scalar": True}),
    ("plain fp32 48x32x48, shell forcing", (48, 32, 48), "fp32", {"forcing": True}),
]


def slab_params(shape, prec, opts):
    nx, ny, nz = shape
    p = ffs.Params(nx=nx, ny=ny, nz=nz, nu=2e-3, integrator="Adams_Bashforth3rdO", variable_dt=True, cfl=0.3,
                   dt_stat=4, dt_spec=1 << 30, write_time=1 << 30, precision=prec)
    if opts.get("les"):
        p.les_model, p.smagorinsky_c = "Smagorinsky", 0.16
    if opts.get("dens"):
        p.density, p.pr, p.density_gradient, p.z_acceleration, p.reference_density = True, 0.7, 1.0, -1.0, 1.0
    if opts.get("ns"):
        p.equation = "NS"
    if opts.get("scalar"):
        p.passive_scalar, p.scalar_pr, p.scalar_gradient = True, 1.0, 1.0
    if opts.get("forcing"):
        p.forcing, p.forcing_kf, p.forcing_power = "shellForcing", 2.5, 0.1
    return p


@pytest.mark.parametrize("name,shape,prec,opts", SLAB_CASES, ids=[c[0] for c in SLAB_CASES])
def test_z_slabs_match_the_whole_grid_bitwise(name, shape, prec, opts):
    """Slabs change which launch does a plane's work, not the work: fields and stats rows equal the
    whole-grid run's bit for bit, for slabs that divide nz and for a thinner last slab, graph-replayed."""
    p = slab_params(shape, prec, opts)
    nz = shape[2]
    results = {}
    for zs in (0, 8, 5, nz - 1):
        print_case(f"{name}: z_slab {zs}")
        with ffs.Simulation(p) as s:
            s.set_plan({"z_slab": zs})
            assert s.plan["z_slab"] == zs
            s.seed_velocity(31)
            if opts.get("dens"):
                s.seed_density(32)
            s.initialize()
            s.advance(13)
            fields = [s.get_field(c, ffs._lib.FFS_SPACE_FOURIER) for c in range(3)]
            results[zs] = (fields, s.pop_stats())
    f0, r0 = results[0]
    assert len(r0) > 0
    for zs, (f, r) in results.items():
        for c in range(3):
            assert np.array_equal(f[c], f0[c]), (zs, c)
        assert len(r) == len(r0)
        for a, b in zip(r, r0):
            for x, y in zip(a, b):
                assert np.array_equal(x, y, equal_nan=True), zs


def test_slab_defaults_and_plan_options():
    """A 256^3 FP64 run defaults to slabs (its y-x-y buffers exceed the cache); a 32^3 run does not."""
    big = make_params(256, "fp64", False, False, True)
    with ffs.Simulation(big) as s:
        opts, plan = s.plan_options(), s.plan
        print_case(f"256^3 fp64: options {opts}, plan {plan}")
        assert opts["z_slab"] > 0 and plan["z_slab"] == opts["z_slab"] and 256 % plan["z_slab"] == 0
    with ffs.Simulation(make_params(32, "fp64", False, False, False)) as s:
        assert s.plan_options()["z_slab"] == 0 and s.plan["z_slab"] == 0


# (name, (nx, ny, nz), z_slab): grids with more row groups than resident blocks, so every block of the looping
# x pass runs several groups (length 256: 4 rows a group; length 64: 16)
LOOP_CASES = [
    ("256x64x128 whole grid", (256, 64, 128), 0),
    ("256x64x128 slabs of 40 planes", (256, 64, 128), 40),
    ("64x128x256 whole grid", (64, 128, 256), 0),
]


@pytest.mark.parametrize("name,shape,zs", LOOP_CASES, ids=[c[0] for c in LOOP_CASES])
def test_looping_x_pass_matches_the_one_group_pass_bitwise(name, shape, zs):
    """x_pair 2 (the FP64 u x w pass looping over row groups, each next group's first field pair copied into
    LDS ahead) does the same arithmetic as x_pair 1: fields and stats rows (the 1/2|u|^2 variant) equal bit
    for bit, graph-replayed."""
    p = slab_params(shape, "fp64", {"forcing": True})
    results = {}
    for pair in (1, 2):
        print_case(f"{name}: x_pair {pair}")
        with ffs.Simulation(p) as s:
            assert s.plan_options()["x_pair"] == 2
            s.set_plan({"x_pair": pair, "z_slab": zs})
            assert s.plan["x_pair"] == pair and s.plan["z_slab"] == zs
            s.seed_velocity(41)
            s.initialize()
            s.advance(9)
            fields = [s.get_field(c, ffs._lib.FFS_SPACE_FOURIER) for c in range(3)]
            results[pair] = (fields, s.pop_stats())
    (f1, r1), (f2, r2) = results[1], results[2]
    assert len(r1) > 0 and len(r1) == len(r2)
    for c in range(3):
        assert np.array_equal(f1[c], f2[c]), c
    for a, b in zip(r1, r2):
        for x, y in zip(a, b):
            assert np.array_equal(x, y, equal_nan=True)


def test_x_pair_options():
    """The looping x pass applies to FP64 u x w rows (x_pair option 2, the default plan); FP32 runs report 1."""
    with ffs.Simulation(make_params(64, "fp64", False, False, False)) as s:
        assert s.plan_options()["x_pair"] == 2 and s.plan["x_pair"] == 2
    with ffs.Simulation(make_params(64, "fp32", False, False, False)) as s:
        assert s.plan_options()["x_pair"] == 1


def test_set_plan_rejects_an_unknown_x_pass_form():
    """x_pair takes 0 (register pass), 1 (mirror-paired) or 2 (its looping form); anything else is an argument
    error and leaves the plan as it was."""
    with ffs.Simulation(make_params(32, "fp64", False, False, False)) as s:
        before = dict(s.plan)
        with pytest.raises(ffs._lib.FfsError):
            s.set_plan({"x_pair": 3})
        assert s.plan == before
