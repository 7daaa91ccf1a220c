"""Hyperviscosity and the spectral barrier on the GPU against the oracle:
fields after 10 steps (1e-12; ETD3 with variable dt 1e-8, as for the plain
ETD3 runs; the barrier within 20x the oracle's own response to a 1e-14
relative change of its f(|k|), which device and host evaluate with different
libm pow), the hdissh/hdissv/hdiss stats columns (1e-10) and the hvis
spectra (1e-11 of each array's largest value).  The density has its own
hyperviscosity (densityHyperViscosity, density_stratification_types.jl:137-149):
the velocity's does not reach it, and its own adds the rhohdiss column."""
import numpy as np
import pytest

import ffs
from ffs import _lib
from ffs.initial import random_scalar, random_spectrum
from oracle.ffs_oracle import Grid, OracleSim

from conftest import print_case

pytestmark = pytest.mark.gpu

N = 32
KXE = N / 2  # KX[end] on a unit box


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


CASES = [
    ("etd3 fixed dt, hyper m=2", "ETD3", False, dict(kind="hyper", nu=2e-6, m=2), False, 1e-12),
    ("etd3 variable dt, hyper m=3", "ETD3", True, dict(kind="hyper", nu=1e-9, m=3), False, 1e-8),
    ("ab3 variable dt, hyper m=2", "AB3", True, dict(kind="hyper", nu=2e-6, m=2), False, 1e-12),
    ("etd3 fixed dt, barrier defaults", "ETD3", False, dict(kind="barrier"), False, 1e-12),
    ("etd3 fixed dt, barrier 4..11 + density", "ETD3", False, dict(kind="barrier", initk=4.0, endk=11.0), True, 1e-12),
]


def barrier_conditioning(make_ref, steps):
    """How far the oracle moves when the barrier's f(|k|) moves by 1e-14 relative.  Device and host
    evaluate pow(|k|, 1/4) with different libm code (one ulp apart for about a fifth of the grid's
    |k|), and endk^(1/4) - |k|^(1/4) cancels, so f differs by up to ~5e-15 relative (measured over
    the retained |k| of this grid).  Modes with c dt near the ETD3 series switch magnify a change
    of c by up to ~1/(c dt)^2."""
    a = make_ref()
    b = make_ref()
    f0 = b.barrier
    b.barrier = lambda k, hv=None: f0(k, hv) * (1 + 1e-14)
    for r in (a, b):
        r.initialize()
        r.advance_in_time(steps)
    return max(rel(b.uh[c], a.uh[c]) for c in range(3))


@pytest.mark.parametrize("name,integ,vdt,hv,dens,tol", CASES, ids=[c[0] for c in CASES])
def test_hyperviscosity_matches_oracle(name, integ, vdt, hv, dens, tol):
    print_case(f"hyperviscosity {name}")
    g = Grid(N, N, N)
    u0 = random_spectrum(N, seed=31, k0=5.0)
    ohv = dict(hv)
    kw = {}
    if hv["kind"] == "hyper":
        kw = dict(hyperviscosity="hyperViscosity", hv_nu=hv["nu"], hv_m=hv["m"])
    else:
        kw = dict(hyperviscosity="spectralBarrier", sb_initk=hv.get("initk", 0.0), sb_endk=hv.get("endk", 0.0))
        # the reference's defaults: initk = cutoff 2 KX[end] / 3, endk = sqrt(3 cutoff) KX[end]
        ohv.setdefault("initk", (2 / 3) * 2 * KXE / 3)
        ohv.setdefault("endk", np.sqrt((2 / 3) * 3) * KXE)
    dkw = {"Pr": 0.7, "gradient": 1.0, "g": -1.0} if dens else None
    r0 = random_scalar(N, N, N, seed=32) if dens else None

    def make_ref():
        r = OracleSim(g, 3e-3, dt=0.0 if vdt else 2e-3, variable_dt=vdt, cfl=0.3, dt_stat=3, dt_spec=4,
                      integrator=integ, hv=ohv, density=dkw)
        r.set_velocity(u0)
        if dens:
            r.set_density(r0)
        return r
    ref = make_ref()
    if hv["kind"] == "barrier":
        tol = max(tol, 20 * barrier_conditioning(make_ref, 10))
        assert tol < 1e-8
    p = ffs.Params(nx=N, ny=N, nz=N, nu=3e-3, integrator="ETD3rdO" if integ == "ETD3" else "Adams_Bashforth3rdO",
                   variable_dt=vdt, dt=0.0 if vdt else 2e-3, cfl=0.3, dt_stat=3, dt_spec=4, write_time=1 << 30, **kw)
    if dens:
        p.density, p.pr, p.density_gradient, p.z_acceleration, p.reference_density = True, 0.7, 1.0, -1.0, 1.0
    with ffs.Simulation(p) as s:
        s.set_velocity(*u0)
        if dens:
            s.set_density(r0)
        ref.initialize()
        s.initialize()
        ref.advance_in_time(10)
        s.advance(10)
        err = max(rel(s.velocity_hat()[c], ref.uh[c]) for c in range(3))
        assert err < tol, err
        if dens:
            assert rel(s.density_hat(), ref.rhoh) < tol
        rows = s.pop_stats()
        assert s.header(_lib.FFS_TABLE_STATS).endswith(",hdissh,hdissv,hdiss")
        assert len(rows) == len(ref.records) == 4
        stol = max(1e-10, 100 * tol)
        for (st, _, _), (ost, _, _) in zip(rows, ref.records):
            assert len(st) == len(ost)
            for col in (-3, -2, -1, 21):   # hdissh, hdissv, hdiss, tdiss
                assert abs(st[col] - ost[col]) <= stol * abs(ost[col]), (col, st[col], ost[col])
            assert st[-1] > 0
        recs = s.pop_spectra()
        assert [r[0] for r in recs] == [r[0] for r in ref.spectra] == [0, 4, 8]
        for (it, got), (_, want) in zip(recs, ref.spectra):
            for q in ("hvis_h", "hvis_v"):
                for b in (0, 1):
                    assert rel(got[q][b], want[q][b]) < max(1e-11, 100 * tol), (it, q, b)


def test_barrier_with_ab3_is_refused():
    p = ffs.Params(nx=16, ny=16, nz=16, nu=3e-3, integrator="Adams_Bashforth3rdO", variable_dt=False, dt=1e-3,
                   dt_stat=10, dt_spec=10, write_time=10, hyperviscosity="spectralBarrier")
    with pytest.raises(_lib.FfsError, match="spectralBarrier needs"):
        ffs.Simulation(p)


DENSITY_CASES = [
    ("density hyper m=2, no velocity hyperviscosity", None, dict(kind="hyper", nu=4e-6, m=2)),
    ("density barrier defaults, velocity hyper m=2", dict(kind="hyper", nu=2e-6, m=2), dict(kind="barrier")),
]


@pytest.mark.parametrize("name,vhv,dhv", DENSITY_CASES, ids=[c[0] for c in DENSITY_CASES])
def test_density_hyperviscosity(name, vhv, dhv):
    """densityHyperViscosity acts on the density's ETD3 operator only, and adds rhohdiss (after bflux)
    to the stats, counted in trhodiss."""
    print_case(f"density hyperviscosity: {name}")
    g = Grid(N, N, N)
    u0 = random_spectrum(N, seed=33, k0=5.0)
    r0 = random_scalar(N, N, N, seed=34)
    odhv = dict(dhv)
    if dhv["kind"] == "barrier":
        odhv.update(initk=(2 / 3) * 2 * KXE / 3, endk=np.sqrt((2 / 3) * 3) * KXE)
    dkw = {"Pr": 0.7, "gradient": 1.0, "g": -1.0, "hv": odhv}

    def make_ref():
        r = OracleSim(g, 3e-3, dt=2e-3, variable_dt=False, dt_stat=3, integrator="ETD3", hv=vhv, density=dkw)
        r.set_velocity(u0)
        r.set_density(r0)
        return r
    ref = make_ref()
    tol = 1e-12
    if dhv["kind"] == "barrier":
        a, b = make_ref(), make_ref()
        f0 = b.barrier
        b.barrier = lambda k, hv=None: f0(k, hv) * (1 + 1e-14)
        for r in (a, b):
            r.initialize()
            r.advance_in_time(10)
        tol = max(tol, 20 * rel(b.rhoh, a.rhoh))
        assert tol < 1e-8
    kw = dict(density_hyperviscosity="hyperViscosity", density_hv_nu=dhv["nu"], density_hv_m=dhv["m"]) \
        if dhv["kind"] == "hyper" else dict(density_hyperviscosity="spectralBarrier")
    if vhv is not None:
        kw.update(hyperviscosity="hyperViscosity", hv_nu=vhv["nu"], hv_m=vhv["m"])
    p = ffs.Params(nx=N, ny=N, nz=N, nu=3e-3, integrator="ETD3rdO", variable_dt=False, dt=2e-3, dt_stat=3,
                   dt_spec=1 << 30, write_time=1 << 30, density=True, pr=0.7, density_gradient=1.0,
                   z_acceleration=-1.0, reference_density=1.0, **kw)
    with ffs.Simulation(p) as s:
        s.set_velocity(*u0)
        s.set_density(r0)
        ref.initialize()
        s.initialize()
        ref.advance_in_time(10)
        s.advance(10)
        assert rel(s.density_hat(), ref.rhoh) < tol
        assert max(rel(s.velocity_hat()[c], ref.uh[c]) for c in range(3)) < tol
        head = s.header(_lib.FFS_TABLE_STATS).split(",")
        assert head[22:31] == ["rho", "rhop2", "drhodxp2", "drhodyp2", "drhodzp2", "rhodiss", "trhodiss", "bflux",
                               "rhohdiss"]
        rows = s.pop_stats()
        assert len(rows) == len(ref.records) == 4
        for (st, _, _), (ost, _, _) in zip(rows, ref.records):
            assert len(st) == len(ost) == len(head)
            for col in (27, 28, 30):  # rhodiss, trhodiss, rhohdiss
                assert abs(st[col] - ost[col]) <= max(1e-10, 100 * tol) * abs(ost[col]), (col, st[col], ost[col])
            assert st[30] > 0 and abs(st[28] - (st[27] + st[30])) <= 1e-14 * st[28]
