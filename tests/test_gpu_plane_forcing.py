"""The reference's kz = 0-plane forcing schemes on the GPU against the oracle:
rfForcing, cRfForcing, nRfForcing, cNRfForcing and aForcing with a prescribed
target and with the target taken from the initial field, QfForcing with and
without the LES filter, the N variants with LES and with density (their N
spectrum reads the full right-hand side), plus graph replay and exact resume
with the oscillator state."""
import numpy as np
import pytest

import ffs
from ffs import _lib
from ffs.initial import random_scalar, random_spectrum
from oracle.ffs_oracle import Grid, OracleSim
from oracle.plane_forcing import PlaneForcing

from conftest import print_case

pytestmark = pytest.mark.gpu

N = 32


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


NAMES = {"rf": "rfForcing", "crf": "cRfForcing", "nrf": "nRfForcing", "cnrf": "cNRfForcing", "a": "aForcing"}


def flat(x):
    """the reference's arrays as the C-ABI lays them out (aForcing's (NX, NY) matrices x fastest)"""
    return np.asarray(x).flatten(order="F")


def forced(pf):
    """positions (in flat() layout) that carry an oscillator: the forced shells, aForcing's forced modes"""
    g = pf.g
    if pf.kind == "a":
        m = np.zeros((g.NX, g.NY), dtype=bool)
        for i in g.XRANGE:
            for j in g.YRANGE:
                m[i, j] = 0 < np.hypot(g.KX[i], g.KY[j]) <= pf.kf
        return flat(m)
    return np.array([0 < i and pf.kh[i] <= pf.kf for i in range(pf.nshells)])


def make(kind, les=False, target=True, dens=False, oracle_only=False, perturb=0.0):
    """(oracle run, device run, oracle forcing) of one case; oracle_only skips the device run and
    perturb scales the oracle's initial coefficients by 1 + perturb * noise"""
    g = Grid(N, N, N)
    u0 = random_spectrum(N, seed=12, k0=3.0)
    delta = (g.LX * 2 * np.pi / g.NRX) / g.cutoff if les else None
    if kind == "qf":
        pf = PlaneForcing(g, "qf", kf=3.0, P=0.1, C=1.0, Kf_alv=2.0, les_delta=delta)
        kw = dict(forcing="QfForcing", forcing_kf=3.0, alvelius_p=0.1, alvelius_c=1.0, alvelius_kf=2.0)
    else:
        tf = 1.0 if kind in ("rf", "nrf", "a") else 0.5
        kf = 3.0 if kind == "a" else 2.5
        pf = PlaneForcing(g, kind, kf=kf, tf=tf, alphac=0.5)
        kw = dict(forcing=NAMES[kind], forcing_kf=kf, forcing_tf=tf, forcing_alpha=0.5)
    dens_kw = {"Pr": 0.7, "gradient": 1.0, "g": -1.0} if dens else None
    ref = OracleSim(g, 2e-3, dt=0.0, variable_dt=True, cfl=0.4, dt_stat=3, plane_forcing=pf,
                    les_c=0.16 if les else None, density=dens_kw)
    ref.set_velocity(u0)
    p = ffs.Params(nx=N, ny=N, nz=N, nu=2e-3, integrator="Adams_Bashforth3rdO", variable_dt=True, cfl=0.4,
                   dt_stat=3, dt_spec=1 << 30, write_time=1 << 30,
                   les_model="Smagorinsky" if les else None, smagorinsky_c=0.16, **kw)
    if dens:
        p.density, p.pr, p.density_gradient, p.z_acceleration, p.reference_density = True, 0.7, 1.0, -1.0, 1.0
    rho0 = random_scalar(N, N, N, seed=13) if dens else None
    if dens:
        ref.set_density(rho0)
    em = None
    if kind != "qf" and target:
        # a prescribed target above the initial spectrum, so the oscillators drive energy in
        em = 3 * (pf.abs2_plane(ref.uh) if kind == "a" else pf.u1u2_spectrum(ref.uh))
        pf.Em = em.copy()
        pf.init = False
    if perturb:
        ref.uh = ref.uh * (1 + perturb * np.random.default_rng(1).standard_normal(ref.uh.shape))
    if kind == "a" and em is not None:
        # aForcing's initialize! always takes the initial field's energies: the target goes in after it
        init = ref.initialize

        def initialize_then_target():
            init()
            pf.Em = em.copy()
        ref.initialize = initialize_then_target
    if oracle_only:
        return ref, None, pf
    s = ffs.Simulation(p)
    s.set_velocity(*u0)
    if dens:
        s.set_density(rho0)
    if em is not None:
        if kind == "a":
            sinit = s.initialize

            def sim_initialize_then_target():
                sinit()
                s.set_forcing_state(_lib.FFS_FSTATE_EM, flat(em))
            s.initialize = sim_initialize_then_target
        else:
            s.set_forcing_state(_lib.FFS_FSTATE_EM, flat(em))
    return ref, s, pf


def conditioning(kind, les, target, dens, steps):
    """How far the oracle itself moves after `steps` steps when its initial coefficients move by
    ~1e-15 relative.  The oscillators' law R += dt (... - w^2 (Ef - Em)) multiplies spectrum
    roundoff by w^2 dt (~1e3-1e4 here), and the force sqrt(R / Ef) dt is not Lipschitz where R
    touches its clamp at 0, so runs whose oscillators start at their target (R ~ roundoff) carry
    roundoff of two correct implementations up to ~1e-8 (aForcing, per mode); planted mistakes of
    the law (w off by 10 %, the N difference term dropped) move the field by ~1e-3."""
    a, _, _ = make(kind, les, target, dens, oracle_only=True)
    b, _, _ = make(kind, les, target, dens, oracle_only=True, perturb=1e-15)
    for r in (a, b):
        r.initialize()
        r.advance_in_time(steps)
    return max(rel(b.uh[c], a.uh[c]) for c in range(3))


CASES = [("rf", False, True, False), ("qf", False, True, False), ("qf", True, True, False),
         ("crf", False, True, False), ("nrf", False, False, False), ("nrf", True, True, False),
         ("cnrf", False, True, True), ("cnrf", False, False, False), ("a", False, False, False),
         ("a", True, True, False)]


@pytest.mark.parametrize("kind,les,target,dens", CASES,
                         ids=[f"{k}{'-les' if l else ''}{'-target' if t else '-init'}{'-density' if d else ''}"
                              for k, l, t, d in CASES])
def test_plane_forcing_matches_oracle(kind, les, target, dens):
    print_case(f"plane forcing {kind} les={les} target={target} density={dens}")
    ref, s, pf = make(kind, les, target, dens)
    with s:
        ref.initialize()
        s.initialize()
        if kind != "a":
            kh = s.forcing_state(_lib.FFS_FSTATE_KH)
            assert np.allclose(kh, pf.kh, rtol=1e-15, atol=0, equal_nan=True)
        m = forced(pf)
        assert m.sum() >= 2
        em = s.forcing_state(_lib.FFS_FSTATE_EM)
        assert not np.any(em[~m])  # the device holds the oscillators' shells / modes only
        if kind != "qf":
            assert np.allclose(em[m], flat(pf.Em)[m], rtol=1e-12, atol=1e-14 * np.abs(pf.Em).max())
        if kind in ("nrf", "cnrf"):
            n0 = s.forcing_state(_lib.FFS_FSTATE_N)[m]
            assert np.allclose(n0, pf.N[m], rtol=1e-11, atol=1e-12 * max(np.abs(pf.N).max(), 1e-300)), (n0, pf.N)
        ref.advance_in_time(15)
        s.advance(15)
        # 1e-12, or 20x the oracle's own conditioning where the oscillators amplify roundoff
        tol = max(1e-12, 20 * conditioning(kind, les, target, dens, 15))
        assert tol < 1e-6
        err = max(rel(s.velocity_hat()[c], ref.uh[c]) for c in range(3))
        assert err < tol, (err, tol)
        if dens:
            assert rel(s.density_hat(), ref.rhoh) < tol
        if kind != "a":  # Ef over every shell (the Eh.N values) from the last forcing call
            ef = s.forcing_state(_lib.FFS_FSTATE_EF)
            assert ef.size == pf.nshells
            assert np.allclose(ef, pf.Ef, rtol=max(1e-12, 1e2 * tol), atol=1e-14 * np.abs(pf.Ef).max())
        R = s.forcing_state(_lib.FFS_FSTATE_R)
        rtol = 1e-11 if tol == 1e-12 else 1e3 * tol
        assert np.allclose(R, flat(pf.R), rtol=rtol, atol=rtol * max(np.abs(pf.R).max(), 1e-300)), (R, pf.R)
        if kind in ("nrf", "cnrf"):
            n1 = s.forcing_state(_lib.FFS_FSTATE_N)[m]
            assert np.allclose(n1, pf.N[m], rtol=1e2 * rtol, atol=rtol * max(np.abs(pf.N).max(), 1e-300)), (n1, pf.N)
        rows = s.pop_stats()
        ncol = 2 if kind == "a" else 3
        assert any(r[0][-ncol] > 0 for r in rows)  # the forcing did work on the flow
        assert len(rows) == len(ref.records) == 5
        # powers relative to their largest value over the run (a row where an oscillator sits at its
        # clamp can read 0 on one side and roundoff on the other)
        scale = [max(abs(o[0][c]) for o in ref.records) for c in range(-ncol, 0)]
        for (st, _, _), (ost, _, _) in zip(rows, ref.records):
            assert len(st) == len(ost)
            for a, b, sc in zip(st[-ncol:], ost[-ncol:], scale):
                assert abs(a - b) <= max(1e-10, 1e3 * tol) * max(abs(b), 1e-30) or \
                    abs(a - b) <= 1e3 * tol * sc, (a, b)
        assert s.header(_lib.FFS_TABLE_STATS).endswith(",hp,vp" if kind == "a" else ",hp,vp,tp")


@pytest.mark.parametrize("kind", ["rf", "cnrf", "a"])
def test_plane_forcing_graph_replay_and_resume_bitwise(kind):
    print_case(f"plane forcing {kind} graph replay and checkpoint")
    _, a, _ = make(kind, target=kind != "cnrf")
    _, b, _ = make(kind, target=kind != "cnrf")
    with a, b:
        a.set_graph(False)
        a.initialize()
        b.initialize()
        a.advance(30)
        b.advance(16)
        blob = b.checkpoint()
        d = ffs.Simulation(b.params)
        d.restore(blob)
        d.advance(14)
        b.advance(14)
        for s in (b, d):
            for x, y in zip(s.velocity_hat(), a.velocity_hat()):
                assert np.array_equal(x, y)
            for what in (_lib.FFS_FSTATE_R, _lib.FFS_FSTATE_EM) + ((_lib.FFS_FSTATE_N,) if kind == "cnrf" else ()):
                assert np.array_equal(s.forcing_state(what), a.forcing_state(what))
        d.close()


def test_run_simulation_writes_the_forcing_files(tmp_path):
    """writeoutput's forcing and density files (output.jl:1-67) from run_simulation in a directory:
    R.N and N.N hold the state of the forcing call before iteration N, Eh.N its spectrum over every
    shell next to the shells' mean |k_h|, rho.N the density."""
    import os
    from ffs import fields
    from ffs.simulation import jl_float

    print_case("run_simulation with nRfForcing and density: forcing state files")
    d = str(tmp_path)
    n = 32
    entries = {"nx": n, "ny": n, "nz": n, "xDomainSize": "1.0", "yDomainSize": "1.0", "zDomainSize": "1.0",
               "kinematicViscosity": 2e-3, "velocityTimeStep": "Adams_Bashforth3rdO", "variableTimestepFlag": "true",
               "cfl": 0.4, "dtStat": 2, "dtSpec": 1000, "writeTime": 3, "nt": 7, "forcing": "nRfForcing",
               "TF": 1.0, "alphac": 0.5, "kf": 2.5, "densityStratification": "true", "Pr": 0.7,
               "densityGradient": 1.0, "zAcceleration": -1.0, "referenceDensity": 1.0}
    with open(os.path.join(d, "global"), "w") as f:
        for k, v in entries.items():
            f.write(f"{k} {v}\n")
    u0 = random_spectrum(n, seed=12, k0=3.0)
    padded = 2 * (n // 2 + 1)
    for c in range(3):
        a = np.zeros((padded, n, n))
        a[:n] = u0[c]
        fields.write_field(os.path.join(d, f"u{c + 1}.0"), a)
    r = np.zeros((padded, n, n))
    r[:n] = random_scalar(n, n, n, seed=13)
    fields.write_field(os.path.join(d, "rho.0"), r)
    with ffs.run_simulation(d) as s:
        kh = s.forcing_state(_lib.FFS_FSTATE_KH)
        for it in (3, 6):
            for name in ("R", "N", "Eh", "rho", "u1"):
                assert os.path.exists(os.path.join(d, f"{name}.{it}")), (name, it)
            R = np.loadtxt(os.path.join(d, f"R.{it}"))
            N = np.loadtxt(os.path.join(d, f"N.{it}"))
            assert R.shape == N.shape == kh.shape
            eh = [line.split("\t") for line in open(os.path.join(d, f"Eh.{it}")).read().splitlines()]
            assert [a for a, _ in eh] == [jl_float(x) for x in kh]   # NaN for an empty shell, as Julia prints it
            assert all(float(b) >= 0 for _, b in eh)
        # the closing writeoutput at iteration 7 (7 % writeTime != 0) writes the state files too
        assert s.iteration == 7 and os.path.exists(os.path.join(d, "N.7"))
    # the N.6 file is the state after the forcing call at iteration 5: rerun to 6 and compare
    with ffs.Simulation(ffs.config.parameters(ffs.config.readglobal(os.path.join(d, "global")))) as t:
        for c in range(3):
            t.set_field(c, fields.read_field(os.path.join(d, f"u{c + 1}.0"), n, n, n))
        t.set_field(_lib.FFS_FIELD_RHO, fields.read_field(os.path.join(d, "rho.0"), n, n, n))
        t.initialize()
        done = 0
        while done < 6:
            k, _ = t.advance(6 - done)
            done += k
        assert t.iteration == 6
        assert np.array_equal(np.loadtxt(os.path.join(d, "N.6")), t.forcing_state(_lib.FFS_FSTATE_N))
        rho6 = t.get_field(_lib.FFS_FIELD_RHO, _lib.FFS_SPACE_REAL)
        assert np.array_equal(fields.read_field(os.path.join(d, "rho.6"), n, n, n), rho6)
