"""The spectra files of write_spec / write_spec_forcing on the GPU against the
oracle (FP64: every bin within 1e-11 of its array's largest value; FP32 within
2e-5), for a plain variable-dt run, LES with density, and the rfForcing
forcing spectrum; and run_simulation end to end in a directory: the Stats,
Scales, Numbers rows, the u?.N dumps and the spectra files it writes, against
the oracle run on the same `global` file and initial fields."""
import os

import numpy as np
import pytest

import ffs
from ffs import _lib, fields
from ffs.initial import random_scalar, random_spectrum
from oracle.ffs_oracle import Grid, OracleSim, taylor_green
from oracle.plane_forcing import PlaneForcing

from conftest import print_case

pytestmark = pytest.mark.gpu

N = 32


def close(a, b, tol, floor=1e-300):
    """|a - b| within tol of b's largest value (or of `floor`, for arrays that are roundoff)"""
    scale = max(np.abs(b).max(), floor)
    return np.abs(a - b).max() <= tol * scale


CASES = {
    "ab3 variable dt": dict(kw=dict(), oracle=dict()),
    "etd3 les + density": dict(kw=dict(integrator="ETD3rdO", variable_dt=False, dt=2e-3, les_model="Smagorinsky",
                                       smagorinsky_c=0.16, density=True, pr=0.7, density_gradient=1.0,
                                       z_acceleration=-1.0, reference_density=1.0),
                               oracle=dict(integrator="ETD3", variable_dt=False, dt=2e-3, les_c=0.16,
                                           density={"Pr": 0.7, "gradient": 1.0, "g": -1.0})),
    "ab3 rfForcing": dict(kw=dict(forcing="rfForcing", forcing_kf=2.5, forcing_tf=1.0, forcing_alpha=0.5),
                          oracle=dict(), plane=True),
    "ab3 fp32": dict(kw=dict(precision="fp32"), oracle=dict()),
}


@pytest.mark.parametrize("name", list(CASES))
def test_spectra_records_match_oracle(name):
    print_case(f"spectra {name}")
    case = CASES[name]
    g = Grid(N, N, N)
    u0 = random_spectrum(N, seed=21, k0=3.0)
    okw = dict(dt=0.0, variable_dt=True, cfl=0.4)
    okw.update(case["oracle"])
    pf = PlaneForcing(g, "rf", kf=2.5, tf=1.0, alphac=0.5) if case.get("plane") else None
    ref = OracleSim(g, 2e-3, dt_stat=1 << 30, dt_spec=3, plane_forcing=pf, **okw)
    ref.set_velocity(u0)
    p = ffs.Params(nx=N, ny=N, nz=N, nu=2e-3, integrator="Adams_Bashforth3rdO", variable_dt=True, cfl=0.4,
                   dt_stat=1 << 30, dt_spec=3, write_time=1 << 30)
    for k, v in case["kw"].items():
        setattr(p, k, v)
    dens = p.density
    with ffs.Simulation(p) as s:
        s.set_velocity(*u0)
        if pf is not None:
            # a prescribed target above the initial spectrum (with the initial field's own as target, R sits at
            # its clamp at 0, where the force sqrt(R / Ef) dt magnifies roundoff)
            em = 3 * pf.u1u2_spectrum(ref.uh)
            pf.Em = em.copy()
            pf.init = False
            s.set_forcing_state(_lib.FFS_FSTATE_EM, em)
        if dens:
            rho0 = random_scalar(N, N, N, seed=22)
            s.set_density(rho0)
            ref.set_density(rho0)
        ref.initialize()
        s.initialize()
        ref.advance_in_time(8)
        s.advance(8)
        recs = s.pop_spectra()
        assert [r[0] for r in recs] == [r[0] for r in ref.spectra] == [0, 3, 6]
        tol = 2e-5 if p.precision == "fp32" else 1e-11
        kh, krz = s.spectra_axes()
        assert np.array_equal(kh, g.KH) and np.array_equal(krz, g.KRZ)
        for (it, got), (_, want) in zip(recs, ref.spectra):
            assert set(got) == set(want), (set(got) ^ set(want))
            for q in want:
                for b in (0, 1):
                    assert close(got[q][b], want[q][b], tol), (it, q, b, np.abs(got[q][b] - want[q][b]).max(),
                                                              np.abs(want[q][b]).max())
        assert s.pop_spectra() == []


def write_global(path, entries):
    with open(path, "w") as f:
        for k, v in entries.items():
            f.write(f"{k} {v}\n")


def test_run_simulation_writes_the_reference_files(tmp_path):
    print_case("run_simulation in a directory: Taylor-Green 32^3 AB3")
    d = str(tmp_path)
    n, nu, dt = 32, 0.0333333, 0.02
    write_global(os.path.join(d, "global"), {
        "nx": n, "ny": n, "nz": n, "xDomainSize": "1.0", "yDomainSize": "1.0", "zDomainSize": "1.0",
        "kinematicViscosity": nu, "dt": dt, "velocityTimeStep": "Adams_Bashforth3rdO",
        "variableTimestepFlag": "false", "dtStat": 2, "dtSpec": 3, "writeTime": 3, "nt": 7})
    g = Grid(n, n, n)
    u0 = taylor_green(g)
    padded = 2 * (n // 2 + 1)
    for c in range(3):
        a = np.zeros((padded, n, n))
        a[:n] = u0[c]
        fields.write_field(os.path.join(d, f"u{c + 1}.0"), a)
    ffs.run_simulation(d).close()
    ref = OracleSim(g, nu, dt=dt, dt_stat=2, dt_spec=3)
    ref.set_velocity(u0)
    ref.initialize()
    ref.advance_in_time(7)
    # the stats rows at 0, 2, 4, 6 and the final row at 7 (run_simulation's closing writestats)
    rows = np.loadtxt(os.path.join(d, "Stats.txt"), delimiter=",", skiprows=1)
    assert [int(r[0]) for r in rows] == [0, 2, 4, 6, 7]
    for r, o in zip(rows, ref.records):
        assert abs(r[11] - o[0][11]) <= 1e-10 * abs(o[0][11])   # k
        assert abs(r[14] - o[0][14]) <= 1e-10 * abs(o[0][14])   # diss
    # field dumps at the output iterations and at the end
    for it in (3, 6, 7):
        for c in range(3):
            assert os.path.exists(os.path.join(d, f"u{c + 1}.{it}")), (c, it)
    # spectra files at 0, 3, 6, written as writedlm rows "k<TAB>value"
    for (it, q) in ref.spectra:
        assert it in (0, 3, 6)
        # Taylor-Green has u3 = 0: its vertical spectra are roundoff, compared on the record's scale
        floor = max(np.abs(q[k][b]).max() for k in q for b in (0, 1))
        for stem in ("u_h", "u_v", "vis_h", "vis_v", "nl_h", "nl_v", "press_h", "press_v"):
            for b, ext in ((0, "specH"), (1, "specV")):
                tab = np.loadtxt(os.path.join(d, f"{stem}.{ext}.{it}"), delimiter="\t")
                want = q[stem][b]
                axis = g.KH if b == 0 else g.KRZ
                assert np.array_equal(tab[:, 0], axis)
                assert close(tab[:, 1], want, 1e-11, floor), (stem, ext, it)
        for n_ in ("u", "vis", "nl", "press"):
            tab = np.loadtxt(os.path.join(d, f"{n_}.specH.{it}"), delimiter="\t")
            assert close(tab[:, 1], q[n_ + "_h"][0] + q[n_ + "_v"][0], 1e-11, floor)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-5)])
def test_les_stress_fields(prec, tol):
    """T11 ... T23: the real-space Smagorinsky stress of the current velocity, and the stored
    transform back in real space, against the oracle's."""
    print_case(f"LES stress fields {prec}")
    g = Grid(N, N, N)
    u0 = random_spectrum(N, seed=61, k0=3.0)
    ref = OracleSim(g, 2e-3, dt=1e-3, les_c=0.16)
    ref.set_velocity(u0)
    p = ffs.Params(nx=N, ny=N, nz=N, nu=2e-3, integrator="Adams_Bashforth3rdO", variable_dt=False, dt=1e-3,
                   dt_stat=1 << 30, dt_spec=1 << 30, write_time=1 << 30, les_model="Smagorinsky",
                   smagorinsky_c=0.16, precision=prec)
    with ffs.Simulation(p) as s:
        s.set_velocity(*u0)
        ref.initialize()
        s.initialize()
        ref.advance_in_time(4)
        s.advance(4)
        ref.calculate_rhs(collect=False)   # the stress of the current velocity (realspace!)
        names = (_lib.FFS_FIELD_T11, _lib.FFS_FIELD_T12, _lib.FFS_FIELD_T13, _lib.FFS_FIELD_T22,
                 _lib.FFS_FIELD_T23)
        for c, f in enumerate(names):
            got = s.get_field(f, _lib.FFS_SPACE_REAL)[:N]
            want = ref.taur[c]
            assert np.abs(got - want).max() <= tol * np.abs(want).max(), (c, np.abs(got - want).max())
        # the stored transform is the last nonlinear pass's (before the 4th step): compare after a pass
        s.writestats_now()
        for c, f in enumerate(names):
            got = s.get_field(f, _lib.FFS_SPACE_FOURIER_TO_REAL)[:N]
            want = g.irfft(ref.tauh[c])
            assert np.abs(got - want).max() <= max(tol, 1e-11) * np.abs(want).max(), c
