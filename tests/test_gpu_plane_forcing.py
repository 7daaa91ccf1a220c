"""The reference's kz = 0-plane forcing schemes on the GPU against the oracle
(rfForcing with a prescribed target spectrum, QfForcing with and without the
LES filter), plus graph replay and exact resume with the oscillator state."""
import numpy as np
import pytest

import ffs
from ffs import _lib
from ffs.initial import random_spectrum
from oracle.ffs_oracle import Grid, OracleSim
from oracle.plane_forcing import PlaneForcing

from conftest import print_case

pytestmark = pytest.mark.gpu

N = 32


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def make(kind, les=False, steps=15):
    g = Grid(N, N, N)
    u0 = random_spectrum(N, seed=12, k0=3.0)
    delta = (g.LX * 2 * np.pi / g.NRX) / g.cutoff if les else None
    if kind == "rf":
        pf = PlaneForcing(g, "rf", kf=2.5, tf=1.0, alphac=0.5)
        kw = dict(forcing="rfForcing", forcing_kf=2.5, forcing_tf=1.0, forcing_alpha=0.5)
    else:
        pf = PlaneForcing(g, "qf", kf=3.0, P=0.1, C=1.0, Kf_alv=2.0, les_delta=delta)
        kw = dict(forcing="QfForcing", forcing_kf=3.0, alvelius_p=0.1, alvelius_c=1.0, alvelius_kf=2.0)
    ref = OracleSim(g, 2e-3, dt=0.0, variable_dt=True, cfl=0.4, dt_stat=3, plane_forcing=pf,
                    les_c=0.16 if les else None)
    ref.set_velocity(u0)
    p = ffs.Params(nx=N, ny=N, nz=N, nu=2e-3, integrator="Adams_Bashforth3rdO", variable_dt=True, cfl=0.4,
                   dt_stat=3, dt_spec=1 << 30, write_time=1 << 30,
                   les_model="Smagorinsky" if les else None, smagorinsky_c=0.16, **kw)
    s = ffs.Simulation(p)
    s.set_velocity(*u0)
    if kind == "rf":
        # a prescribed target above the initial spectrum, so the oscillator drives energy in
        em = 3 * pf.u1u2_spectrum(ref.uh)
        pf.Em = em.copy()
        pf.init = False
        s.set_forcing_state(_lib.FFS_FSTATE_EM, em)
    return ref, s, pf


@pytest.mark.parametrize("kind,les", [("rf", False), ("qf", False), ("qf", True)])
def test_plane_forcing_matches_oracle(kind, les):
    print_case(f"plane forcing {kind} les={les}")
    ref, s, pf = make(kind, les)
    with s:
        ref.initialize()
        s.initialize()
        kh = s.forcing_state(_lib.FFS_FSTATE_KH)
        assert np.allclose(kh, pf.kh, rtol=1e-15, atol=0, equal_nan=True)
        ref.advance_in_time(15)
        s.advance(15)
        err = max(rel(s.velocity_hat()[c], ref.uh[c]) for c in range(3))
        assert err < 1e-12, err
        R = s.forcing_state(_lib.FFS_FSTATE_R)
        assert np.allclose(R, pf.R, rtol=1e-12, atol=1e-12 * max(np.abs(pf.R).max(), 1e-300)), (R, pf.R)
        rows = s.pop_stats()
        assert any(r[0][-3] > 0 for r in rows)  # the forcing did work on the flow
        assert len(rows) == len(ref.records) == 5
        for (st, _, _), (ost, _, _) in zip(rows, ref.records):
            assert len(st) == len(ost)
            hp, vp, tp = st[-3:]
            ohp, ovp, otp = ost[-3:]
            for a, b in ((hp, ohp), (vp, ovp), (tp, otp)):
                assert abs(a - b) <= 1e-10 * max(abs(b), 1e-30), (a, b)
        assert s.header(_lib.FFS_TABLE_STATS).endswith(",hp,vp,tp")


def test_plane_forcing_graph_replay_and_resume_bitwise():
    print_case("plane forcing graph replay and checkpoint")
    _, a, _ = make("rf")
    _, b, _ = make("rf")
    _, c, _ = make("rf")
    with a, b, c:
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
            assert np.array_equal(s.forcing_state(_lib.FFS_FSTATE_R), a.forcing_state(_lib.FFS_FSTATE_R))
        d.close()
