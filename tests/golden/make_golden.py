"""Writes the golden fixtures of tests/golden from the CPU oracle.

Taylor-Green vortex at 64^3 (nu = 0.0333333, 2/3 dealiasing, no forcing,
no LES), Stats.txt rows every 20 steps:

  taylor_green_64_ab3_dt0.05.csv  Adams_Bashforth3rdO, dt = 0.05, steps 0-20.
      This is the configuration as specified; explicit AB3 is unstable for
      nu |k|^2 dt > ~0.545 and the dealiased corner mode has nu |k|^2 dt = 2.2,
      so the run diverges after step ~30 (oracle and GPU alike).
  taylor_green_64_ab3_dt0.01.csv  Adams_Bashforth3rdO, dt = 0.01, 500 steps (stable).
  taylor_green_64_etd3_dt0.05.csv ETD3rdO (the reference's default integrator),
      dt = 0.05, 500 steps (stable: the viscous term is integrated exactly).

Usage: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from oracle.ffs_oracle import Grid, OracleSim, taylor_green  # noqa: E402

HEADER = "iteration,time,u1,u2,u3,u1p2,u2p2,u3p2,kh,kv,kaniso,k,dissh,dissv,diss,nlh,nlv,pressh,pressv,tnlh,tnlv,tdiss"
CASES = {
    "taylor_green_64_ab3_dt0.05": dict(integrator="AB3", dt=0.05, nsteps=21),
    "taylor_green_64_ab3_dt0.01": dict(integrator="AB3", dt=0.01, nsteps=500),
    "taylor_green_64_etd3_dt0.05": dict(integrator="ETD3", dt=0.05, nsteps=500),
}


def taylor_green_64(integrator, dt, nsteps, dt_stat=20):
    g = Grid(64, 64, 64)
    s = OracleSim(g, 0.0333333, dt=dt, dt_stat=dt_stat, integrator=integrator)
    s.set_velocity(taylor_green(g))
    s.initialize()
    s.advance_in_time(nsteps)
    return np.array([r[0] for r in s.records])


if __name__ == "__main__":
    for name, kw in CASES.items():
        rows = taylor_green_64(**kw)
        np.savetxt(os.path.join(HERE, name + ".csv"), rows, delimiter=",", header=HEADER, comments="", fmt="%.17g")
        print(f"{name}: {len(rows)} rows, final k = {rows[-1][11]:.6g}")
