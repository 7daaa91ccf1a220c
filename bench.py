"""Benchmark: forced isotropic turbulence DNS, 256^3, FP64, AB3 (CFL 0.5),
constant-power shell forcing (|k| <= 2.5, P = 0.1), nu from k_max eta = 1.5.

One step = advance_in_time! (models.jl:3-17) on one GPU: 9 component
transforms fused into 5 pencil kernels, the AB3 update, device-side variable
dt and forcing, stats every `--dt-stat` steps.  Inputs are resident in HBM
before the timed region starts.

Prints ONE JSON line (rank 0) with the throughput in N^3*steps/s, the HBM
roofline of the dominant kernel (HIP events on the solver's stream) and the
CPU baseline (the FP64 oracle on the host cores, bounded sample).
"""
import argparse
import json
import math
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "fluidflowsimulation.jl_amd"))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
METRIC = "N³·steps/sec at 256³ (FP64 plain vs FP32 batched); ensemble p99 step latency"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def nu_for_kmax_eta(n, eps=0.1, kmax_eta=1.5):
    """nu with k_max eta = 1.5, k_max = n/3, eta = (nu^3/eps)^(1/4)."""
    eta = kmax_eta / (n / 3.0)
    return (eta**4 * eps) ** (1.0 / 3.0)


def kernel_bytes(g, name, elem, with_ke_frac):
    """Algorithmic HBM bytes of one launch of each step kernel (minimum traffic
    of the pruned pencil layout: every array read once and written once).
    g = (nrx, ny, nz, mx, my, mz); elem = bytes per complex value."""
    nrx, ny, nz, mx, my, mz = g
    S = mx * my * mz            # retained modes
    S1 = mx * my * nz           # after/before the z pass
    S2 = mx * ny * nz           # after/before the y pass
    nfw = 3 + with_ke_frac      # forward fields (u x w, plus 1/2|u|^2 on stats steps)
    return {
        "inv_z_curl": (3 * S + 6 * S1) * elem,
        "inv_y": (6 * S1 + 6 * S2) * elem,
        "x_nl": (6 * S2 + nfw * S2) * elem,
        "fwd_y": (nfw * S2 + nfw * S1) * elem,
        "fwd_z": (nfw * S1 + nfw * S) * elem,
        "update": (12 * S + 6 * S) * elem,
    }.get(name)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--dt-stat", type=int, default=100)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--fp32", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first time a plain 1 GiB device copy (HBM calibration for the traffic counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(runs shard as independent jobs; see the design notes)", file=sys.stderr)
        sys.exit(2)
    raise SystemExit(bench(args))


def retained(n, cutoff=2 / 3):
    """(mx, my, mz) retained modes per direction for an n^3 grid (Globals.jl:132-169)."""
    kx = [i for i in range(n // 2 + 1)]
    cx2 = (cutoff * (n // 2)) ** 2
    mx = sum(1 for k in kx if not k * k > cx2)
    ky = [i if i < n // 2 else i - n for i in range(n)]
    cy2 = (cutoff * max(abs(k) for k in ky)) ** 2
    my = sum(1 for k in ky if not k * k > cy2)
    return mx, my, my


def reference_ranges(n, cutoff=2 / 3):
    """|XRANGE|, |YRANGE|, |ZRANGE| of the reference (3 extra rows per signed axis)."""
    mx, my, _ = retained(n, cutoff)
    f = mx  # first dealiased index (0-based) along y equals mx for cubic grids
    return mx, 2 * f + 2, 2 * f + 2


def copy_probe():
    """Plain device-to-device copy of 1 GiB (torch's copy kernel): reads 1 GiB and
    writes 1 GiB; its FETCH_SIZE / WRITE_SIZE calibrate the traffic counters and
    its rate is the measured copy bandwidth of the box."""
    import torch

    n = (1 << 30) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 10
    gbs = 2 * (1 << 30) / (ms * 1e-3) / 1e9
    log(f"copy probe: 1 GiB copy {ms:.3f} ms -> {gbs:.0f} GB/s (read+write)")
    del a, b
    torch.cuda.empty_cache()
    return gbs


def bench(args):
    import numpy as np

    copy_gbs = copy_probe() if args.copy_probe else None

    import ffs
    from ffs.initial import random_spectrum

    n = args.n
    nu = nu_for_kmax_eta(n)
    prec = "fp32" if args.fp32 else "fp64"
    b = 4 if args.fp32 else 8
    p = ffs.Params(nx=n, ny=n, nz=n, nu=nu, dt=0.0, integrator="Adams_Bashforth3rdO", variable_dt=True, cfl=0.5,
                   dt_stat=args.dt_stat, dt_spec=1 << 30, write_time=1 << 30, precision=prec,
                   forcing="shellForcing", forcing_kf=2.5, forcing_power=0.1)
    log(f"workload n={n} nu={nu:.6e} precision={prec}")
    t0 = time.time()
    u0 = random_spectrum(n, seed=args.seed)
    log(f"initial field generated in {time.time() - t0:.1f}s")
    s = ffs.Simulation(p)
    s.set_velocity(*u0)
    s.initialize()
    s.synchronize()
    log(f"initialized; device bytes {s.device_bytes() / 2**30:.2f} GiB")
    t0 = time.time()
    s.advance_in_time(args.warmup)
    s.synchronize()
    log(f"warmup {args.warmup} steps in {time.time() - t0:.2f}s")
    s.pop_stats()
    # timed region: K steps bracketed by synchronisation; per-kernel HIP events on the solver's stream
    s.profile(True)
    s.profile_reset()
    s.synchronize()
    t0 = time.perf_counter()
    s.advance_in_time(args.steps)
    s.synchronize()
    t1 = time.perf_counter()
    prof = s.profile_get()
    s.profile(False)
    rows = s.pop_stats()
    elapsed = t1 - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = n**3 * args.steps / elapsed
    log(f"timed {args.steps} steps: {ms_per_step:.3f} ms/step, {value:.4e} N^3*steps/s, dt={s.dt:.4e}")
    for r in rows[-2:]:
        log(f"stats it={int(r[0][0])} t={r[0][1]:.4f} k={r[0][11]:.6f} diss={r[0][14]:.6f}")
    finite = all(np.all(np.isfinite(r[0][2:22])) for r in rows)

    mx, my, mz = retained(n)
    g = (n, n, n, mx, my, mz)
    ke_frac = (args.steps // args.dt_stat) / args.steps if args.dt_stat > 0 else 0.0
    kernels = {}
    for name, (ms, cnt) in prof.items():
        avg_us = 1e3 * ms / cnt
        kb = kernel_bytes(g, name, 2 * b, ke_frac)
        kernels[name] = {"launches": cnt, "avg_us": round(avg_us, 3), "share": round(ms / (1e3 * elapsed), 4)}
        if kb:
            kernels[name]["alg_bytes"] = int(kb)
            kernels[name]["GBs"] = round(kb / (avg_us * 1e-6) / 1e9, 1)
    dom = max((k for k in kernels if "alg_bytes" in kernels[k]), key=lambda k: kernels[k]["avg_us"] * kernels[k]["launches"])
    d = kernels[dom]
    achieved = d["GBs"]
    # survey byte model for the whole step: 54 F_b + 15 R_b
    F = 2 * (n // 2 + 1) * n * n * b
    xr, yr, zr = reference_ranges(n)
    R = 2 * b * xr * yr * zr
    step_bytes = 54 * F + 15 * R
    step_eff = step_bytes / (ms_per_step * 1e-3) / 1e9
    step_actual = sum(kernel_bytes(g, k, 2 * b, ke_frac) for k in ("inv_z_curl", "inv_y", "x_nl", "fwd_y", "fwd_z",
                                                                  "update"))
    log(f"dominant kernel {dom}: {d['avg_us']:.1f} us, {achieved:.0f} GB/s; step model {step_eff:.0f} GB/s effective")
    for k, v in sorted(kernels.items(), key=lambda kv: -kv[1]["avg_us"] * kv[1]["launches"]):
        log(f"  {k:12s} {v}")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "N^3*steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.fp32 else "f64",
        "data": "synthetic: seeded random-phase solenoidal field, E(k)~k^4 exp(-2(k/4)^2), energy 0.5",
        "config": {
            "workload": f"forced isotropic turbulence DNS {n}^3 {prec.upper()} AB3 CFL 0.5, shell forcing |k|<=2.5 "
                        f"P=0.1, nu={nu:.4e} (k_max eta=1.5), stats every {args.dt_stat} steps",
            "n": n, "precision": prec, "integrator": "Adams_Bashforth3rdO", "cfl": 0.5, "dt_stat": args.dt_stat,
            "final_dt": s.dt, "finite": bool(finite),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "alg_bytes_per_launch": d["alg_bytes"],
            "avg_us": d["avg_us"],
        },
        "step_roofline": {
            "model_bytes_per_step": step_bytes,
            "effective_GBs": step_eff,
            "effective_frac": step_eff / HBM_PEAK_GBS,
            "pruned_bytes_per_step": step_actual,
            "pruned_GBs": step_actual / (ms_per_step * 1e-3) / 1e9,
        },
        "kernels": kernels,
        "copy_probe_GBs": copy_gbs,
    }
    s.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, nu, args)
    print(json.dumps(out), flush=True)
    return 0


def cpu_baseline(n, nu, args):
    """The FP64 oracle (numpy + multi-threaded scipy.fft) on the host cores:
    initialize + `cpu_steps` steps of the same workload."""
    from ffs.initial import random_spectrum
    from oracle.ffs_oracle import Grid, OracleSim, workers

    cores = workers()
    log(f"cpu baseline: oracle, {args.cpu_steps} steps at {n}^3 on {cores} threads")
    g = Grid(n, n, n)
    o = OracleSim(g, nu, dt=0.0, variable_dt=True, cfl=0.5, dt_stat=args.dt_stat, forcing_kf=2.5,
                  forcing_power=0.1)
    o.set_velocity(random_spectrum(n, seed=args.seed))
    o.initialize()
    t0 = time.perf_counter()
    o.advance_in_time(args.cpu_steps)
    el = time.perf_counter() - t0
    v = n**3 * args.cpu_steps / el
    log(f"cpu baseline: {el / args.cpu_steps:.2f} s/step, {v:.4e} N^3*steps/s")
    return {"value": v, "unit": "N^3*steps/s", "cores": cores, "kind": "port",
            "sample": f"{args.cpu_steps} steps of the same {n}^3 FP64 workload after initialize (numpy/scipy.fft "
                      f"oracle, {el:.1f}s)",
            "s_per_step": el / args.cpu_steps}


if __name__ == "__main__":
    main()
