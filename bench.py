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
        # the forward z pass of u x w fused with the update: no rhs round trip
        "fwd_z_update": (3 * S1 + 9 * S + 6 * S) * elem,
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
    ap.add_argument("--legs", default="fp32,ensemble,serve,strat,features",
                    help="extra measurements after the main line: fp32, ensemble, serve, strat, features "
                         "(comma separated)")
    ap.add_argument("--feature-steps", type=int, default=100)
    ap.add_argument("--strat-n", type=int, default=512)
    ap.add_argument("--strat-steps", type=int, default=500)
    ap.add_argument("--fp32-batch", type=int, default=2)
    ap.add_argument("--fp32-err-steps", type=int, default=20)
    ap.add_argument("--ensemble-runs", type=int, default=64)
    ap.add_argument("--ensemble-steps", type=int, default=1000)
    ap.add_argument("--latency-samples", type=int, default=50)
    ap.add_argument("--serve-requests", type=int, default=40)
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--serve-max-runs", type=int, default=64)
    ap.add_argument("--serve-chunk", type=int, default=25)
    ap.add_argument("--no-fusion", action="store_true",
                    help="run the forward z pass and the update of plain steps as two kernels")
    ap.add_argument("--copy-probe", action="store_true", default=True,
                    help="first time a plain 1 GiB device copy (the box's measured copy bandwidth, beside the 8 TB/s "
                         "peak, and the calibration of the traffic counters); on by default")
    ap.add_argument("--no-copy-probe", dest="copy_probe", action="store_false")
    ap.add_argument("--plan", choices=["auto", "default"], default="auto",
                    help="kernel plans: tuned per workload on this device (cached, see ffs/tune.py) or the defaults")
    ap.add_argument("--leg-child", default=None, help=argparse.SUPPRESS)  # run one leg, print its JSON (internal)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(runs shard as independent jobs; see the design notes)", file=sys.stderr)
        sys.exit(2)
    global PLAN_MODE
    PLAN_MODE = "auto" if args.plan == "auto" else None
    if args.leg_child == "serve":  # the serving leg in a process of its own (see leg_in_child)
        res = leg_serve(args)
        res["_plans"] = PLANS
        print(json.dumps(res), flush=True)
        return 0
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


def measured_traffic(kernel, prec, n, launches_per_step=1):
    """HBM bytes per launch of `kernel`: its bytes per step from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (profiles/step_kernels_<n>_<prec>.md) over this run's launches per
    step (z slabs), or None when none is recorded."""
    path = os.path.join(REPO, "profiles", f"step_kernels_{n}_{prec}.md")
    if not os.path.exists(path):
        return None
    key = f"traffic_bytes_per_step {kernel} {prec} {n} ="
    for line in open(path):
        if line.startswith(key):
            return int(float(line.split("=")[1]) / launches_per_step)
    return None


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


# the kernel plan of every run the bench creates: "auto" (ffs.tune: timed on a twin before the run is
# created, then cached) or None (the library's defaults); PLANS records what each plan key ran with
PLAN_MODE = "auto"
PLANS = {}


def record_plan(s):
    from ffs import tune

    PLANS[tune.plan_key(s.params)] = s.plan


def make_forced(n, nu, prec, dt_stat, seed, les=False):
    """One run of the forced workload, its initial field generated on the device from `seed`."""
    import ffs

    p = ffs.Params(nx=n, ny=n, nz=n, nu=nu, dt=0.0, integrator="Adams_Bashforth3rdO", variable_dt=True, cfl=0.5,
                   dt_stat=dt_stat, dt_spec=1 << 30, write_time=1 << 30, precision=prec,
                   forcing="shellForcing", forcing_kf=2.5, forcing_power=0.1,
                   les_model="Smagorinsky" if les else None, smagorinsky_c=0.16)
    s = ffs.Simulation(p, plan=PLAN_MODE)
    record_plan(s)
    s.seed_velocity(seed, k0=4.0, energy=0.5)
    s.initialize()
    return s


def timed_steps(s, steps):
    """Wall time of `steps` steps on the default (graph-replay) path, bracketed by synchronisation."""
    s.synchronize()
    t0 = time.perf_counter()
    s.advance_in_time(steps)
    s.synchronize()
    return time.perf_counter() - t0


def profiled_steps(s, steps):
    """Per-kernel device time (HIP events on the solver's stream around every launch) over `steps` steps."""
    s.profile(True)
    s.profile_reset()
    s.synchronize()
    t0 = time.perf_counter()
    s.advance_in_time(steps)
    s.synchronize()
    el = time.perf_counter() - t0
    prof = s.profile_get()
    s.profile(False)
    return prof, el


def bench(args):
    import numpy as np

    copy_gbs = copy_probe() if args.copy_probe else None

    n = args.n
    nu = nu_for_kmax_eta(n)
    prec = "fp32" if args.fp32 else "fp64"
    b = 4 if args.fp32 else 8
    log(f"workload n={n} nu={nu:.6e} precision={prec}")
    s = make_forced(n, nu, prec, args.dt_stat, args.seed)
    main_plan = s.plan
    if args.no_fusion:
        s.set_fusion(False)
    s.synchronize()
    log(f"initialized; device bytes {s.device_bytes() / 2**30:.2f} GiB")
    t0 = time.time()
    s.advance_in_time(args.warmup)
    s.synchronize()
    log(f"warmup {args.warmup} steps in {time.time() - t0:.2f}s")
    s.pop_stats()
    # timed region: K steps of the product path (steady steps replayed from HIP graphs)
    elapsed = timed_steps(s, args.steps)
    rows = s.pop_stats()
    ms_per_step = 1e3 * elapsed / args.steps
    value = n**3 * args.steps / elapsed
    log(f"timed {args.steps} steps: {ms_per_step:.3f} ms/step, {value:.4e} N^3*steps/s, dt={s.dt:.4e}")
    # kernel timing: the same K steps again with HIP events around every launch (eager launches)
    prof, el_prof = profiled_steps(s, args.steps)
    rows += s.pop_stats()
    log(f"profiled {args.steps} steps: {1e3 * el_prof / args.steps:.3f} ms/step")
    for r in rows[-2:]:
        log(f"stats it={int(r[0][0])} t={r[0][1]:.4f} k={r[0][11]:.6f} diss={r[0][14]:.6f}")
    finite = all(np.all(np.isfinite(r[0][2:22])) for r in rows)

    mx, my, mz = retained(n)
    g = (n, n, n, mx, my, mz)
    n_ke = args.steps // args.dt_stat if args.dt_stat > 0 else 0  # steps that also transform 1/2|u|^2
    kernels = {}
    for name, (ms, cnt) in prof.items():
        avg_us = 1e3 * ms / cnt
        # the share of this kernel's launches that carry 1/2|u|^2 (the separate forward z pass runs on
        # those steps only when the plain steps are fused)
        ke_frac = min(1.0, n_ke / cnt) if name == "fwd_z" else n_ke / args.steps
        kb = kernel_bytes(g, name, 2 * b, ke_frac)
        # with z slabs the inverse y, x and forward y passes launch once per slab, each on its share of planes
        slabs = max(1, round(cnt / args.steps)) if name in ("inv_y", "x_nl", "fwd_y") else 1
        if kb:
            kb /= slabs
        kernels[name] = {"launches": cnt, "avg_us": round(avg_us, 3), "share": round(ms / (1e3 * el_prof), 4)}
        if slabs > 1:
            kernels[name]["launches_per_step"] = slabs
        if kb:
            kernels[name]["alg_bytes"] = int(kb)
            kernels[name]["GBs"] = round(kb / (avg_us * 1e-6) / 1e9, 1)
    dom = max((k for k in kernels if "alg_bytes" in kernels[k]),
              key=lambda k: kernels[k]["avg_us"] * kernels[k]["launches"])
    d = kernels[dom]
    achieved = d["GBs"]
    # survey byte model for the whole step: 54 F_b + 15 R_b
    F = 2 * (n // 2 + 1) * n * n * b
    xr, yr, zr = reference_ranges(n)
    R = 2 * b * xr * yr * zr
    step_bytes = 54 * F + 15 * R
    # the pruned-layout bytes of the step as launched (fused or not), per step
    step_actual = sum(v["alg_bytes"] * v["launches"] for v in kernels.values() if "alg_bytes" in v) / args.steps
    log(f"dominant kernel {dom}: {d['avg_us']:.1f} us, {achieved:.0f} GB/s; step {step_actual / 1e9:.3f} GB at "
        f"{step_actual / (ms_per_step * 1e-3) / 1e9:.0f} GB/s ({step_actual / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS:.3f} of peak)")
    for k, v in sorted(kernels.items(), key=lambda kv: -kv[1]["avg_us"] * kv[1]["launches"]):
        log(f"  {k:12s} {v}")
    final_dt = s.dt
    s.close()

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
        "data": "synthetic: seeded random-phase solenoidal field generated on the device (SplitMix64 + Box-Muller "
                "per mode), E(k)~k^4 exp(-2(k/4)^2), energy 0.5",
        "config": {
            "workload": f"forced isotropic turbulence DNS {n}^3 {prec.upper()} AB3 CFL 0.5, shell forcing |k|<=2.5 "
                        f"P=0.1, nu={nu:.4e} (k_max eta=1.5), stats every {args.dt_stat} steps",
            "n": n, "precision": prec, "integrator": "Adams_Bashforth3rdO", "cfl": 0.5, "dt_stat": args.dt_stat,
            "final_dt": final_dt, "finite": bool(finite),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(dom, prec, n, d.get("launches_per_step", 1)),
            "traffic_source": f"profiles/step_kernels_{n}_{prec}.md (FETCH_SIZE x2.0 + WRITE_SIZE x1.0, calibrated "
                              "by a 1 GiB copy and fill)",
            "alg_bytes_per_launch": d["alg_bytes"],
            "avg_us": d["avg_us"],
            "timing": "HIP events on the solver's stream around every launch, over a second stretch of the same "
                      "K steps launched eagerly",
            "profiled_ms_per_step": 1e3 * el_prof / args.steps,
        },
        # the whole step against the HBM roofline: the bytes its kernels must move (each array of the
        # pruned pencil layout read and written once) over the timed ms/step; the survey's 54 F + 15 R
        # model of the reference's unpruned layout is kept for reference only (it is not a fraction of peak
        # for this design, which moves fewer bytes)
        "step_roofline": {
            "bytes_per_step": step_actual,
            "achieved_GBs": step_actual / (ms_per_step * 1e-3) / 1e9,
            "peak_GBs": HBM_PEAK_GBS,
            "frac": step_actual / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "copy_probe_frac": (step_actual / (ms_per_step * 1e-3) / 1e9 / copy_gbs) if copy_gbs else None,
            "survey_model_bytes_per_step": step_bytes,
        },
        "kernels": kernels,
        "copy_probe_GBs": copy_gbs,
    }
    out["config"]["plan"] = main_plan
    legs = [x for x in args.legs.split(",") if x]
    if "fp32" in legs and not args.fp32:
        out["fp32"] = leg_fp32(args, n, nu, value)
    if "ensemble" in legs:
        out["ensemble"] = leg_ensemble(args)
    if "serve" in legs:
        out["serving"] = leg_in_child("serve")
    if "strat" in legs:
        out["stratified"] = leg_strat(args)
    if "features" in legs:
        out["features"] = leg_features(args, n, nu)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, nu, args)
    from ffs import tune

    out["plans"] = {"mode": args.plan, "cache": tune.cache_path() if PLAN_MODE else None, "by_key": PLANS}
    print(json.dumps(out), flush=True)
    return 0


def spectrum_rel_err(a, b):
    """max over the four spectra of max_k |a - b| / max_k |b|."""
    import numpy as np

    err = 0.0
    for x, y in zip(a, b):
        for u, v in zip(x, y):
            err = max(err, float(np.abs(u - v).max() / max(np.abs(v).max(), 1e-300)))
    return err


def leg_fp32(args, n, nu, fp64_value):
    """FP32 bulk path at the same workload: one run, and a lockstep batch of `fp32_batch` runs
    (each its own seed); the spectrum error against the FP64 path after `fp32_err_steps` steps."""
    import ffs

    out = {}
    s = make_forced(n, nu, "fp32", 1 << 30, args.seed)
    s.advance_in_time(args.warmup)
    el = timed_steps(s, args.steps)
    out["single"] = {"value": n**3 * args.steps / el, "ms_per_step": 1e3 * el / args.steps,
                     "vs_fp64_value": n**3 * args.steps / el / fp64_value}
    s.close()
    log(f"fp32 single: {1e3 * el / args.steps:.3f} ms/step, {n**3 * args.steps / el:.4e} N^3*steps/s")
    nb = args.fp32_batch
    sims = [make_forced(n, nu, "fp32", 1 << 30, args.seed + i) for i in range(nb)]
    with ffs.Ensemble(sims) as ens:
        ens.advance_in_time(args.warmup)
        ens.synchronize()
        t0 = time.perf_counter()
        ens.advance_in_time(args.steps)
        ens.synchronize()
        el = time.perf_counter() - t0
    for x in sims:
        x.close()
    v = nb * n**3 * args.steps / el
    out["batched"] = {"runs": nb, "value": v, "ms_per_lockstep_step": 1e3 * el / args.steps,
                      "vs_fp64_value": v / fp64_value}
    log(f"fp32 batched x{nb}: {1e3 * el / args.steps:.3f} ms/lockstep step, {v:.4e} N^3*steps/s "
        f"({v / fp64_value:.2f}x FP64)")
    # precision: FP32 vs FP64 from the same field after K steps (each with its own variable dt)
    k = args.fp32_err_steps
    a = make_forced(n, nu, "fp64", 1 << 30, args.seed)
    c = make_forced(n, nu, "fp32", 1 << 30, args.seed)
    a.advance_in_time(k)
    c.advance_in_time(k)
    err = spectrum_rel_err(c.spectrum(), a.spectrum())
    a.writestats_now()
    c.writestats_now()
    ra, rc = a.pop_stats()[-1][0], c.pop_stats()[-1][0]
    out["spectrum_rel_err_vs_fp64"] = err
    out["energy_rel_err_vs_fp64"] = float(abs(rc[11] - ra[11]) / abs(ra[11]))
    out["dissipation_rel_err_vs_fp64"] = float(abs(rc[14] - ra[14]) / abs(ra[14]))
    out["err_after_steps"] = k
    log(f"fp32 vs fp64 after {k} steps: spectrum {err:.2e}, energy {out['energy_rel_err_vs_fp64']:.2e}")
    a.close()
    c.close()
    return out


def leg_ensemble(args):
    """Batched LES ensemble: `ensemble_runs` Smagorinsky (c=0.16) runs, sizes drawn uniformly from
    {32,48,64,96,128}^3, nu log-uniform in [1e-4, 1e-2], own seeded field, shell forcing, FP32,
    AB3 CFL 0.5, advanced in lockstep (one graph launch per steady step; runs of one size share one launch
    per pass).  The same runs timed as one graph branch per run are reported beside it."""
    import numpy as np

    import ffs

    rng = np.random.default_rng(args.seed)
    R = args.ensemble_runs
    sizes = [int(x) for x in rng.choice([32, 48, 64, 96, 128], size=R)]
    nus = np.exp(rng.uniform(np.log(1e-4), np.log(1e-2), size=R))
    t0 = time.time()
    sims = [make_forced(n, float(nus[i]), "fp32", 1 << 30, args.seed + 100 + i, les=True) for i, n in enumerate(sizes)]
    log(f"ensemble: {R} runs created in {time.time() - t0:.1f}s; sizes {sorted(set(sizes))}")
    pts = float(sum(n**3 for n in sizes))
    K = args.ensemble_steps
    with ffs.Ensemble(sims) as ens:
        # the same runs with one graph branch per run (no merged launches), over a shorter stretch
        ens.set_merge(False)
        ens.advance_in_time(args.warmup)
        ens.synchronize()
        kb = max(1, K // 4)
        t0 = time.perf_counter()
        ens.advance_in_time(kb)
        ens.synchronize()
        el_branches = time.perf_counter() - t0
        branch_info = ens.launch_info()
        ens.set_merge(True)
        ens.advance_in_time(args.warmup)
        ens.synchronize()
        t0 = time.perf_counter()
        ens.advance_in_time(K)
        ens.synchronize()
        el = time.perf_counter() - t0
        merged_info = ens.launch_info()
        lat = []
        for _ in range(args.latency_samples):
            a = time.perf_counter()
            ens.advance(1)
            ens.synchronize()
            lat.append(1e3 * (time.perf_counter() - a))
    finite = sum(1 for x in sims if math.isfinite(x.dt))
    fp = sum(ffs.footprint(x.params) for x in sims)
    for x in sims:
        x.close()
    out = {"runs": R, "steps": K, "value": pts * K / el, "unit": "N^3*steps/s (sum over runs)",
           "ms_per_lockstep_step": 1e3 * el / K, "step_ms_p50": float(np.percentile(lat, 50)),
           "step_ms_p99": float(np.percentile(lat, 99)), "latency_samples": len(lat), "finite_runs": finite,
           "device_bytes": fp, "size_counts": {str(n): sizes.count(n) for n in sorted(set(sizes))},
           "launches_per_step": merged_info[0] if merged_info else None,
           "merged_groups": merged_info[1] if merged_info else None,
           "merged_runs": merged_info[2] if merged_info else None,
           "branches": {"ms_per_lockstep_step": 1e3 * el_branches / kb, "steps": kb,
                        "launches_per_step": branch_info[0] if branch_info else None}}
    log(f"ensemble: {1e3 * el / K:.3f} ms/lockstep step, {pts * K / el:.4e} N^3*steps/s, "
        f"p50 {out['step_ms_p50']:.3f} ms p99 {out['step_ms_p99']:.3f} ms; launches/step {out['launches_per_step']} "
        f"in {out['merged_groups']} groups ({out['merged_runs']} runs) vs {out['branches']['launches_per_step']} "
        f"as branches at {out['branches']['ms_per_lockstep_step']:.3f} ms/step")
    return out


def leg_in_child(name):
    """Runs one auxiliary leg in a child process (the same arguments, the leg's JSON on the child's last stdout
    line), so that a crash in that leg cannot take the bench's line down; a failed child is reported as
    {"error": ...}.  The child is started as a new process (nothing is exec-replaced)."""
    import subprocess

    cmd = [sys.executable, "-u", os.path.abspath(__file__), *sys.argv[1:], "--leg-child", name]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, text=True)
    lines = (r.stdout or "").strip().splitlines()
    if r.returncode != 0 or not lines:
        log(f"{name}: the leg's process exited with status {r.returncode}")
        return {"error": f"the leg's process exited with status {r.returncode}"}
    res = json.loads(lines[-1])
    PLANS.update(res.pop("_plans", {}))
    return res


def leg_serve(args):
    """Serving: Poisson arrivals (mean `serve_rate`/s) of `serve_requests` solves, size from
    {64,128,192,256}^3 weighted 4:3:2:1, FP32:FP64 3:1, nu log-uniform in [1e-4, 1e-2], horizon
    uniform in [200, 2000] steps, Smagorinsky LES with probability 1/2, admitted under an HBM budget
    with state resident on the device; continuous batching in lanes of one size and precision that run
    side by side.  Step latency comes from device timestamps around each lane's chunks.  Each request's
    initial field is generated on the device from its own seed."""
    import numpy as np

    import ffs
    from ffs.serve import Job, Server

    rng = np.random.default_rng(args.seed + 7)
    R = args.serve_requests
    sizes = [int(x) for x in rng.choice([64, 128, 192, 256], size=R, p=[0.4, 0.3, 0.2, 0.1])]
    precs = ["fp32" if x < 0.75 else "fp64" for x in rng.uniform(size=R)]
    nus = np.exp(rng.uniform(np.log(1e-4), np.log(1e-2), size=R))
    horizons = [int(x) for x in rng.integers(200, 2001, size=R)]
    les = [bool(x < 0.5) for x in rng.uniform(size=R)]
    arrivals = np.cumsum(rng.exponential(1.0 / args.serve_rate, size=R))
    log(f"serving: {R} requests")

    def job(i):
        n = sizes[i]
        p = ffs.Params(nx=n, ny=n, nz=n, nu=float(nus[i]), integrator="Adams_Bashforth3rdO", variable_dt=True,
                       cfl=0.5, dt_stat=100, dt_spec=1 << 30, write_time=1 << 30, precision=precs[i],
                       forcing="shellForcing", forcing_kf=2.5, forcing_power=0.1,
                       les_model="Smagorinsky" if les[i] else None, smagorinsky_c=0.16)
        return Job(p, horizons[i], lambda s, k=args.seed + 500 + i: s.seed_velocity(k), name=f"req{i}",
                   keep_fields=False)

    jobs = [job(i) for i in range(R)]
    if PLAN_MODE == "auto":  # tune every plan key before the clock starts (later lookups hit the cache)
        from ffs import tune

        t0 = time.perf_counter()
        keys = {}
        for j in jobs:
            keys.setdefault(tune.plan_key(j.params), j.params)
        for k, p in keys.items():
            PLANS[k] = tune.tuned_plan(p)
        log(f"serving: {len(keys)} plan keys tuned or read from the cache in {time.perf_counter() - t0:.1f}s")
    srv = Server(budget=int(args.serve_budget_gb * 1e9), max_runs=args.serve_max_runs, chunk=args.serve_chunk,
                 plan=PLAN_MODE)
    t0 = time.perf_counter()
    done = srv.run(arrivals=[(float(arrivals[i]), jobs[i]) for i in range(R)])
    el = time.perf_counter() - t0
    pts_steps = float(sum(j.params.nx**3 * j.nsteps for j in done))
    lat = srv.latency()
    out = {"requests": R, "rate_per_s": args.serve_rate, "wall_s": el, "value": pts_steps / el,
           "unit": "N^3*steps/s (sum over requests)", **lat, "latency_clock": "device timestamps per lane chunk",
           "max_active": srv.max_active, "budget_bytes": srv.budget, "peak_bytes": srv.peak_used,
           "memory_releases": srv.releases,
           "finite": sum(1 for j in done if all(np.all(np.isfinite(r[0][2:12])) for r in j.stats))}
    log(f"serving: {R} requests in {el:.1f}s, {pts_steps / el:.4e} N^3*steps/s, step p50 "
        f"{lat.get('step_ms_p50', 0):.3f} ms p99 {lat.get('step_ms_p99', 0):.3f} ms, max active {srv.max_active}")
    return out


def leg_strat(args):
    """Stably stratified Boussinesq DNS: seeded velocity (energy 0.5, k0 4) and density (rms 0.01, k0 4),
    N = 1 (Pr 1, densityGradient 1, zAcceleration 1, referenceDensity 1), nu = 1e-4, ETD3 with a fixed dt
    from the initial CFL 0.18 (so both precisions share dt); the FP32 bulk path against the FP64 path
    after `strat_steps` steps: energy, potential energy (<rho^2>) and spectrum errors, and each one's rate."""
    import numpy as np

    import ffs
    from ffs import _lib

    n = args.strat_n
    nu = 1e-4

    def params(prec, variable):
        return ffs.Params(nx=n, ny=n, nz=n, nu=nu, integrator="ETD3rdO", variable_dt=variable, cfl=0.18,
                          dt=dt if not variable else 0.0, dt_stat=1 << 30, dt_spec=1 << 30, write_time=1 << 30,
                          precision=prec, density=True, pr=1.0, density_gradient=1.0, z_acceleration=1.0,
                          reference_density=1.0)

    def seeded(p):
        s = ffs.Simulation(p, plan=PLAN_MODE)
        record_plan(s)
        s.seed_velocity(args.seed, k0=4.0, energy=0.5)
        s.seed_density(args.seed + 1, k0=4.0, rms=0.01)
        s.initialize()
        return s

    dt = 0.0
    with seeded(params("fp64", True)) as probe:  # max|u_i| of the initial field
        umax = probe.scalar(_lib.FFS_SCALAR_UMAX)
    dt = 0.18 * (2 * math.pi / n) / umax
    out = {"n": n, "steps": args.strat_steps, "dt": dt, "integrator": "ETD3rdO", "nu": nu}
    res = {}
    for prec in ("fp64", "fp32"):
        s = seeded(params(prec, False))
        s.advance_in_time(3)
        s.synchronize()
        t0 = time.perf_counter()
        s.advance_in_time(args.strat_steps - 3)
        s.synchronize()
        el = time.perf_counter() - t0
        s.writestats_now()
        row = s.pop_stats()[-1][0]
        spec = s.spectrum()
        res[prec] = (row, spec)
        out[prec] = {"value": n**3 * (args.strat_steps - 3) / el, "ms_per_step": 1e3 * el / (args.strat_steps - 3),
                     "device_bytes": s.device_bytes(), "finite": bool(np.all(np.isfinite(row[2:30])))}
        log(f"stratified {n}^3 {prec}: {out[prec]['ms_per_step']:.3f} ms/step, {out[prec]['value']:.4e} N^3*steps/s")
        s.close()
    (r64, s64), (r32, s32) = res["fp64"], res["fp32"]
    out["energy_rel_err"] = float(abs(r32[11] - r64[11]) / abs(r64[11]))
    out["potential_energy_rel_err"] = float(abs(r32[23] - r64[23]) / abs(r64[23]))
    errs = []
    for x, y in zip(s32, s64):
        for u, v in zip(x, y):
            sel = v >= 1e-6 * v.max()
            errs.append(float(np.max(np.abs(u[sel] - v[sel]) / v[sel])))
    out["spectrum_rel_err"] = max(errs)
    out["fp32_vs_fp64_value"] = out["fp32"]["value"] / out["fp64"]["value"]
    log(f"stratified fp32 vs fp64: energy {out['energy_rel_err']:.2e}, potential energy "
        f"{out['potential_energy_rel_err']:.2e}, spectrum {out['spectrum_rel_err']:.2e}")
    return out


def leg_features(args, n, nu):
    """Step cost of the options beyond the plain AB3 step, at the main workload's grid (FP64, seeded
    field, variable dt, graph replay): ETD3, hyperviscosity, the spectral barrier, three plane forcing
    schemes, Smagorinsky, the conservation form, a passive scalar; and the device time of one spectra record (the 17 write_spec quantities and
    their marginals) at a spectra iteration, from HIP events around its launches."""
    import ffs

    base = dict(nx=n, ny=n, nz=n, nu=nu, dt=0.0, integrator="Adams_Bashforth3rdO", variable_dt=True, cfl=0.5,
                dt_stat=1 << 30, dt_spec=1 << 30, write_time=1 << 30)
    cases = {
        "ab3": {},
        "etd3": dict(integrator="ETD3rdO", cfl=0.4),
        "etd3 hyperviscosity m=2": dict(integrator="ETD3rdO", cfl=0.4, hyperviscosity="hyperViscosity",
                                        hv_nu=1e-9, hv_m=2),
        "etd3 spectral barrier": dict(integrator="ETD3rdO", cfl=0.4, hyperviscosity="spectralBarrier"),
        "ab3 rfForcing": dict(forcing="rfForcing", forcing_kf=2.5, forcing_tf=1.0, forcing_alpha=0.5),
        "ab3 nRfForcing": dict(forcing="nRfForcing", forcing_kf=2.5, forcing_tf=1.0, forcing_alpha=0.5),
        "ab3 aForcing": dict(forcing="aForcing", forcing_kf=3.0, forcing_tf=1.0, forcing_alpha=0.5),
        "ab3 smagorinsky": dict(les_model="Smagorinsky", smagorinsky_c=0.16),
        "ab3 vreman": dict(les_model="Vreman"),
        "ab3 dynamic smagorinsky": dict(les_model="dynamicSmagorinsky"),
        "ab3 conservation form": dict(equation="NS"),
        "ab3 passive scalar": dict(passive_scalar=True, scalar_pr=1.0, scalar_gradient=1.0),
    }
    out = {"n": n, "steps": args.feature_steps, "ms_per_step": {}}
    for name, kw in cases.items():
        p = ffs.Params(**dict(base, **kw))
        with ffs.Simulation(p) as s:
            s.seed_velocity(args.seed, k0=4.0, energy=0.5)
            s.initialize()
            s.advance_in_time(10)
            el = timed_steps(s, args.feature_steps)
            out["ms_per_step"][name] = 1e3 * el / args.feature_steps
            log(f"features {n}^3 {name}: {out['ms_per_step'][name]:.3f} ms/step")
    p = ffs.Params(**dict(base, dt_spec=5, les_model="Smagorinsky", smagorinsky_c=0.16,
                          forcing="rfForcing", forcing_kf=2.5, forcing_tf=1.0, forcing_alpha=0.5))
    with ffs.Simulation(p) as s:
        s.seed_velocity(args.seed, k0=4.0, energy=0.5)
        s.initialize()
        s.advance_in_time(10)
        s.pop_spectra()
        s.profile(True)
        s.profile_reset()
        s.advance_in_time(20)
        s.synchronize()
        prof = s.profile_get()
        s.profile(False)
        recs = s.pop_spectra()
        ms, cnt = prof.get("spectra", (0.0, 0))
        out["spectra_record_ms"] = ms / max(cnt, 1)
        out["spectra_records"] = len(recs)
        log(f"features {n}^3: one spectra record {out['spectra_record_ms']:.3f} ms device time ({cnt} records)")
    return out


def cpu_model():
    """The host CPU's model string (/proc/cpuinfo) and its logical core count."""
    name = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return name, os.cpu_count()


def cpu_baseline(n, nu, args):
    """The FP64 oracle (numpy + multi-threaded scipy.fft) on the host cores:
    initialize + `cpu_steps` steps of the same workload."""
    from oracle.ffs_oracle import Grid, OracleSim, workers
    from oracle.seeded import seeded_velocity_hat

    cores = workers()
    log(f"cpu baseline: oracle, {args.cpu_steps} steps at {n}^3 on {cores} threads")
    g = Grid(n, n, n)
    o = OracleSim(g, nu, dt=0.0, variable_dt=True, cfl=0.5, dt_stat=args.dt_stat, forcing_kf=2.5,
                  forcing_power=0.1)
    o.set_velocity([g.irfft(x) for x in seeded_velocity_hat(g, args.seed, k0=4.0, energy=0.5)])
    o.initialize()
    t0 = time.perf_counter()
    o.advance_in_time(args.cpu_steps)
    el = time.perf_counter() - t0
    v = n**3 * args.cpu_steps / el
    log(f"cpu baseline: {el / args.cpu_steps:.2f} s/step, {v:.4e} N^3*steps/s")
    model, ncpu = cpu_model()
    return {"value": v, "unit": "N^3*steps/s", "cores": cores, "kind": "port", "cpu_model": model,
            "host_logical_cpus": ncpu,
            "sample": f"{args.cpu_steps} steps of the same {n}^3 FP64 workload after initialize (numpy/scipy.fft "
                      f"oracle, {el:.1f}s)",
            "s_per_step": el / args.cpu_steps}


if __name__ == "__main__":
    main()
