#!/usr/bin/env python3
"""Benchmark of the D3Q19 fluid + temperature streaming step on one MI355X.

Workload (default): the 3-D Rayleigh-Benard configuration, 512^3 lattice
(y vertical), D3Q19 fluid + D3Q19 temperature populations, Boussinesq
buoyancy at Ra = 1e9, Pr = 1, no-slip and fixed-T walls at y, periodic x and
z, initial T linear plus seeded uniform noise 1e-3 -- the largest
configuration that fits one GPU (41 GB in FP32-deviation storage, 83 GB in
FP64).  The timed path is the FP32-deviation fused stream+collide kernel
(value); the FP64 plain path is timed beside it.  A step = every node of both
population sets advanced one time step; inputs are resident in HBM.  The 2-D
512 x 256 configuration (40 MB, resident in the 256 MB last-level cache) is
reported beside it under its own key.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "ch4-project_amd"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

METRIC = "D3Q19 fluid+T MLUPS, % HBM roofline; served-ensemble MLUPS + p99 latency"
HBM_PEAK_GBS = 8000.0
BYTES_PER_NODE = {"fp32dev": 2 * (19 + 19) * 4, "fp64": 2 * (19 + 19) * 8}  # populations read once + written once


WORKLOADS = {
    # name: (description, config factory(precision))
    "rb2d": ("2-D Rayleigh-Benard 512x1x256 (NX=512, NY=256 vertical), D3Q19 fluid + D3Q19 temperature, "
             "Ra=1e8, Pr=1", lambda c, p: c.rb2d(precision=p)),
    "rb3d-256": ("3-D Rayleigh-Benard 256^3, D3Q19 fluid + D3Q19 temperature, Ra=1e9, Pr=1",
                 lambda c, p: c.rb2d(NX=256, NY=256, NZ=256, Ra=1e9, precision=p)),
    "rb3d-512": ("3-D Rayleigh-Benard 512^3, D3Q19 fluid + D3Q19 temperature, Ra=1e9, Pr=1",
                 lambda c, p: c.rb2d(NX=512, NY=512, NZ=512, Ra=1e9, precision=p)),
    "channel-192": ("3-D channel 192^3, D3Q19 fluid, body force 1e-6", lambda c, p: c.channel3d(192, 192, 192,
                                                                                               precision=p)),
}


def pmc_traffic(workload, kernel):
    """HBM bytes per launch of the workload's step kernel, from the rocprofv3
    FETCH_SIZE / WRITE_SIZE runs recorded in profiles/pmc_traffic.json
    (calibrated on a plain copy; the matching profiles/*_pmc_traffic.md gives
    the commands).  Counters cannot be read inside this process, so the
    figure is the recorded one: it is used only when the record names the
    same kernel, and the line says which commit and configuration it comes
    from."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as fh:
            rec = json.load(fh)[workload]
    except (OSError, KeyError, ValueError):
        return None, None
    if rec.get("kernel") != kernel:
        return None, None
    src = (f"profiles/pmc_traffic.json[{workload}]: {rec.get('kernel')} at commit {rec.get('commit', '?')}, "
           f"{rec.get('config', workload)}")
    return rec["traffic_bytes"], src


def gpu_leg(precision, steps, warmup, workload="rb2d"):
    import configs
    from ch4g import Lattice
    cfg = WORKLOADS[workload][1](configs, precision)
    lat = Lattice(cfg)
    lat.initial_conditions()
    lat.step(warmup)
    lat.synchronize()
    t0 = time.perf_counter()
    ms = lat.step_timed(steps)
    lat.synchronize()
    wall = time.perf_counter() - t0
    nodes = cfg.nodes()
    spl = lat.steps_per_launch()
    out = {"nodes": nodes, "thermal": cfg.has(1 << 1), "wall_s": wall, "event_ms": ms, "mlups": nodes * steps / wall / 1e6,
           "us_per_step": ms * 1e3 / steps, "steps_per_launch": spl, "avg_launch_us": ms * 1e3 / steps * spl,
           "kernel": "k_step_fp64" if precision == 0 else (
               ("k_step2z_fp32" if cfg.NZ > 1 else "k_step2_fp32") if spl == 2 else "k_step_fp32")}
    h = lat.dump_averages()
    out["nusselt"] = h["nuy"]
    lat.close()
    return out


def cpu_leg(workload, target_s=12.0):
    """The oracle (FP64, reference-structured un-fused sweeps, OpenMP) on
    this host's cores, on a bounded sample of the same workload: the 3-D
    workloads are sampled on a 128^3 lattice of the same physics (fluid + T,
    Ra = 1e9; 512^3 needs ~90 GB of host memory in the reference's ghost-cell
    layout), the 2-D one at its own size."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import configs
    from oracle import Oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", min(16, os.cpu_count() or 1)))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    if workload.startswith("rb3d"):
        cfg = configs.rb2d(NX=128, NY=128, NZ=128, Ra=1e9)
        what = "3-D Rayleigh-Benard 128^3 fluid+T FP64 (Ra=1e9)"
    else:
        cfg = configs.rb2d()
        what = "2-D Rayleigh-Benard 512x256 fluid+T FP64"
    o = Oracle(cfg)
    o.initial_conditions()
    t0 = time.perf_counter()
    o.step(2)
    probe = time.perf_counter() - t0
    n = max(2, min(2000, int(target_s / max(probe / 2, 1e-6))))
    t0 = time.perf_counter()
    o.step(n)
    dt = time.perf_counter() - t0
    return {"value": round(cfg.nodes() * n / dt / 1e6, 3), "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"{what}, {n} steps ({dt:.1f} s), oracle/ch4_oracle.c un-fused reference-structured sweeps, "
                      f"OpenMP {threads} threads"}


def _serve_both(make_jobs):
    """ch4g_serve on the same request stream three ways: batched (continuous
    batching, lock-step chunks, FP32-deviation storage); one job at a time
    with the same storage (ch4g_step with graphs); one job at a time on the
    plain FP64 path (the reference's method, no added pattern)."""
    import numpy as np
    from ch4g import abi, serve, workloads
    res = {}
    for name, mode, prec in (("batched", abi.SERVE_BATCHED, abi.FP32_DEV),
                             ("sequential", abi.SERVE_SEQUENTIAL, abi.FP32_DEV),
                             ("sequential_plain_fp64", abi.SERVE_SEQUENTIAL, abi.FP64)):
        done = serve(make_jobs(prec), mode=mode)
        ok = [j for j in done if j.status == 0]
        work = sum(workloads.job_work(j) for j in ok)
        span = max(j.finish_s for j in ok)
        lat = np.array([j.finish_s - j.arrival_s for j in ok])
        res[name] = {"mlups": round(work / span / 1e6, 1), "makespan_s": round(span, 3),
                     "p50_latency_s": round(float(np.percentile(lat, 50)), 3),
                     "p99_latency_s": round(float(np.percentile(lat, 99)), 3), "jobs_done": len(ok),
                     "node_updates": work}
    res["speedup"] = round(res["batched"]["mlups"] / res["sequential"]["mlups"], 3)
    res["speedup_vs_plain_fp64"] = round(res["batched"]["mlups"] / res["sequential_plain_fp64"]["mlups"], 3)
    return res


def serve_leg(njobs, seed=2026, sweep_jobs=48):
    """Served ensembles through ch4g_serve.  (1) The request stream of
    workloads.ensemble: Poisson arrivals at 20 jobs/s, 2-D RB and 3-D channel
    jobs of log-uniform sizes and step counts 1e3..1e5, FP32-deviation (most
    of its work is in lattices large enough to stream at HBM speed alone).
    (2) A parameter sweep (workloads.sweep): a burst of small 2-D RB jobs,
    each too small to fill the GPU alone."""
    from ch4g import workloads
    res = _serve_both(lambda prec: workloads.ensemble(njobs, seed=seed, precision=prec))
    res["workload"] = (f"{njobs} jobs of the served-ensemble stream (workloads.ensemble seed {seed}): Poisson "
                       "20 jobs/s, 2-D Rayleigh-Benard 64x64..1024x512 Ra 1e6..1e9 and 3-D channel 64^3..192^3, "
                       "steps log-uniform 1e3..1e5, FP32-deviation")
    sw = _serve_both(lambda prec: workloads.sweep(sweep_jobs, precision=prec))
    sw["workload"] = (f"{sweep_jobs} jobs of a parameter sweep (workloads.sweep seed 7): burst at 200 jobs/s of "
                      "2-D Rayleigh-Benard 64x32..256x128, Ra 1e5..1e7, steps log-uniform 2e3..2e4, FP32-deviation")
    res["sweep"] = sw
    return res


def batch_compare(steps=320, graphs=False):
    """Throughput (GLUPS) of sets of lattices stepped one by one (one
    ch4g_step call per lattice, each lattice on its own stream) and together
    through ch4g_step_batch (one launch per population kind and step, 32 steps
    per call).  graphs=False launches every single-lattice step, so a kernel
    trace sees every launch."""
    import time
    from ch4g import abi, Lattice, step_batch, workloads
    ch, rb = workloads.channel3d, workloads.rb2d
    cases = {
        "channel 78^3": [ch(78, 78, 78)],
        "4 x channel 78^3": [ch(78, 78, 78)] * 4,
        "RB 234x169": [rb(234, 169, 1, Ra=1e7, precision=abi.FP32_DEV)],
        "channel 160^3": [ch(160, 160, 160)],
        "channel 192^3": [ch(192, 192, 192)],
        "8 x channel 80^3": [ch(80, 80, 80)] * 8,
        "RB 512x256": [rb(512, 256, 1, Ra=1e8, precision=abi.FP32_DEV)],
        "16 x RB 512x256": [rb(512, 256, 1, Ra=1e8, precision=abi.FP32_DEV)] * 16,
        "64 x RB 128x64": [rb(128, 64, 1, Ra=1e6, precision=abi.FP32_DEV)] * 64,
        "mix": [ch(113, 113, 113), ch(78, 78, 78), rb(584, 336, 1, Ra=1e8, precision=abi.FP32_DEV),
                ch(133, 133, 133)],
    }

    def started(cfg):
        lat = Lattice(cfg)
        lat.set_graphs(graphs)
        lat.initial_conditions()
        lat.step(1)
        return lat

    out = {}
    for name, cfgs in cases.items():
        nodes = sum(c.nodes() for c in cfgs)
        ls = [started(c) for c in cfgs]
        for lat in ls:
            lat.step(32)
            lat.synchronize()
        t0 = time.perf_counter()
        for lat in ls:
            lat.step(steps)
        for lat in ls:
            lat.synchronize()
        single = nodes * steps / (time.perf_counter() - t0) / 1e9
        for lat in ls:
            lat.close()
        ls = [started(c) for c in cfgs]
        step_batch(ls, 32)
        t0 = time.perf_counter()
        for _ in range(steps // 32):
            step_batch(ls, 32)
        batch = nodes * (steps // 32) * 32 / (time.perf_counter() - t0) / 1e9
        for lat in ls:
            lat.close()
        out[name] = {"single_glups": round(single, 2), "batch_glups": round(batch, 2)}
    return out


def particle_leg(steps=40, npart=2_000_000, N=256):
    """HIT 256^3 (FP32-deviation, LES Cs 0.1) with 1e6 tracers + 1e6 inertial
    particles: device time per step without and with the particles."""
    import numpy as np
    import configs
    from ch4g import Lattice, abi
    cfg = configs.hit(N, tau=0.5015, amp=1e-5, les_cs=0.1, precision=abi.FP32_DEV)
    lat = Lattice(cfg)
    lat.initial_conditions()
    lat.step(5)
    fluid_ms = lat.step_timed(steps) / steps
    lat.close()
    lat = Lattice(cfg)
    lat.initial_conditions()
    lat.particles_init(abi.make_particle_config(abi.LAGRANGE_WRAP, 1, tau_drag=(1, 1.0, 1.0)))
    rng = np.random.default_rng(7)
    a = np.zeros(npart, dtype=abi.particle_dtype())
    a["name"] = np.arange(npart)
    for c in ("x", "y", "z"):
        a[c] = rng.uniform(0.0, N, npart)
    half = npart // 2
    a["tau_drag"][half:] = rng.uniform(1.0, 50.0, npart - half)
    lat.set_particles(a)
    lat.step(5)
    both_ms = lat.step_timed(steps) / steps
    alive = lat.particles_count()
    lat.close()
    pms = both_ms - fluid_ms
    return {"workload": f"HIT {N}^3 FP32-deviation LES Cs=0.1 + {half} tracers + {npart - half} inertial "
                        "particles (tau_p uniform in 1..50, positions uniform), trilinear interpolation, AB2",
            "fluid_ms_per_step": round(fluid_ms, 4), "with_particles_ms_per_step": round(both_ms, 4),
            "particle_ms_per_step": round(pms, 4),
            "particle_updates_per_s": round(npart / (pms * 1e-3), 1) if pms > 0 else None, "alive": alive}


def twoway_leg(steps=6, npart=4_000_000, N=512):
    """The particle part of the 3-D configuration: 512^3 fluid + T
    (FP32-deviation, Ra = 1e9) with 4e6 heavy particles (beta = 0.6, i.e.
    density ratio 2; tau_p = 516 steps, Stokes number ~0.5 on the Kolmogorov
    time of Ra = 1e9) coupled back onto the fluid (per-node feedback sums in
    particle order).  The particle gravity is 2e-5 with beta_t
    rescaled so that the fluid's beta_t g is unchanged.  Device time per step
    without and with the particles."""
    import configs
    from ch4g import Lattice, abi
    c = configs.rb2d(NX=N, NY=N, NZ=N, Ra=1e9, precision=abi.FP32_DEV)
    g = 2e-5
    cfg = c.copy(gravity_y=g, beta_t=c.beta_t * c.gravity_y / g)
    lat = Lattice(cfg)
    lat.initial_conditions()
    lat.step(3)
    fluid_ms = lat.step_timed(steps) / steps
    lat.close()
    flags = (abi.LAGRANGE_WRAP | abi.LAGRANGE_GRADIENT | abi.LAGRANGE_ADDEDMASS | abi.LAGRANGE_GRAVITY |
             abi.LAGRANGE_TWOWAY_MOMENTUM)
    lat = Lattice(cfg)
    lat.initial_conditions()
    lat.particles_init(abi.make_particle_config(flags, npart, tau_drag=(1, 516.0, 516.0), beta_coeff=(1, 0.6, 0.6)))
    lat.step(3)
    both_ms = lat.step_timed(steps) / steps
    alive = lat.particles_count()
    lat.close()
    pms = both_ms - fluid_ms
    return {"workload": f"3-D Rayleigh-Benard {N}^3 FP32-deviation + {npart} two-way heavy particles (beta 0.6, "
                        "tau_p 516), gradients, added mass, gravity, feedback momentum",
            "fluid_ms_per_step": round(fluid_ms, 4), "with_particles_ms_per_step": round(both_ms, 4),
            "particle_ms_per_step": round(pms, 4),
            "particle_updates_per_s": round(npart / (pms * 1e-3), 1) if pms > 0 else None, "alive": alive}


def roofline_of(r, bpn):
    """Algorithmic bytes per launch (bytes per node update x nodes x steps
    per launch) over the average launch duration (HIP events on the lattice
    stream around the timed steps)."""
    alg_launch = bpn * r["nodes"] * r["steps_per_launch"]
    achieved = alg_launch / (r["avg_launch_us"] * 1e-6) / 1e9
    return alg_launch, achieved


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 plain-path leg")
    ap.add_argument("--workload", default="rb3d-512", choices=sorted(WORKLOADS))
    ap.add_argument("--serve-jobs", type=int, default=16, help="jobs of the served-ensemble leg (0: skip)")
    ap.add_argument("--no-particles", action="store_true", help="skip the particle legs")
    ap.add_argument("--no-rb2d", action="store_true", help="skip the 2-D (cache-resident) leg")
    ap.add_argument("--batch-compare", action="store_true",
                    help="only compare single-lattice and batched stepping of sets of lattices (no graphs)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run plain 1 GiB device-to-device copies (PMC byte-counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        a = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1.0)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": a.numel() * 4, "copies": 5}))
        return
    if args.batch_compare:
        print(json.dumps({"batch_compare": batch_compare()}), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: runs on one GPU only (one lattice per GPU; a slab decomposition over several GPUs "
              "is not built). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)
    from ch4g import abi
    r32 = gpu_leg(abi.FP32_DEV, args.steps, args.warmup, args.workload)
    print(f"bench: {args.workload} FP32-deviation {r32['mlups']:.1f} MLUPS", file=sys.stderr, flush=True)
    bpn = BYTES_PER_NODE["fp32dev"] if r32["thermal"] else BYTES_PER_NODE["fp32dev"] // 2
    alg_launch, achieved = roofline_of(r32, bpn)
    traffic, traffic_src = pmc_traffic(args.workload, r32["kernel"])
    line = {
        "metric": METRIC,
        "value": round(r32["mlups"], 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["wall_s"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": WORKLOADS[args.workload][0] + ", FP32-deviation storage",
                   "nodes": r32["nodes"], "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": r32["kernel"], "steps_per_launch": r32["steps_per_launch"],
                     "bytes_per_node": bpn, "algorithmic_bytes_per_launch": alg_launch,
                     "avg_launch_us": round(r32["avg_launch_us"], 3)},
    }
    if not args.no_fp64:
        r64 = gpu_leg(abi.FP64, args.steps, args.warmup, args.workload)
        b64 = BYTES_PER_NODE["fp64"] if r64["thermal"] else BYTES_PER_NODE["fp64"] // 2
        _, a64 = roofline_of(r64, b64)
        line["fp64_plain"] = {"mlups": round(r64["mlups"], 2), "kernel": r64["kernel"],
                              "avg_launch_us": round(r64["avg_launch_us"], 3),
                              "roofline_frac": round(a64 / HBM_PEAK_GBS, 4),
                              "fp32dev_speedup": round(r32["mlups"] / r64["mlups"], 3),
                              "fp32dev_speedup_kernel_time": round(r64["us_per_step"] / r32["us_per_step"], 3)}
        print(f"bench: {args.workload} FP64 {r64['mlups']:.1f} MLUPS", file=sys.stderr, flush=True)
    if args.workload != "rb2d" and not args.no_rb2d:
        # the 2-D configuration: 40 MB of state, resident in the 256 MB
        # last-level cache, so its fraction is not an HBM figure
        k = max(args.steps, 4000)
        a = gpu_leg(abi.FP32_DEV, k, 200, "rb2d")
        _, ach = roofline_of(a, BYTES_PER_NODE["fp32dev"])
        rb2d = {"workload": WORKLOADS["rb2d"][0] + ", FP32-deviation storage; state resident in the last-level cache",
                "steps": k, "mlups": round(a["mlups"], 2), "kernel": a["kernel"],
                "steps_per_launch": a["steps_per_launch"], "avg_launch_us": round(a["avg_launch_us"], 3),
                "algorithmic_frac_cache_resident": round(ach / HBM_PEAK_GBS, 4)}
        if not args.no_fp64:
            b = gpu_leg(abi.FP64, k, 200, "rb2d")
            rb2d["fp64_mlups"] = round(b["mlups"], 2)
            rb2d["fp32dev_speedup"] = round(a["mlups"] / b["mlups"], 3)
        line["rb2d_cache_resident"] = rb2d
    if not args.no_particles:
        line["particles"] = particle_leg()
        print("bench: HIT particle leg done", file=sys.stderr, flush=True)
        if args.workload == "rb3d-512":
            line["particles_twoway_rb3d"] = twoway_leg()
            print("bench: two-way particle leg done", file=sys.stderr, flush=True)
    if args.serve_jobs > 0:
        line["serving"] = serve_leg(args.serve_jobs)
        print("bench: serving leg done", file=sys.stderr, flush=True)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_leg(args.workload)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
