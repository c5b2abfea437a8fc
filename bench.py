#!/usr/bin/env python3
"""Benchmark of the MI355X DG hydro step: DOF-stage updates/s, FP64, p=1, ~1M elements.

Workload (the plain method at scale): seeded Delaunay basin of ~1M triangles (500k nodes) over
1000 km x 600 km, Gaussian-random-field bathymetry 5-4000 m with a shelf rising to +2 m onshore,
5 tidal constituents on the open boundary, per-node Manning n in [0.02, 0.03], SSP(2,2), slope
limiter and wetting/drying on, dt from CFL 0.3.  A "step" is one full time step (both RK stages
with edge fluxes, volume quadrature, RK update, limiter and wet/dry) over all elements.

    python bench.py [--gpus 1] [--steps 2000] [--warmup 50] [--p 1] [--elements 1000000]
                    [--ens-steps 200] [--adj-steps 2000] [--no-ensemble] [--no-adjoint]

Prints one JSON line.  `value` = NE * 3 * Np * NRK * steps / wall time of the timed steps, with the
state resident in HBM.  `roofline` is for the stage kernel: algorithmic bytes per launch
(300 B per element-stage for p=1, 480 B for p=2) over its mean launch time from HIP events
recorded on the library's stream during the timed region (the launches of every 10th step are
bracketed -- of every step on short runs, at least 20 launches -- so the events barely perturb the
wall time that `value` comes from); `roofline.step_level` is the same byte count over the whole
timed loop.  `cpu_baseline` times dgswem_serial, built from the reference's sources, on one core of
this host (two runs differenced to remove set-up), with the CPU model and a full-host estimate
(rate x cores); `cpu_port` times the CPU oracle (a bit-exact restatement of dgswem_serial's step)
with OpenMP on this host's cores on a short sample of the same mesh.

The same line carries the other two parts of the metric:
* `ensemble` -- 64 uneven spherical meshes (NE log-uniform in 20k-300k) under their own Holland
  hurricanes, limiter and wetting/drying on: wall time of the members run one after another over
  the wall time of the batched run (one stage-kernel launch per RK stage for all members), same
  number of steps on the same GPU, both replaying HIP graphs;
* `adjoint` -- gradient of a 50-gauge misfit with respect to per-node Manning n and depth on a
  200k-element all-wet basin over 2000 steps: wall(forward + checkpointed adjoint) / wall(forward).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0            # MI355X HBM3E
PROFILE_EVERY = 10               # kernel launches of every 10th timed step are bracketed by events
B_PER_ELEM_STAGE = {1: 300.0, 2: 480.0}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def traffic_from_profiles(p: int):
    """HBM bytes per stage-kernel launch from the committed rocprofv3 --pmc summary, if any."""
    path = os.path.join(ROOT, "profiles", f"pmc_stage_kernel_delaunay1m_p{p}.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return float(json.load(f)["hbm_bytes_per_launch"])
    except Exception:
        return None


def cpu_baseline(case, sample_steps: int, threads: int):
    from oracle.oracle import Oracle
    o = Oracle(case, threads=threads)
    o.step(1)                                          # warm caches / first touch
    t0 = time.perf_counter()
    o.step(sample_steps)
    dt = time.perf_counter() - t0
    dof = case.ne * 3 * case.dofh * case.dg.rk_stage * sample_steps
    return dof / dt, dt


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def reference_serial_rate(case, n1: int, n2: int):
    """dgswem_serial (built from the reference's sources into oracle/_ref; a serial program) on one
    core: two runs of the same mesh with n1 and n2 steps; the difference of their wall times
    removes the set-up (reading the mesh, PREP_DG).  Returns (DOF-stage/s, seconds per step, total
    seconds), or None when the binary is not there.  Called before this process touches the GPU
    (the runs are child processes)."""
    import subprocess
    import tempfile
    from dgswem_amd import mesh as _mesh
    exe = os.path.join(ROOT, "oracle", "_ref", "dgswem_serial")
    if not os.path.exists(exe):
        return None
    walls = []
    t_all = time.perf_counter()
    with tempfile.TemporaryDirectory() as d:
        for n in (n1, n2):
            case.params.nt = n
            rd = os.path.join(d, f"run{n}")
            _mesh.write_run_dir(rd, case, nspoolge=n)
            env = dict(os.environ, OMP_NUM_THREADS="1")
            t0 = time.perf_counter()
            r = subprocess.run([exe], cwd=rd, capture_output=True, text=True, env=env)
            walls.append(time.perf_counter() - t0)
            if r.returncode != 0 or "CPU_TIME" not in r.stdout:
                log(f"dgswem_serial failed: {r.stdout[-300:]} {r.stderr[-300:]}")
                return None
    per_step = (walls[1] - walls[0]) / (n2 - n1)
    dof = case.ne * 3 * case.dofh * case.dg.rk_stage
    return dof / per_step, per_step, time.perf_counter() - t_all


def ensemble_leg(gpu, mesh, n_members: int, steps: int, warmup: int, seed: int):
    """Batched ensemble of uneven meshes under Holland hurricanes (NE log-uniform in 20k-300k,
    limiter and wetting/drying on, p=1) against the same members run one after another on the
    same GPU.  Both legs replay HIP graphs and use autotuned block sizes."""
    import numpy as np
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(20_000), np.log(300_000), n_members)).astype(int)
    t0 = time.perf_counter()
    cases = [mesh.storm_basin(int(n), seed=seed + 100 + i, nsteps=warmup + steps)
             for i, n in enumerate(sizes)]
    log(f"ensemble: {n_members} meshes, {sum(c.ne for c in cases)} elements "
        f"({time.perf_counter() - t0:.1f} s to generate)")
    members = [gpu.GpuSolver(c) for c in cases]
    seq = 0.0
    dof_stage = 0
    for m, c in zip(members, cases):
        m.use_graph(True)
        m.step(warmup)
        m.autotune()
        m.synchronize()
        t1 = time.perf_counter()
        m.step(steps, sync=False)
        m.synchronize()
        seq += time.perf_counter() - t1
        dof_stage += c.ne * 3 * c.dofh * c.dg.rk_stage * steps
    fresh = [gpu.GpuSolver(c) for c in cases]
    ens = gpu.Ensemble(fresh)
    for f in fresh:
        f.close()
    ens.use_graph(True)
    ens.step(warmup)
    block = ens.autotune()
    ens.synchronize()
    t1 = time.perf_counter()
    ens.step(steps, sync=False)
    ens.synchronize()
    bat = time.perf_counter() - t1
    ze, qx, qy, _ = ens.state()
    finite = bool(np.isfinite(ze).all() and np.isfinite(qx).all() and np.isfinite(qy).all())
    ens.close()
    for m in members:
        m.close()
    log(f"ensemble: sequential {seq:.3f} s, batched {bat:.3f} s, speedup {seq / bat:.2f}")
    return {"speedup": seq / bat, "members": n_members, "elements": int(sum(c.ne for c in cases)),
            "steps": steps, "sequential_s": seq, "batched_s": bat,
            "batched_dof_stage_per_s": dof_stage / bat, "block": block, "state_finite": finite,
            "workload": "64 spherical Delaunay basins, NE log-uniform 20k-300k, own Holland "
                        "hurricane each (Pmin 920-990 hPa, Rmax 20-60 km, straight/curved track), "
                        "p=1, limiter and wetting/drying on; both legs replay HIP graphs"}


def adjoint_leg(gpu, mesh, n_elem: int, steps: int, seed: int):
    """Gauge-misfit gradient with respect to per-node Manning n and depth on an all-wet Delaunay
    basin without limiter: wall(forward + checkpointed adjoint) / wall(forward)."""
    import numpy as np
    from dgswem_amd import adjoint
    case = mesh.delaunay_basin(n_elem=n_elem, seed=seed, depth_range=(10.0, 3000.0), shelf=False,
                               wetdry=False, slopeflag=0, nsteps=steps, cf=1e-4)
    prob = adjoint.twin_problem(case, nsteps=steps, ng=50, seed=seed)
    s = prob.solver
    s.autotune()
    prob.loss(case.manning, case.dp)            # warm-up
    t0 = time.perf_counter()
    J = prob.loss(case.manning, case.dp)
    fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    J2, gn, gdp = prob.gradient(case.manning, case.dp)
    grad = time.perf_counter() - t0
    interval = max(1, int(round(steps ** 0.5)))
    s.close()
    log(f"adjoint: forward {fwd:.3f} s, forward+adjoint {grad:.3f} s, ratio {grad / fwd:.2f}")
    return {"cost_ratio": grad / fwd, "forward_s": fwd, "gradient_s": grad, "elements": case.ne,
            "nodes": case.np_, "steps": steps, "checkpoint_interval": interval, "loss": J2,
            "params": 2 * case.np_,
            "workload": "all-wet seeded Delaunay basin (GRF depth 10-3000 m, CF 1e-4), p=1, no limiter, "
                        "50 gauges, misfit against a perturbed-Manning twin; gradient w.r.t. "
                        "per-node Manning n and depth"}


def batched_adjoint_leg(mesh, n_members: int, n_elem: int, steps: int, seed: int):
    """Per-member gauge-misfit gradients (Manning n and depth) of a p=2 storm ensemble: every member
    a seeded spherical basin with a dry coast under its own Holland vortex with its own gauges and
    perturbed-Manning twin data, from the cold start, slope limiter and wetting/drying on with their
    active sets frozen from the forward run, all members in one batched forward and one batched
    checkpointed reverse sweep: wall(forward + adjoint) / wall(forward)."""
    import numpy as np
    from dgswem_amd import adjoint
    t0 = time.perf_counter()
    cases = [mesh.storm_basin(n_elem, seed=seed + 300 + i, nsteps=steps, p=2, slopeflag=5, wetdry=True,
                              shelf=True, cfl=0.15, depth_range=(10.0, 300.0), cf=1e-4)
             for i in range(n_members)]
    eprob, probs = adjoint.ensemble_problem(cases, nsteps=steps, ng=20, seed=seed + 400)
    for p in probs:
        p.solver.close()
    mn = np.concatenate([c.manning for c in cases])
    dp = np.concatenate([c.dp for c in cases])
    log(f"batched adjoint: {n_members} members, {sum(c.ne for c in cases)} elements, p=2 "
        f"({time.perf_counter() - t0:.1f} s to set up)")
    eprob.solver.autotune()
    eprob.loss(mn, dp)                              # warm-up
    t0 = time.perf_counter()
    eprob.loss(mn, dp)
    fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    J, gn, gdp = eprob.gradient(mn, dp)
    grad = time.perf_counter() - t0
    finite = bool(np.isfinite(gn).all() and np.isfinite(gdp).all())
    eprob.solver.close()
    log(f"batched adjoint: forward {fwd:.3f} s, forward+adjoint {grad:.3f} s, ratio {grad / fwd:.2f}")
    return {"cost_ratio": grad / fwd, "forward_s": fwd, "gradient_s": grad, "members": n_members,
            "elements": int(sum(c.ne for c in cases)), "p": 2, "steps": steps,
            "checkpoint_interval": max(1, int(round(steps ** 0.5))), "loss": J, "gradient_finite": finite,
            "workload": f"{n_members} spherical storm basins of {n_elem} elements with a shelf rising 2 m "
                        "above the datum along the coast (dry land, a shoreline WETDRY redistributes), "
                        "p=2, limiter 5 and wetting/drying on (active sets frozen), from the cold start, "
                        "own Holland vortex, 20 gauges and perturbed-Manning twin each; per-member "
                        "gradients w.r.t. Manning n and depth"}


def copy_calibration() -> int:
    """2 GiB (larger than the 256 MiB last-level cache) copied device to device three times."""
    import torch
    n = 2 * 1024 ** 3 // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": a.numel() * 8, "copies": 3}), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--p", type=int, default=1)
    ap.add_argument("--elements", type=int, default=1_000_000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-steps", type=int, default=0, help="oracle sample steps (0 = auto)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--block", type=int, default=0,
                    help="stage-kernel block size (0 = autotune on the warmed-up state)")
    ap.add_argument("--ens-members", type=int, default=64)
    ap.add_argument("--ens-steps", type=int, default=200,
                    help="timed steps of each ensemble leg (the ratio is per step)")
    ap.add_argument("--adj-elements", type=int, default=200_000)
    ap.add_argument("--adj-steps", type=int, default=2000)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--badj-members", type=int, default=32)
    ap.add_argument("--badj-elements", type=int, default=100_000)
    ap.add_argument("--badj-steps", type=int, default=1000)
    ap.add_argument("--no-batched-adjoint", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy a 2 GiB device buffer three times (a plain copy kernel of known "
                         "size, to scale the FETCH_SIZE / WRITE_SIZE counters of a profiler run)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(one GPU per run; a multi-GPU run would shard the mesh by element partition)", file=sys.stderr)
        return 2

    from dgswem_amd import gpu, mesh

    t0 = time.perf_counter()
    # CFL number 0.3 at p=1, 0.15 at p=2 (the explicit DG limit shrinks like 1/(2p+1))
    case = mesh.delaunay_basin(n_elem=args.elements, seed=args.seed, p=args.p,
                               cfl=0.3 if args.p == 1 else 0.15, nsteps=args.steps + args.warmup)
    log(f"mesh: {case.ne} elements, {case.np_} nodes, dt {case.params.dtdp:.4f} s "
        f"({time.perf_counter() - t0:.1f} s)")
    ne = case.ne
    model = cpu_model()
    threads = len(os.sched_getaffinity(0))
    ref = None
    if not args.no_cpu:
        # the reference's own serial solver on one core (steps differenced), before any GPU work
        n1, n2 = (2, 8) if ne >= 500_000 else (5, 45)
        log(f"cpu baseline: dgswem_serial, 1 core, {n1} and {n2} steps ({model})")
        ref = reference_serial_rate(case, n1, n2)
        case.params.nt = args.steps + args.warmup
    g = gpu.GpuSolver(case)
    log("library initialised")
    g.step(args.warmup)
    g.synchronize()
    if args.block:
        g.set_block(args.block)
        block = args.block
    else:
        block = g.autotune()
    log(f"stage kernel block size {block} ({'set' if args.block else 'autotuned'})")
    # time the launches of every n-th step, with at least 20 stage launches sampled
    every = max(1, min(PROFILE_EVERY, (args.steps * case.dg.rk_stage) // 40))
    g.profile(True, every=every)
    g.synchronize()
    t1 = time.perf_counter()
    g.step(args.steps, sync=False)
    g.synchronize()
    wall = time.perf_counter() - t1
    stats = {k: g.kernel_stats(k) for k in (gpu.K_STAGE, gpu.K_MINMAX, gpu.K_LIMIT)}
    g.profile(False)
    ze, qx, qy, wd = g.state()
    import numpy as np
    finite = bool(np.isfinite(ze).all() and np.isfinite(qx).all() and np.isfinite(qy).all())
    ne, dof, nrk = case.ne, case.dofh, case.dg.rk_stage
    dof_stage = ne * 3 * dof * nrk * args.steps
    value = dof_stage / wall
    name, tot_ms, launches = stats[gpu.K_STAGE]
    avg_s = tot_ms / 1e3 / max(launches, 1)
    alg_bytes = B_PER_ELEM_STAGE[args.p] * ne
    traffic = traffic_from_profiles(args.p)
    achieved = alg_bytes / avg_s / 1e9
    step_gbs = alg_bytes * nrk * args.steps / wall / 1e9
    kernels = {}
    for k, (nm, ms, cnt) in stats.items():
        if cnt:
            kernels[nm] = {"launches": cnt, "avg_us": round(ms * 1e3 / cnt, 2),
                           "share_of_wall": round(ms / 1e3 / wall, 4)}
    log(f"timed {args.steps} steps: {wall * 1e3:.1f} ms, stage kernel {avg_s * 1e6:.1f} us/launch, "
        f"{achieved:.0f} GB/s algorithmic, finite={finite}")

    cpu = port = None
    if not args.no_cpu:
        # the OpenMP oracle on every core this process may use, beside dgswem_serial's rate
        if ref is not None:
            rate, per_step, secs = ref
            cpu = {"value": rate, "unit": "DOF-stage updates/s", "cores": 1, "kind": "reference",
                   "sample": f"dgswem_serial (built from the reference's sources) on the same {ne}-element "
                             f"mesh, runs of {n1} and {n2} steps differenced ({per_step:.3f} s/step, "
                             f"{secs:.1f} s in all)", "cpu_model": model, "host_threads": threads,
                   "full_host_estimate": rate * threads}
        steps_cpu = args.cpu_steps or max(2, int(150 * 1_000_000 / max(ne, 1)))
        log(f"cpu port: oracle, {threads} threads, {steps_cpu} steps")
        prate, psecs = cpu_baseline(case, steps_cpu, threads)
        port = {"value": prate, "unit": "DOF-stage updates/s", "cores": threads, "kind": "port",
                "sample": f"{steps_cpu} steps of the same {ne}-element mesh ({psecs:.1f} s) with the "
                          f"OpenMP oracle (bit-exact restatement of dgswem_serial's step)",
                "cpu_model": model}
        if cpu is None:
            cpu, port = port, None

    g.close()
    ens = None if args.no_ensemble else ensemble_leg(gpu, mesh, args.ens_members, args.ens_steps,
                                                     10, args.seed)
    adj = None if args.no_adjoint else adjoint_leg(gpu, mesh, args.adj_elements, args.adj_steps,
                                                   args.seed)
    badj = None if args.no_batched_adjoint else batched_adjoint_leg(mesh, args.badj_members, args.badj_elements,
                                                                    args.badj_steps, args.seed)

    out = {
        "metric": "DOF-stage updates/s FP64 p=1 1M elems; ensemble speedup; adjoint/forward cost",
        "value": value,
        "unit": "DOF-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"seeded Delaunay tidal basin with shelf, {ne} elements, p={args.p}, "
                               f"SSP(2,2), slope limiter 5, wetting/drying, 5 constituents",
                   "elements": ne, "nodes": case.np_, "p": args.p, "rk_stages": nrk,
                   "dt_s": case.params.dtdp, "limiter": case.dg.slopeflag,
                   "wetdry": case.params.nolifa >= 2, "seed": args.seed, "state_finite": finite,
                   "block": block},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": name, "avg_launch_us": avg_s * 1e6, "timed_launches": launches,
                     "algorithmic_bytes_per_launch": alg_bytes,
                     # the metric's own definition: B_p per element-stage over the whole timed loop
                     "step_level": {"achieved": step_gbs, "frac": step_gbs / HBM_PEAK_GBS},
                     "limited_by": "FP64 VALU issue" + (f" (HBM traffic {traffic / alg_bytes:.2f}x the algorithmic bytes)"
                                                        if traffic else "")},
        "kernels": kernels,
        "cpu_baseline": cpu,
        "ensemble": ens,
        "adjoint": adj,
        "batched_adjoint": badj,
    }
    if cpu:
        out["speedup_vs_cpu_baseline"] = value / cpu["value"]
        if cpu.get("full_host_estimate"):
            out["speedup_vs_full_host_estimate"] = value / cpu["full_host_estimate"]
    if port:
        out["cpu_port"] = port
        out["speedup_vs_cpu_port"] = value / port["value"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
