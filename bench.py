#!/usr/bin/env python3
"""Benchmark: HAADF MBIR reconstruction, 512x64x256 volume, 91 tilts over +-70 deg.

A step is one outer iteration of the reconstruction engine's loop in steady
state: one ICD sweep over all N_x*N_y*N_z voxels followed by the joint
gain/offset estimation and the noise-variance update
(HAADF/HAADF_ReconstructionEngine.cpp:706-815 with NumIter = 1).  The sinogram,
volume and nuisance state are resident on the GPU before timing starts.

value = voxel visits per second (N_x*N_y*N_z per sweep; zero-skipped visits are
reported separately).  roofline: the ICD line-update kernel (k_icd_step), its
algorithmic bytes per launch over its average launch duration measured with HIP
events on the engine's stream.  cpu_baseline: the FP64 C oracle's sweep with the
reference's y-slab thread decomposition on this host's cores.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "openmbir-tem_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="haadf_512x64x256_91v")
    ap.add_argument("--schedule", type=int, default=0)
    ap.add_argument("--rows-per-step", type=int, default=1)
    ap.add_argument("--kernel-mode", type=int, default=0,
                    help="0 = the job's own step kernel, 1 = the batched step kernel")
    ap.add_argument("--batch", type=int, default=8,
                    help="jobs of the same workload swept together for the batched figure (0 = skip)")
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--traffic-json", default=os.path.join(REPO, "profiles", "icd_step_traffic.json"))
    return ap.parse_args()


def cpu_baseline(wl, ang, y, threads):
    """Time one steady-state sweep of the FP64 oracle (bounded sample: one sweep)."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import numpy as np

    import oracle_py as O

    g = O.geometry(wl.n_r, wl.n_t, ang, wl.thickness)
    prm = O.params(sigma_x=0.1, mrf_q=1.2, mrf_c=1e-3, joint_estimation=wl.joint_estimation,
                   noise_model=wl.noise_model, stop_threshold=wl.stop_threshold, seed=wl.seed)
    t0 = time.time()
    eng = O.Engine(g, ang, np.asarray(y, np.float64), prm)
    log(f"cpu: oracle setup {time.time() - t0:.1f}s")
    eng.sweep(0, 0, threads=threads)  # first sweep (no zero skipping), untimed
    t0 = time.time()
    _, st = eng.sweep(1, 0, threads=threads)
    dt = time.time() - t0
    V = g.N_x * g.N_y * g.N_z
    log(f"cpu: steady sweep {dt:.2f}s on {threads} threads, {V / dt:.3g} visits/s")
    return {"value": V / dt, "unit": "voxel-updates/s", "cores": threads, "kind": "port",
            "sample": f"one steady-state ICD sweep (second sweep, zero skipping on) of the FP64 "
                      f"oracle over the full {g.N_x}x{g.N_y}x{g.N_z} volume, {threads} y-slab threads"}


def batched(ctx, wl, n, steps, single_value):
    """Throughput of n independent jobs of the workload swept together (one batched
    step kernel launch carries one conflict-free step of every job)."""
    import time as _t

    from mbirtem import capi, workloads

    jobs = []
    for i in range(n):
        w = workloads.make(ctx, wl, seed=wl.seed + 1000 * (i + 1), kernel_mode=1)
        jobs.append(w["job"])
    g = jobs[0].geom
    V = g.N_x * g.N_y * g.N_z
    capi.batch_sweep(ctx, jobs, inner=[0] * n)  # first sweeps (no zero skipping)
    capi.batch_sweep(ctx, jobs, inner=[1] * n)
    t0 = _t.perf_counter()
    for k in range(steps):
        capi.batch_sweep(ctx, jobs, inner=[2 + k] * n)
        if wl.joint_estimation or wl.noise_model:
            for j in jobs:
                j.nuisance_update()
    dt = _t.perf_counter() - t0
    val = n * V * steps / dt
    for j in jobs:
        j.close()
    log(f"batched: {n} jobs x {steps} steps in {dt:.3f}s -> {val:.4g} visits/s")
    return {"jobs": n, "value": val, "unit": "voxel-updates/s", "ms_per_step": 1e3 * dt / steps,
            "vs_single_job": val / single_value}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine measures one GPU; --gpus N>1 is not run "
                                   "(multi-GPU sharding is a design, not a measured path)"}))
        sys.exit(2)
    import numpy as np
    import torch

    from mbirtem import capi, workloads

    assert torch.cuda.is_available(), "no GPU visible"
    wl = workloads.WORKLOADS[args.workload]
    ctx = capi.Context(0)
    t0 = time.time()
    w = workloads.make(ctx, wl, schedule=args.schedule, rows_per_step=args.rows_per_step,
                       kernel_mode=args.kernel_mode)
    job, geom = w["job"], w["geom"]
    V = geom.N_x * geom.N_y * geom.N_z
    log(f"setup {time.time() - t0:.1f}s: {geom.N_x}x{geom.N_y}x{geom.N_z}, {geom.N_theta} views")

    it = 0

    def step():
        nonlocal it
        _, st = job.sweep(it, 0)
        if wl.joint_estimation or wl.noise_model:
            job.nuisance_update()
        it += 1
        return st

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = [step() for _ in range(args.steps)]
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    value = V * args.steps / dt
    upd = sum(s.updated for s in stats)
    zs = sum(s.zero_skipped for s in stats)
    log(f"timed: {args.steps} steps in {dt:.3f}s -> {value:.4g} visits/s "
        f"(updated {upd}, zero-skipped {zs})")

    # live roofline of k_icd_step: HIP events around every launch on the job's stream
    job.set_timing(True)
    rl_bytes, rl_ms, rl_launches = 0.0, 0.0, 0
    for _ in range(2):
        st = step()
        rl_bytes += 12.0 * st.nnz_updated + 8.0 * st.updated + 4.0 * st.zero_skipped
        rl_ms += st.kernel_ms
        rl_launches += st.steps
    job.set_timing(False)
    achieved = rl_bytes / (rl_ms * 1e-3) / 1e9
    traffic = None
    if os.path.exists(args.traffic_json):
        with open(args.traffic_json) as f:
            traffic = json.load(f).get(wl.name)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "k_icd_step", "avg_launch_us": round(1e3 * rl_ms / rl_launches, 3),
                "bytes_per_launch": round(rl_bytes / rl_launches, 1)}
    cost = job.cost()
    out = {
        "metric": "voxel-updates/s",
        "value": value,
        "unit": "voxel-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": wl.name, "volume": [geom.N_x, geom.N_y, geom.N_z],
                   "views": geom.N_theta, "tilt_deg": wl.max_deg,
                   "step": "ICD sweep + joint gain/offset + noise update",
                   "schedule": "conflict-free" if args.schedule == 0 else
                               f"relaxed rows_per_step={args.rows_per_step}",
                   "kernel": "batched step kernel" if args.kernel_mode == 1 else "step kernel",
                   "parallelism": "single GPU"},
        "updated_frac": upd / max(1, upd + zs),
        "final_cost": cost,
        "roofline": roofline,
    }
    if args.batch > 0:
        out["batched"] = batched(ctx, wl, args.batch, args.batch_steps, value)
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(wl, w["angles"], w["sinogram"], threads)
    else:
        out["cpu_baseline"] = None
    job.close()
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
