#!/usr/bin/env python3
"""Benchmark: HAADF MBIR reconstruction, 512x64x256 volume, 91 tilts over +-70 deg.

A step is one outer iteration of the reconstruction engine's loop in steady
state: one ICD sweep over all N_x*N_y*N_z voxels followed by the joint
gain/offset estimation and the noise-variance update
(HAADF/HAADF_ReconstructionEngine.cpp:706-815 with NumIter = 1).  The sinogram,
volume and nuisance state are resident on the GPU before timing starts.

value = voxel visits per second (N_x*N_y*N_z per sweep; zero-skipped visits are
reported separately).  batched_slabs: 48 independent slab problems of uneven size
reconstructed to the stop threshold together in the request server, against the
same problems run one after another with the plain per-job path (mbir_job_run with
the job's own step kernel).  served: the same mix plus 5% 1024x128x512 jobs arriving
as a Poisson stream (4 jobs/s) into the server; time-to-converged p50/p99 against a
first-come-first-served queue of plain per-job runs (their measured service times).
roofline: the ICD line-update kernel (k_icd_step), its algorithmic bytes per launch
over its average launch duration measured with HIP events on the engine's stream.  cpu_baseline: the FP64 C oracle's sweep with the
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


# Algorithmic HBM bytes of the line update per voxel visit (float32): an updated voxel
# reads e and w for its theta sums (8 B per column entry and detector row) and
# read-modify-writes e (8 B), plus its own read and write (8 B); a zero-skipped voxel is
# only read (4 B).  The column table and the neighbour lines (cache-resident) are not
# counted.
BYTES_MODEL = "16 B per column entry and detector row of an updated voxel + 8 B per updated voxel + 4 B per zero-skipped voxel"


def algorithmic_bytes(nnz_updated, updated, zero_skipped):
    return 16.0 * nnz_updated + 8.0 * updated + 4.0 * zero_skipped


def traffic_probe(args):
    """Run under rocprofv3 --pmc by measure_traffic: a 512 MiB device copy (the counter's
    scale), then three sweeps of the workload's job."""
    import torch

    from mbirtem import capi, workloads

    n = (512 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b
    ctx = capi.Context(0)
    w = workloads.make(ctx, workloads.WORKLOADS[args.workload])
    for it in range(3):
        w["job"].sweep(it, 0)
    w["job"].close()
    ctx.close()


def _pmc_rows(d):
    import csv
    out = []
    for root, _, files in os.walk(d):
        for f in files:
            if f.endswith("counter_collection.csv"):
                with open(os.path.join(root, f)) as fh:
                    out.extend(csv.DictReader(fh))
    return out


def measure_traffic(workload):
    """HBM bytes per k_icd_step launch, measured now: two rocprofv3 --pmc passes
    (FETCH_SIZE, then WRITE_SIZE) over a child process running traffic_probe, each
    counter scaled by its reading on the 512 MiB copy of the same process."""
    import shutil
    import subprocess
    import tempfile

    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        return None, "rocprofv3 not found"
    per_launch = {}
    factors = {}
    for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="mbir_pmc_", dir="/tmp")
        cmd = [prof, "--pmc", ctr, "--output-format", "csv", "-d", d, "-o", "run", "--",
               sys.executable, os.path.abspath(__file__), "--traffic-probe", "--workload", workload]
        try:
            r = subprocess.run(cmd, capture_output=True, text=True, timeout=240,
                               env=dict(os.environ, TMPDIR="/tmp"))
        except subprocess.TimeoutExpired:
            return None, f"rocprofv3 --pmc {ctr} timed out"
        rows = _pmc_rows(d)
        shutil.rmtree(d, ignore_errors=True)
        if r.returncode != 0 and not rows:
            return None, f"rocprofv3 --pmc {ctr} failed ({r.returncode})"
        disp = {}
        for row in rows:
            if row.get("Counter_Name") != ctr:
                continue
            k = (row.get("Kernel_Name", ""), row.get("Dispatch_Id", ""))
            disp[k] = disp.get(k, 0.0) + float(row.get("Counter_Value", 0) or 0)
        icd = [v for (name, _), v in disp.items() if "k_icd_step" in name]
        # the copy: a kernel named for copying, else the run's largest mover of bytes
        cp = sorted(v for (name, _), v in disp.items() if "copy" in name.lower())
        big = cp or sorted(v for (name, _), v in disp.items() if "mbir::" not in name)
        if not icd or not big:
            return None, f"no {ctr} rows for the kernels"
        copy_kib = big[-1]
        factors[ctr] = copy_kib * 1024.0 / (512 << 20)
        per_launch[ctr] = sum(icd) / len(icd) * 1024.0
    bytes_per_launch = (per_launch["FETCH_SIZE"] / factors["FETCH_SIZE"] +
                        per_launch["WRITE_SIZE"] / factors["WRITE_SIZE"])
    note = (f"rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE child runs of this bench (3 sweeps), "
            f"scaled by a 512 MiB copy: factors {factors['FETCH_SIZE']:.3f} / "
            f"{factors['WRITE_SIZE']:.3f}")
    return round(bytes_per_launch, 1), note


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
    ap.add_argument("--slabs", type=int, default=48,
                    help="independent slab problems for the batched and served figures (0 = skip)")
    ap.add_argument("--arrival-rate", type=float, default=4.0)
    ap.add_argument("--big-frac", type=float, default=0.05)
    ap.add_argument("--slab-only-batched", action="store_true",
                    help="only the batched run of the slab problems (no plain runs, no served stream)")
    ap.add_argument("--no-bf", action="store_true", help="skip the bright-field figure")
    ap.add_argument("--bf-steps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true",
                    help="skip the float32-vs-FP64 comparison at the full workload size")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 --pmc child runs that measure HBM traffic")
    ap.add_argument("--traffic-probe", action="store_true", help=argparse.SUPPRESS)
    return ap.parse_args()


def _host_info():
    model = ""
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except Exception:
        pass
    return os.cpu_count() or 1, model


def fp32_vs_fp64(ctx, wl, w):
    """The float32 engine against the FP64 oracle from identical states at the full
    workload size: setup (error sinogram = the forward projection of the phantom, cost),
    one gain/offset + variance step, and one ICD sweep replayed by the oracle in the
    GPU's own update order (the exported schedule)."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import numpy as np

    import oracle_py as O
    from mbirtem import capi

    ang, y = w["angles"], np.asarray(w["sinogram"], np.float64)
    x0 = np.asarray(w["volume"], np.float32)
    kw = dict(w["job_kw"])
    job = capi.Job(ctx, w["geom"], ang, w["sinogram"], init_volume=x0, **kw)
    g = O.geometry(wl.n_r, wl.n_t, ang, wl.thickness)
    prm = O.params(sigma_x=kw["sigma_x"], mrf_q=kw["mrf_q"], mrf_c=kw["mrf_c"],
                   joint_estimation=wl.joint_estimation, noise_model=wl.noise_model)
    t0 = time.time()
    eng = O.Engine(g, ang, y, prm, x0=x0.astype(np.float64))
    err = {}
    er = eng.e()
    err["error_sinogram_at_setup_rel"] = float(np.max(np.abs(job.error_sinogram() - er)) /
                                               np.max(np.abs(y)))
    err["cost_at_setup_rel"] = abs(job.cost() - eng.cost()) / abs(eng.cost())
    eng.joint_estimation()
    eng.update_weights()
    job.nuisance_update()
    I0r, mur, alr = eng.nuisance()
    I0g, mug, alg = job.nuisance()
    err["gain_step_rel"] = float(np.max(np.abs(I0g - I0r) / np.abs(I0r)))
    err["offset_step_abs"] = float(np.max(np.abs(mug - mur)))
    err["variance_step_rel"] = float(np.max(np.abs(alg - alr) / alr))
    xs = job.volume().astype(np.float64)
    job.sweep(0, 0)
    steps = job.schedule()
    eng.update_units(np.concatenate(steps), 0, 0)
    xr, xg = eng.x(), job.volume().astype(np.float64)
    d = np.abs(xg - xr)
    err["sweep_volume_max_rel"] = float(d.max() / np.abs(xr).max())
    err["sweep_volume_mean_rel"] = float(d.mean() / np.abs(xr).max())
    err["cost_after_sweep_rel"] = abs(job.cost() - eng.cost()) / abs(eng.cost())
    job.close()
    log(f"fp32 vs fp64 ({time.time() - t0:.1f}s): {err}")
    del xs
    return err


def cpu_baseline(wl, ang, y, threads, ctx=None):
    """The FP64 oracle (a C restatement of the reference's sweep, with its y-slab thread
    decomposition) timed on this host, bounded samples:
    - the configured workload: one steady-state sweep on `threads` threads, and the
      y-slab [0, 8) of a steady-state sweep on one thread;
    - the 128x16x64 two-sphere workload run end to end (20 sweeps) on `threads` threads
      and on one, beside the same run on the GPU."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import numpy as np

    import oracle_py as O

    nproc, model = _host_info()
    g = O.geometry(wl.n_r, wl.n_t, ang, wl.thickness)
    prm = O.params(sigma_x=0.1, mrf_q=1.2, mrf_c=1e-3, joint_estimation=wl.joint_estimation,
                   noise_model=wl.noise_model, stop_threshold=wl.stop_threshold, seed=wl.seed)
    t0 = time.time()
    eng = O.Engine(g, ang, np.asarray(y, np.float64), prm)
    log(f"cpu: oracle setup {time.time() - t0:.1f}s")
    eng.sweep(0, 0, threads=threads)  # first sweep (no zero skipping), untimed
    t0 = time.time()
    eng.sweep(1, 0, threads=threads)
    dt = time.time() - t0
    V = g.N_x * g.N_y * g.N_z
    log(f"cpu: steady sweep {dt:.2f}s on {threads} threads, {V / dt:.3g} visits/s")
    ys = min(8, g.N_y)
    t0 = time.time()
    eng.sweep_slab(2, 0, 0, ys)
    dt1 = time.time() - t0
    v1 = g.N_x * ys * g.N_z / dt1
    log(f"cpu: slab [0, {ys}) of a steady sweep on 1 thread {dt1:.2f}s, {v1:.3g} visits/s")
    out = {"value": V / dt, "unit": "voxel-updates/s", "cores": threads, "kind": "port",
           "sample": f"one steady-state ICD sweep (second sweep, zero skipping on) of the FP64 "
                     f"oracle over the full {g.N_x}x{g.N_y}x{g.N_z} volume, {threads} y-slab "
                     f"threads (the reference's TBB split)",
           "host_nproc": nproc, "host_cpu": model,
           "one_core": {"value": v1, "unit": "voxel-updates/s", "cores": 1,
                        "sample": f"the y-slab [0, {ys}) of a steady-state sweep "
                                  f"({g.N_x}x{ys}x{g.N_z} voxel visits) on one thread"}}
    del eng
    # the small two-sphere workload end to end, CPU and GPU
    from mbirtem import workloads
    s1 = workloads.HAADF_128x16x64
    d1 = workloads.make(ctx, s1)
    g1 = O.geometry(s1.n_r, s1.n_t, d1["angles"], s1.thickness)
    runs = {}
    for nt in (threads, 1):
        e1 = O.Engine(g1, d1["angles"], np.asarray(d1["sinogram"], np.float64),
                      O.params(sigma_x=0.1, mrf_q=1.2, mrf_c=1e-3, num_outer=1, num_inner=20,
                               stop_threshold=s1.stop_threshold, compute_cost=0, seed=1))
        t0 = time.time()
        n = 0
        for it in range(20):
            st, _ = e1.sweep(0, it, threads=nt)
            n += 1
            if st == 0:
                break
        runs[nt] = (time.time() - t0, n)
    t0 = time.time()
    rs, _ = d1["job"].run(cap=0)
    tg = time.time() - t0
    d1["job"].close()
    V1 = s1.n_r * s1.n_t * int(2 * s1.thickness)
    out["haadf_128x16x64_61v"] = {
        "cpu_seconds": runs[threads][0], "cpu_sweeps": runs[threads][1], "cpu_threads": threads,
        "cpu_voxel_updates_per_s": V1 * runs[threads][1] / runs[threads][0],
        "cpu_1core_seconds": runs[1][0], "cpu_1core_voxel_updates_per_s": V1 * runs[1][1] / runs[1][0],
        "gpu_seconds": tg, "gpu_sweeps": rs.sweeps,
        "gpu_voxel_updates_per_s": V1 * rs.sweeps / tg}
    log(f"cpu: 128x16x64 run {out['haadf_128x16x64_61v']}")
    return out


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
    # the batched line kernel on this batch: algorithmic bytes of one more sweep over
    # the device time of its launches (HIP events around the sweep's launches)
    jobs[0].set_timing(True)
    sts = capi.batch_sweep(ctx, jobs, inner=[2 + steps] * n)
    nbytes = sum(algorithmic_bytes(st.nnz_updated, st.updated, st.zero_skipped) for st in sts)
    ms = sts[0].kernel_ms
    achieved = nbytes / (ms * 1e-3) / 1e9
    for j in jobs:
        j.close()
    log(f"batched: {n} jobs x {steps} steps in {dt:.3f}s -> {val:.4g} visits/s; "
        f"kernel {achieved:.0f} GB/s")
    return {"jobs": n, "value": val, "unit": "voxel-updates/s", "ms_per_step": 1e3 * dt / steps,
            "vs_single_job": val / single_value,
            "roofline": {"kernel": "k_icd_batch_step", "bound": "hbm", "achieved": round(achieved, 1),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBS, 4), "bytes_model": BYTES_MODEL,
                         "bytes_per_sweep": nbytes, "kernel_ms_per_sweep": ms,
                         "timing": "HIP events around one batched sweep's launches"}}


def _pct(v, q):
    import numpy as np
    return float(np.percentile(np.asarray(v, np.float64), q))


def _rmse(a, b):
    import numpy as np
    d = np.asarray(a, np.float64) - np.asarray(b, np.float64)
    return float(np.sqrt(np.mean(d * d)))


def plain_runs(ctx, inputs):
    """Each problem alone with the plain per-job path: setup + mbir_job_run (own step
    kernel, one launch per conflict-free step).  Returns service time, sweeps, RMSE."""
    from mbirtem import capi
    out = []
    for w in inputs:
        t0 = time.perf_counter()
        job = capi.Job(ctx, w["geom"], w["angles"], w["sinogram"], **w["job_kw"])
        rs, _ = job.run(cap=0)
        t1 = time.perf_counter()
        vol = job.volume()
        job.close()
        g = w["geom"]
        out.append({"service": t1 - t0, "sweeps": rs.sweeps, "rmse": _rmse(vol, w["volume"]),
                    "visits": g.N_x * g.N_y * g.N_z * rs.sweeps})
    return out


def server_runs(ctx, inputs, arrivals):
    """The problems submitted to the request server at the given arrival times (s)."""
    import threading

    from mbirtem import capi
    srv = capi.Server(ctx, max_jobs=min(256, len(inputs)), setup_threads=4)
    order = sorted(range(len(inputs)), key=lambda i: arrivals[i])
    tickets = {}
    base = time.perf_counter()

    def feeder():
        for i in order:
            dt = base + arrivals[i] - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            w = inputs[i]
            tickets[srv.submit(w["geom"], w["angles"], w["sinogram"], **w["job_kw"])] = i

    th = threading.Thread(target=feeder)
    th.start()
    res = [None] * len(inputs)
    for _ in range(len(inputs)):
        t = srv.wait_any(timeout=600.0)
        if t is None:
            raise RuntimeError("request server: no job finished within 600 s")
        while t not in tickets:
            time.sleep(0.001)
        i = tickets[t]
        rep = srv.wait(t, timeout=0.0)
        if rep.state != capi.JOB_DONE:
            raise RuntimeError(f"request server: job {i} failed with {rep.error}")
        vol = srv.job(t).volume()
        srv.release(t)
        g = inputs[i]["geom"]
        res[i] = {"latency": rep.t_done - rep.t_submit, "t_submit": rep.t_submit,
                  "t_done": rep.t_done, "sweeps": rep.sweeps, "rmse": _rmse(vol, inputs[i]["volume"]),
                  "visits": g.N_x * g.N_y * g.N_z * rep.sweeps,
                  "queue": rep.t_admit - rep.t_submit, "setup": rep.t_start - rep.t_admit}
    th.join()
    st = srv.stats()
    srv.close()
    return res, st


def tune_on(ctx, inputs, k=16):
    """Autotune the batched line kernels' form on the first k problems (each job runs
    two sweeps, one per form); returns the choice and the timings for the JSON line."""
    from mbirtem import capi
    jobs = [capi.Job(ctx, w["geom"], w["angles"], w["sinogram"], **dict(w["job_kw"], kernel_mode=1))
            for w in inputs[:k]]
    rep = capi.autotune(ctx, jobs, sweeps=1)
    for j in jobs:
        j.close()
    names = ["<=256", "<=512", "<=1024", "<=2048", ">2048"]
    out = {}
    for a in range(2):
        for b in range(2):
            for q in range(5):
                if rep.launches[a][b][q][0] and rep.launches[a][b][q][1]:
                    out[f"ypl{a + 1}_nv{b + 1}_wg{names[q]}"] = {
                        "prefetch": int(rep.prefetch[a][b][q]),
                        "us_per_wg": [round(rep.us_per_wg[a][b][q][0], 4),
                                      round(rep.us_per_wg[a][b][q][1], 4)]}
    log(f"autotune on {len(jobs)} slab problems: {out}")
    return out


def batched_roofline(ctx, inputs):
    """The batched line kernel over all the slab problems at once (mbir_batch_sweep: each
    launch carries one conflict-free step of every job): algorithmic bytes of two steady
    sweeps over the device time of their launches (HIP events around each batch sweep)."""
    from mbirtem import capi
    jobs = [capi.Job(ctx, w["geom"], w["angles"], w["sinogram"], **w["job_kw"]) for w in inputs]
    n = len(jobs)
    capi.batch_sweep(ctx, jobs, inner=[0] * n)
    jobs[0].set_timing(True)
    nbytes, ms, launches = 0.0, 0.0, 0
    for k in range(2):
        sts = capi.batch_sweep(ctx, jobs, inner=[1 + k] * n)
        nbytes += sum(algorithmic_bytes(st.nnz_updated, st.updated, st.zero_skipped) for st in sts)
        ms += sts[0].kernel_ms
        launches += max(st.steps for st in sts)
    for j in jobs:
        j.close()
    achieved = nbytes / (ms * 1e-3) / 1e9
    log(f"batched roofline: {nbytes / 1e9:.2f} GB in {ms:.1f} ms -> {achieved:.0f} GB/s")
    return {"kernel": "k_icd_batch_step", "bound": "hbm", "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "bytes_model": BYTES_MODEL, "bytes_per_sweep": nbytes / 2, "kernel_ms_per_sweep": ms / 2,
            "steps_per_sweep": launches // 2,
            "timing": "HIP events around each batched sweep's launches"}


def slab_figures(ctx, n, rate, big_frac, only_batched=False):
    """batched_slabs (all n problems at once) and served (Poisson arrivals) figures."""
    import numpy as np

    from mbirtem import workloads
    specs = workloads.slab_mix(n, seed=7)
    if only_batched:  # the server's launches alone (for profiling)
        inputs = [workloads.make_slab(ctx, s) for s in specs]
        t0 = time.perf_counter()
        res, st = server_runs(ctx, inputs, [0.0] * n)
        dt = time.perf_counter() - t0
        visits = sum(r["visits"] for r in res)
        log(f"batched slabs only: {n} jobs in {dt:.2f}s, {visits / dt:.4g} visits/s")
        return {"jobs": n, "value": visits / dt, "unit": "voxel-updates/s",
                "launches": int(st.launches)}, None
    rng = np.random.default_rng(11)
    served_specs = [workloads.slab_mix(1, seed=1000 + i, big_frac=1.0)[0]
                    if rng.uniform() < big_frac else s for i, s in enumerate(specs)]
    t0 = time.time()
    cache = {}
    for s in specs + served_specs:
        if s.name not in cache:
            cache[s.name] = workloads.make_slab(ctx, s)
            if len(cache) % 8 == 0:
                log(f"slabs: {len(cache)} problems prepared ({time.time() - t0:.1f}s)")
    log(f"slabs: {len(cache)} problems prepared in {time.time() - t0:.1f}s")
    plain = {}
    t0 = time.time()
    for k, (name, w) in enumerate(cache.items()):
        plain[name] = plain_runs(ctx, [w])[0]
        if k % 8 == 7:
            log(f"slabs: {k + 1} plain runs done ({time.time() - t0:.1f}s)")
    log(f"slabs: plain per-job runs {time.time() - t0:.1f}s")

    # batched: every problem submitted at t = 0, with the kernel forms autotuned first
    inputs = [cache[s.name] for s in specs]
    tuning = tune_on(ctx, inputs)
    res, st = server_runs(ctx, inputs, [0.0] * n)
    p = [plain[s.name] for s in specs]
    for a, b in zip(res, p):
        if a["sweeps"] != b["sweeps"] or a["rmse"] != b["rmse"]:
            raise RuntimeError("request server result differs from the plain per-job run")
    make = max(r["t_done"] for r in res) - min(r["t_submit"] for r in res)
    visits = sum(r["visits"] for r in res)
    plain_total = sum(q["service"] for q in p)
    plain_lat = np.cumsum([q["service"] for q in p])  # one after another, submitted at once
    lat = [r["latency"] for r in res]
    batched = {
        "jobs": n, "value": visits / make, "unit": "voxel-updates/s",
        "plain_value": visits / plain_total, "vs_plain": plain_total / make,
        "makespan_s": make, "plain_total_s": plain_total,
        "time_to_converged_p50_s": _pct(lat, 50), "time_to_converged_p99_s": _pct(lat, 99),
        "plain_time_to_converged_p50_s": _pct(plain_lat, 50),
        "plain_time_to_converged_p99_s": _pct(plain_lat, 99),
        "p99_vs_plain": _pct(lat, 99) / _pct(plain_lat, 99),
        "rmse_vs_phantom_mean": float(np.mean([r["rmse"] for r in res])),
        "final_rmse_equal_to_plain": True, "sweeps_total": int(sum(r["sweeps"] for r in res)),
        "launches": int(st.launches), "entries_per_launch": st.entries / max(1, st.launches),
        "peak_resident": int(st.peak_resident), "peak_device_gb": st.peak_reserved_bytes / 1e9,
        "autotuned_forms": tuning,
    }
    batched["roofline"] = batched_roofline(ctx, inputs)
    log(f"batched slabs: {batched['value']:.4g} visits/s vs plain {batched['plain_value']:.4g} "
        f"({batched['vs_plain']:.2f}x), p99 {batched['time_to_converged_p99_s']:.2f}s vs "
        f"{batched['plain_time_to_converged_p99_s']:.2f}s")

    # served: Poisson arrivals
    inputs = [cache[s.name] for s in served_specs]
    arr = np.cumsum(rng.exponential(1.0 / rate, len(inputs)))
    arr -= arr[0]
    res, st = server_runs(ctx, inputs, list(arr))
    p = [plain[s.name] for s in served_specs]
    for a, b in zip(res, p):
        if a["sweeps"] != b["sweeps"] or a["rmse"] != b["rmse"]:
            raise RuntimeError("request server result differs from the plain per-job run")
    fifo, done = [], 0.0
    for a, q in zip(arr, p):  # first come, first served, one job at a time
        done = max(done, a) + q["service"]
        fifo.append(done - a)
    lat = [r["latency"] for r in res]
    span = max(r["t_done"] for r in res) - min(r["t_submit"] for r in res)
    served = {
        "jobs": len(inputs), "arrival_rate_per_s": rate,
        "big_jobs": sum(1 for s in served_specs if s.n_r == 1024 and s.n_t == 128),
        "time_to_converged_p50_s": _pct(lat, 50), "time_to_converged_p99_s": _pct(lat, 99),
        "plain_fifo_p50_s": _pct(fifo, 50), "plain_fifo_p99_s": _pct(fifo, 99),
        "p99_vs_plain": _pct(lat, 99) / _pct(fifo, 99),
        "value": sum(r["visits"] for r in res) / span, "unit": "voxel-updates/s",
        "queue_wait_p99_s": _pct([r["queue"] for r in res], 99),
        "setup_p50_s": _pct([r["setup"] for r in res], 50),
        "final_rmse_equal_to_plain": True,
        "peak_resident": int(st.peak_resident), "peak_device_gb": st.peak_reserved_bytes / 1e9,
    }
    log(f"served: p50 {served['time_to_converged_p50_s']:.2f}s p99 "
        f"{served['time_to_converged_p99_s']:.2f}s vs plain FIFO p99 {served['plain_fifo_p99_s']:.2f}s")
    return batched, served


def bf_figures(ctx, steps):
    """Bright field 512x64x256, 121 views: throughput of sweep + offset + variance steps,
    and the float32 engine's error against the FP64 oracle from identical states on
    the 128x16x64, 61-view bright-field sample."""
    import numpy as np

    from mbirtem import capi, workloads
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_py as O

    wl = workloads.BF_512x64x256
    t0 = time.time()
    w = workloads.make_bf(ctx, wl)
    job, g = w["job"], w["geom"]
    V = g.N_x * g.N_y * g.N_z
    log(f"bf: setup {time.time() - t0:.1f}s")
    job.sweep(0, 0)
    job.sweep(0, 1)
    job.set_bragg(workloads.BF_KW["bragg_T"])
    t0 = time.perf_counter()
    for k in range(steps):
        job.sweep(1 + k, 0)
        job.nuisance_update()
    dt = time.perf_counter() - t0
    value = V * steps / dt
    log(f"bf: {steps} steps in {dt:.3f}s -> {value:.4g} visits/s")
    job.close()
    # NHICD calls, the reference engine's bright-field schedule (BrightField/BFConstants.h:51
    # defines NHICD): 42 calls (21 homogeneous, 21 non-homogeneous) from the start
    kw_n = dict(w["job_kw"], bf_nhicd=1, num_outer=1, num_inner=42, stop_threshold=0.0,
                joint_estimation=0, noise_model=0)
    jn = capi.Job(ctx, w["geom"], w["angles"], w["counts"], **kw_n)
    jn.set_nhicd_log(True)
    t0 = time.perf_counter()
    rs_n, _ = jn.run(cap=0)  # no per-call cost: the reference computes it in debug builds only
    dt_n = time.perf_counter() - t0
    lines_n = sum(len(c) for c in jn.nhicd_lists())
    jn.close()
    nhicd = {"calls": int(rs_n.sweeps), "ms_per_call": 1e3 * dt_n / rs_n.sweeps,
             "lines_per_call": lines_n / rs_n.sweeps, "voxel_updates_per_s": lines_n * g.N_y / dt_n}
    log(f"bf: NHICD {nhicd}")

    # float32 engine vs FP64 oracle from identical states
    s = workloads.BF_128x16x64
    ws = workloads.make_bf(ctx, s)
    gs = O.geometry(s.n_r, s.n_t, ws["angles"], s.thickness)
    kw = ws["job_kw"]
    prm = O.params(bf_mode=1, mrf_c=kw["mrf_c"], sigma_x=kw["sigma_x"], mrf_q=kw["mrf_q"],
                   bragg_T=kw["bragg_T"], bragg_delta=kw["bragg_delta"], joint_estimation=1,
                   noise_model=1, num_outer=s.num_outer, num_inner=s.num_inner,
                   stop_threshold=s.stop_threshold, seed=s.seed)
    x = ws["volume"].astype(np.float64)
    offs = np.full(s.n_views, -np.log(s.dose))
    eng = O.Engine(gs, ws["angles"], ws["counts"].astype(np.float64), prm, x0=x, offsets=offs)
    jb = capi.Job(ctx, ws["geom"], ws["angles"], ws["counts"], init_volume=ws["volume"],
                  init_offset=offs, **{k: v for k, v in kw.items() if k != "default_offset"})
    pr = eng.project(x)
    pg = jb.forward_project(ws["volume"])
    err = {"projection_rel": float(np.max(np.abs(pg - pr)) / np.max(np.abs(pr))),
           "cost_rel_at_setup": abs(jb.cost() - eng.cost()) / abs(eng.cost())}
    eng.bragg(kw["bragg_T"])
    jb.set_bragg(kw["bragg_T"])
    eng.joint_estimation()
    eng.update_selector()
    eng.update_weights()
    eng.update_selector()
    jb.nuisance_update()
    _, mu, al = eng.nuisance()
    _, mug, alg = jb.nuisance()
    err["offset_step_abs"] = float(np.max(np.abs(mug - mu)))
    err["variance_step_rel"] = float(np.max(np.abs(alg - al) / al))
    err["cost_rel_after_steps"] = abs(jb.cost() - eng.cost()) / abs(eng.cost())
    jb.close()
    # whole runs (the update orders differ): volume RMS difference relative to the
    # oracle's own seed-to-seed difference
    e1 = O.Engine(gs, ws["angles"], ws["counts"].astype(np.float64), prm, offsets=offs)
    e1.run()
    prm2 = O.params(**{f: getattr(prm, f) for f, _ in O.Params._fields_})
    prm2.seed = s.seed + 1
    e2 = O.Engine(gs, ws["angles"], ws["counts"].astype(np.float64), prm2, offsets=offs)
    e2.run()
    jr = capi.Job(ctx, ws["geom"], ws["angles"], ws["counts"], init_offset=offs,
                  **{k: v for k, v in kw.items() if k != "default_offset"})
    jr.run(cap=0)
    xr = e1.x()
    rms = lambda a: float(np.sqrt(np.mean(a * a)))
    err["run_volume_rel_rms"] = rms(jr.volume() - xr) / rms(xr)
    err["oracle_seed_spread_rel_rms"] = rms(e2.x() - xr) / rms(xr)
    jr.close()
    ws["job"].close()
    log(f"bf: fp32 vs fp64 {err}")
    return {"workload": wl.name, "value": value, "unit": "voxel-updates/s",
            "ms_per_step": 1e3 * dt / steps,
            "step": "ICD sweep + offset (log dose) + shared variance update",
            "nhicd": nhicd, "fp32_vs_fp64": err, "fp32_vs_fp64_sample": s.name}


def main():
    args = parse()
    if args.traffic_probe:
        traffic_probe(args)
        return
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
        rl_bytes += algorithmic_bytes(st.nnz_updated, st.updated, st.zero_skipped)
        rl_ms += st.kernel_ms
        rl_launches += st.steps
    job.set_timing(False)
    achieved = rl_bytes / (rl_ms * 1e-3) / 1e9
    traffic, traffic_note = None, "not measured (--no-traffic)"
    if not args.no_traffic:
        traffic, traffic_note = measure_traffic(wl.name)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "k_icd_step", "avg_launch_us": round(1e3 * rl_ms / rl_launches, 3),
                "bytes_per_launch": round(rl_bytes / rl_launches, 1),
                "bytes_model": BYTES_MODEL, "timing": "HIP events around each sweep's launches "
                "on the job's stream, divided by the launches", "traffic_source": traffic_note}
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
    if args.slabs > 0:
        out["batched_slabs"], out["served"] = slab_figures(ctx, args.slabs, args.arrival_rate,
                                                           args.big_frac, args.slab_only_batched)
    if not args.no_fp64:
        out["fp32_vs_fp64"] = fp32_vs_fp64(ctx, wl, w)
        out["fp32_vs_fp64_sample"] = (f"{wl.name} from identical states (the phantom volume "
                                      "and its nuisance parameters), one sweep in the GPU's order")
    if not args.no_bf:
        out["bright_field"] = bf_figures(ctx, args.bf_steps)
    if not args.no_cpu:
        # every core of the host's share (16 per GPU on the benchmark boxes)
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(wl, w["angles"], w["sinogram"], threads, ctx)
    else:
        out["cpu_baseline"] = None
    job.close()
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
