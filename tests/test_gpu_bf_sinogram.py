"""GPU: HAADF with a bright-field sinogram (every projection entry scaled by its
detector element's bright-field counts + 32768, the reference's BF_Flag) against the
FP64 oracle, whose scaled path the CPU tests pin by identities."""
import numpy as np
import pytest

import oracle_py as O
from mbirtem import capi
from problems import small_problem

pytestmark = pytest.mark.gpu

BF_OFFSET = 32768.0


def _case(seed=12):
    g, ang, vol, y0 = small_problem(seed=seed)
    nt = len(ang)
    rng = np.random.default_rng(seed)
    # scales on a 1/256 grid: the raw counts s - 32768 are then exact in float32 (the
    # problem carries them as float32; near 32768 its spacing is 1/256)
    s = (205 + rng.integers(0, 103, (nt, g.N_r, g.N_t))) / 256.0
    y = (s * y0.reshape(nt, g.N_r, g.N_t)).reshape(y0.shape)  # measurements of the scaled model
    return g, ang, y, s - BF_OFFSET, s


def _job(ctx, g, ang, y, **kw):
    geom = capi.geometry_init(g.N_r, g.N_t, ang, thickness=g.LengthZ / 2)
    return capi.Job(ctx, geom, ang, np.asarray(y, np.float32), **kw)


def test_scaled_setup_sweeps_and_convergence_match_oracle(gpu_ctx):
    g, ang, y, bf, s = _case()
    eng = O.Engine(g, ang, y, O.params(), default_recon=0.2)
    eng.set_bf_sinogram(bf)
    job = _job(gpu_ctx, g, ang, y, bf_sinogram=bf, default_recon=0.2)
    er, eg = eng.e(), job.error_sinogram()
    assert np.max(np.abs(eg - er)) <= 1e-5 * np.abs(er).max()
    assert job.cost() == pytest.approx(eng.cost(), rel=1e-6)
    costs = [job.cost()]
    for it in range(60):
        eng.sweep(0, it)
        job.sweep(0, it)
        if it < 8:
            costs.append(job.cost())
    for a, b in zip(costs, costs[1:]):
        assert b <= a * (1 + 1e-6), costs
    # e stays y - s A x - mu for the current volume (fp32 accuracy)
    xg = job.volume()
    e_ref = y.reshape(s.shape) - s * job.forward_project(xg).reshape(s.shape)
    assert np.max(np.abs(job.error_sinogram().reshape(s.shape) - e_ref)) <= 2e-4 * np.abs(y).max()
    xr = eng.x()
    rmse = np.sqrt(np.mean((xr - xg) ** 2))
    scale = np.sqrt(np.mean(xr ** 2))
    assert rmse <= 1e-2 * scale, (rmse, scale)
    assert job.cost() == pytest.approx(eng.cost(), rel=1e-4)


def test_scaled_job_offset_step_matches_oracle(gpu_ctx):
    """Joint estimation with bright-field data estimates the offsets only
    (HAADF/HAADF_ReconstructionEngine_Extra.cpp:636-669): from the same state, one
    offset step and one variance step on the GPU equal the oracle's to 1e-6, the gains
    stay, and a whole run keeps the gains while the offsets move."""
    g, ang, y, bf, s = _case(seed=13)
    nt = len(ang)
    rng = np.random.default_rng(13)
    gains = np.full(nt, 1.0)
    offs = rng.uniform(-0.02, 0.02, nt)
    var = rng.uniform(0.5, 2.0, nt)
    kw = dict(joint_estimation=1, noise_model=1)
    eng = O.Engine(g, ang, y, O.params(**kw), gains=gains, offsets=offs, variances=var,
                   default_recon=0.2)
    eng.set_bf_sinogram(bf)
    job = _job(gpu_ctx, g, ang, y, bf_sinogram=bf, init_gain=gains, init_offset=offs,
               init_variance=var, default_recon=0.2, **kw)
    scale = np.abs(y).max()
    assert np.max(np.abs(job.error_sinogram() - eng.e())) <= 1e-5 * scale
    for step in range(2):
        eng.joint_estimation()
        eng.update_weights()
        job.nuisance_update()
        I0r, mur, alr = eng.nuisance()
        I0g, mug, alg = job.nuisance()
        np.testing.assert_array_equal(I0g, gains)
        np.testing.assert_allclose(mug, mur, rtol=0, atol=1e-6 * scale)
        np.testing.assert_allclose(alg, alr, rtol=1e-6)
        assert np.max(np.abs(job.error_sinogram() - eng.e())) <= 1e-5 * scale
    assert np.max(np.abs(mug - offs)) > 1e-3  # the offsets moved
    # after the step every view's error has zero weighted mean
    assert np.abs(job.error_sinogram().reshape(nt, -1).mean(axis=1)).max() <= 1e-6 * scale
    assert job.cost() == pytest.approx(eng.cost(), rel=1e-6)
    run = _job(gpu_ctx, g, ang, y, bf_sinogram=bf, num_outer=3, num_inner=4,
               init_gain=np.full(nt, 0.98), init_offset=offs, target_gain=10.0, **kw)
    run.run()
    I0, mu, al = run.nuisance()
    np.testing.assert_array_equal(I0, np.full(nt, 0.98))
    assert np.max(np.abs(mu - offs)) > 1e-4


def test_scaled_jobs_are_single_job_only(gpu_ctx):
    g, ang, y, bf, s = _case(seed=14)
    job = _job(gpu_ctx, g, ang, y, bf_sinogram=bf)
    with pytest.raises(capi.MbirError) as ei:
        capi.batch_sweep(gpu_ctx, [job])
    assert ei.value.code == -3
    srv = capi.Server(gpu_ctx, max_jobs=2)
    with pytest.raises(capi.MbirError) as ei:
        srv.submit(capi.geometry_init(g.N_r, g.N_t, ang, thickness=g.LengthZ / 2), ang,
                   np.asarray(y, np.float32), bf_sinogram=bf)
    assert ei.value.code == -3
    srv.close()
    job.close()


def test_adapter_reads_bright_field_tilt_series(gpu_ctx, tmp_path):
    """The C++ adapter given a bright-field MRC beside the HAADF one (the reference's
    BFTomoInputs / BFSinogram) must write exactly the C-ABI's volume for the same
    problem with the counts as bf_sinogram; a bright-field file of another size or
    type is an error."""
    import os
    import subprocess
    from test_io import make_tilt_series, parse_mrc

    cli = os.path.join(os.path.dirname(capi.LIB_PATH), "engine_cli")
    g, ang, y, bf, s = _case(seed=15)
    nv = len(ang)
    ang32 = ang.astype(np.float32)
    y32 = y.astype(np.float32).reshape(nv, g.N_r, g.N_t)
    bf32 = bf.astype(np.float32)
    assert np.array_equal(bf32.astype(np.float64), bf)  # the counts are exact in float32
    mrc, bmrc = str(tmp_path / "tilts.mrc"), str(tmp_path / "bf.mrc")
    make_tilt_series(mrc, y32.transpose(0, 2, 1), mode=2, tilts=ang32, pixel=1e-9)
    make_tilt_series(bmrc, bf32.transpose(0, 2, 1), mode=2, tilts=ang32, pixel=1e-9)
    d = float(np.float32(1e-9)) * 1e9
    th = g.LengthZ / 2
    args = [str(th), "0.1", "1.2", "4", "3", "1", "1", "7"]
    out = str(tmp_path / "out")
    os.makedirs(out)
    r = subprocess.run([cli, "--mrc", mrc] + args + [out, bmrc], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    geom = capi.geometry_init(g.N_r, g.N_t, ang32.astype(np.float64), th, delta_r=d, delta_t=d)
    job = capi.Job(gpu_ctx, geom, ang32.astype(np.float64), y32.astype(np.float64), seed=7,
                   num_outer=4, num_inner=3, joint_estimation=1, noise_model=1,
                   stop_threshold=1e-4, bf_sinogram=bf32)
    job.run()
    x = job.volume()
    nz, nx, ny = x.shape
    raw = np.fromfile(os.path.join(out, "ReconstructedObject.bin"), "<f8").reshape(ny, nx, nz)
    np.testing.assert_array_equal(raw, x.transpose(2, 1, 0).astype(np.float64))
    h, data = parse_mrc(os.path.join(out, "recon.mrc"))
    assert h["nz"] == nz
    I0, mu, _ = job.nuisance()
    np.testing.assert_array_equal(np.fromfile(os.path.join(out, "gains.bin"), "<f8"), I0)
    np.testing.assert_array_equal(np.fromfile(os.path.join(out, "offsets.bin"), "<f8"), mu)
    job.close()
    # a bright-field series with one view fewer, and one that is no MRC file
    short = str(tmp_path / "short.mrc")
    make_tilt_series(short, bf32[1:].transpose(0, 2, 1), mode=2, tilts=ang32[1:], pixel=1e-9)
    for bad in (short, str(tmp_path / "bf.bin")):
        o = str(tmp_path / ("bad_" + os.path.basename(bad)))
        os.makedirs(o)
        r = subprocess.run([cli, "--mrc", mrc] + args + [o, bad], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode != 0, bad
        assert "error -1" in open(os.path.join(o, "run.txt")).read()


def test_scaled_job_graph_replay_and_y_slabs(gpu_ctx):
    """The scaled line update in its other single-job forms: sweeps replayed from a HIP
    graph (kernel_mode 3) end bit for bit as the per-step launches, across offset steps
    (the descriptor changes: the graph is captured again); and y-slab units (two per
    line) lower the cost on every sweep and keep e = y - s A x - mu."""
    g, ang, y, bf, s = _case(seed=14)
    kw = dict(bf_sinogram=bf, default_recon=0.2, joint_estimation=1, seed=4)
    a = _job(gpu_ctx, g, ang, y, kernel_mode=0, **kw)
    b = _job(gpu_ctx, g, ang, y, kernel_mode=3, **kw)
    for outer in range(3):
        for it in range(3):
            a.sweep(outer, it)
            b.sweep(outer, it)
            np.testing.assert_array_equal(a.volume(), b.volume())
            np.testing.assert_array_equal(a.error_sinogram(), b.error_sinogram())
        a.nuisance_update()
        b.nuisance_update()
    for u, v in zip(a.nuisance(), b.nuisance()):
        np.testing.assert_array_equal(u, v)
    a.close()
    b.close()
    job = _job(gpu_ctx, g, ang, y, bf_sinogram=bf, default_recon=0.2, y_slabs=2)
    costs = [job.cost()]
    for it in range(8):
        job.sweep(0, it)
        costs.append(job.cost())
    assert all(d <= c * (1 + 1e-6) for c, d in zip(costs, costs[1:])), costs
    xg = job.volume()
    e_ref = y.reshape(s.shape) - s * job.forward_project(xg).reshape(s.shape)
    assert np.max(np.abs(job.error_sinogram().reshape(s.shape) - e_ref)) <= 2e-4 * np.abs(y).max()
    units = np.concatenate(job.schedule())
    assert units[:, 2].max() < g.N_y  # units are y-ranges, not whole lines
    job.close()
