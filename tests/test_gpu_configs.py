"""GPU: the benchmark configurations themselves.

* 128 x 16 x 64, two spheres, 61 views over +-60 degrees, 20 ICD sweeps: the GPU run
  against the FP64 oracle's run of the same sinogram, within three times the oracle's own
  seed-to-seed spread (volume) and 1e-3 (final cost).
* 512 x 64 x 256, 91 views over +-70 degrees, joint gain/offset and noise estimation (too
  large for the oracle's whole run): size-independent properties at full size: the cost
  never rises over sweeps and the gain/offset step, the gains keep their mean at the
  target, the error sinogram stays y - I0 A x - mu for the GPU's own volume.
* bright field 512 x 64 x 256, 121 views: the cost never rises at a fixed threshold,
  e stays -log(counts) - A x - mu, and the selector is the rule |e sqrt(w)| < T.
* the batched and served mixes, reduced: three slabs of the 48-slab mix and one
  1024 x 128 x 512 job in the request server end bit for bit as each alone.
"""
import numpy as np
import pytest

import oracle_py as O
from mbirtem import capi, workloads as W

pytestmark = pytest.mark.gpu


def _rms(a):
    return float(np.sqrt(np.mean(np.asarray(a, np.float64) ** 2)))


def test_config_128x16x64_two_spheres_matches_oracle(gpu_ctx):
    wl = W.HAADF_128x16x64
    d = W.make(gpu_ctx, wl)
    job, y, ang = d["job"], d["sinogram"], d["angles"]
    kw = d["job_kw"]
    assert (job.geom.N_x, job.geom.N_y, job.geom.N_z) == (128, 16, 64)
    g = O.geometry(wl.n_r, wl.n_t, ang, wl.thickness)
    # exactly 20 sweeps on both sides (no early stop: the stop test fires at different
    # sweeps for different update orders)
    runs = []
    for seed in (1, 2, 3):
        eng = O.Engine(g, ang, np.asarray(y, np.float64),
                       O.params(sigma_x=kw["sigma_x"], mrf_q=kw["mrf_q"], mrf_c=kw["mrf_c"],
                                stop_threshold=0.0, num_outer=1, num_inner=20, seed=seed))
        costs, n = eng.run()
        assert n == 20
        runs.append((eng.x(), costs[-1], n))
    job.close()
    job = capi.Job(gpu_ctx, d["geom"], ang, y, **dict(kw, stop_threshold=0.0))
    rs, costs_g = job.run()
    assert rs.sweeps == 20
    c = np.asarray(costs_g)
    assert np.all(np.diff(c) <= 1e-6 * np.abs(c[1:]))  # every sweep lowers the cost
    xg = job.volume().astype(np.float64)
    xr = runs[0][0]
    spread = max(_rms(r[0] - xr) for r in runs[1:])
    assert _rms(xg - xr) <= max(3 * spread, 1e-3 * _rms(xr)), (_rms(xg - xr), spread)
    cr = np.array([r[1] for r in runs])
    assert abs(c[-1] - cr.mean()) <= max(3 * (cr.max() - cr.min()), 1e-4 * cr.mean()), (c[-1], cr)
    # against the phantom the reconstruction is as good as the oracle's runs
    assert _rms(xg - d["volume"]) <= 1.5 * max(_rms(r[0] - d["volume"]) for r in runs)
    job.close()


def test_config_512x64x256_haadf_properties(gpu_ctx):
    wl = W.HAADF_512x64x256
    d = W.make(gpu_ctx, wl)
    job, y = d["job"], np.asarray(d["sinogram"], np.float64)
    g = job.geom
    assert (g.N_x, g.N_y, g.N_z, g.N_theta) == (512, 64, 256, 91)
    costs = [job.cost()]
    for it in range(3):
        job.sweep(0, it)
        costs.append(job.cost())
    c = np.asarray(costs)
    assert np.all(np.diff(c) <= 1e-6 * np.abs(c[1:])), c
    # the gain/offset step minimises the weighted data term over gains and offsets
    # (with the mean gain fixed): it cannot raise it
    data0 = job.cost(terms=True)[1][0]
    job.nuisance_update()
    I0, mu, al = job.nuisance()
    assert np.mean(I0) == pytest.approx(1.0, abs=1e-9)
    assert np.all(al > 0)
    # e == y - I0 A x - mu for the GPU's own volume (the forward projection carries I0)
    yest = job.forward_project(job.volume()).reshape(g.N_theta, g.N_r, g.N_t)
    e_ref = y.reshape(yest.shape) - yest - mu[:, None, None]
    err = np.max(np.abs(job.error_sinogram().reshape(yest.shape) - e_ref))
    assert err <= 2e-4 * np.abs(y).max(), err
    # the noise step sets alpha to the mean squared error of each view
    e = job.error_sinogram().reshape(yest.shape).astype(np.float64)
    np.testing.assert_allclose(al, np.mean(e * e, axis=(1, 2)), rtol=1e-5)
    assert data0 > 0
    job.close()


def test_config_512x64x256_bright_field_properties(gpu_ctx):
    wl = W.BF_512x64x256
    d = W.make_bf(gpu_ctx, wl)
    job, counts = d["job"], np.asarray(d["counts"], np.float64)
    g = job.geom
    assert (g.N_x, g.N_y, g.N_z, g.N_theta) == (512, 64, 256, 121)
    job.set_bragg(3.0)
    costs = [job.cost()]
    for it in range(3):  # the first sweep updates (and refreshes the selector of) all
        job.sweep(0, it)
        costs.append(job.cost())
    c = np.asarray(costs)
    assert np.all(np.diff(c) <= 1e-6 * np.abs(c[1:])), c
    shape = (g.N_theta, g.N_r, g.N_t)
    # the offset and shared-variance step refreshes every entry's selector (detector
    # bins that no voxel projects onto keep theirs from the setup until then)
    job.nuisance_update()
    _, mu, _ = job.nuisance()
    yest = job.forward_project(job.volume()).reshape(shape)
    e_ref = -np.log(counts.reshape(shape)) - yest - mu[:, None, None]
    e = job.error_sinogram().reshape(shape)
    assert np.max(np.abs(e - e_ref)) <= 2e-4 * np.abs(e_ref).max()
    w, sel = job.weights().reshape(shape), job.selector().reshape(shape)
    rule = (np.abs(e * np.sqrt(w)) < 3.0).astype(np.uint8)
    assert np.mean(sel != rule) < 1e-3
    assert sel.mean() < 1.0  # the threshold rejects entries at this dose
    job.close()


def test_served_mix_with_a_large_job_equals_alone(gpu_ctx):
    slabs = W.slab_mix(48, seed=7)[:3]
    slabs.append(W.Slab("big_1024x128x512_61v", 1024, 128, 61, 65.0, 256.0, 12345))
    inputs = [W.make_slab(gpu_ctx, s) for s in slabs]
    alone = []
    for w in inputs:
        j = capi.Job(gpu_ctx, w["geom"], w["angles"], w["sinogram"], **w["job_kw"])
        rs, _ = j.run(cap=0)
        alone.append((j.volume(), j.nuisance(), rs.sweeps))
        j.close()
    srv = capi.Server(gpu_ctx, max_jobs=8)
    tickets = [srv.submit(w["geom"], w["angles"], w["sinogram"], **w["job_kw"]) for w in inputs]
    for t, (vol, nuis, sweeps) in zip(tickets, alone):
        rep = srv.wait(t, timeout=300.0)
        assert rep.state == capi.JOB_DONE, rep.error
        j = srv.job(t)
        np.testing.assert_array_equal(j.volume(), vol)
        assert rep.sweeps == sweeps
        srv.release(t)
    srv.close()
