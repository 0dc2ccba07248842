"""Serving loop on the GPU: jobs of different image sizes, subset counts,
radii and warp orders submitted together to a multi-slot server (their work
overlaps on the device, references shared from the cache) give exactly the
results each job gives alone; a tight HBM budget still completes every job
with the same results."""

import numpy as np
import pytest

from geopyv_dev import serving, workloads

pytestmark = pytest.mark.gpu


def _jobs():
    rng = np.random.Generator(np.random.PCG64(21))
    refs = {}
    jobs = []
    for k in range(10):
        shape = (192, 224) if k % 2 else (256, 256)
        if shape not in refs:
            sp = workloads.Speckle(shape, int(0.012 * shape[0] * shape[1]), 2.0, 40 + len(refs))
            refs[shape] = (sp, sp.reference(2.0))
        sp, ref = refs[shape]
        g = rng.uniform(-0.02, 0.02, 4)
        field = workloads.Affine(((shape[1] - 1) / 2, (shape[0] - 1) / 2), g[0], g[1], g[2], g[3])
        tar = sp.target(field, 2.0, rng=np.random.Generator(np.random.PCG64(100 + k)))
        nsub = int(rng.integers(20, 120))
        coords = np.column_stack([rng.uniform(45, shape[1] - 46, nsub), rng.uniform(45, shape[0] - 46, nsub)])
        coords = np.rint(coords * 4) / 4
        radii = rng.integers(12, 26, nsub)
        order = int(rng.integers(1, 3))
        truth = field.warp(coords[:, 0], coords[:, 1])
        p0 = np.zeros((nsub, 12))
        p0[:, :2] = np.rint(truth[:, :2])
        jobs.append(serving.Job(k, shape, ref, tar, coords, radii, order=order, p0=p0))
    return jobs


def _same(a, b):
    assert np.array_equal(a["p"], b["p"])
    assert np.array_equal(a["iterations"], b["iterations"])
    assert np.array_equal(a["c_zncc"], b["c_zncc"])


def test_results_alone_equal_results_in_the_mix():
    jobs = _jobs()
    alone = {}
    srv1 = serving.Server(n_slots=1)
    for j in jobs:
        alone.update(srv1.run([j]))
    srv1.close()
    srv = serving.Server(n_slots=3)
    mix = srv.run(jobs, arrivals=[0.002 * k for k in range(len(jobs))])
    srv.close()
    for j in jobs:
        _same(alone[j.job_id][0], mix[j.job_id][0])
        assert np.mean(mix[j.job_id][0]["converged"]) > 0.9
    lat = np.array([mix[j.job_id][1] for j in jobs])
    print("latency ms p50", np.percentile(lat, 50), "p99", np.percentile(lat, 99))


def test_tight_budget_admits_in_turn():
    jobs = _jobs()[:6]
    big = max(j.device_bytes() + j.image_bytes() for j in jobs)
    srv = serving.Server(n_slots=3, hbm_budget=int(1.5 * big))
    mix = srv.run(jobs)
    srv.close()
    ref = serving.Server(n_slots=1)
    for j in jobs:
        _same(ref.run([j])[j.job_id][0], mix[j.job_id][0])
    ref.close()
