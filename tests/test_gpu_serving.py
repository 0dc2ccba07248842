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


def test_served_results_match_the_oracle():
    # jobs at the serving workload's smallest image size (1024^2) with
    # 1 000-1 200 subsets of radii 15-40 and orders 1 and 2, served together
    # with two small jobs; every job's results in the mix agree with the
    # FP64 oracle's batch solve of the same subsets and initial warps
    import oracle as O

    rng = np.random.Generator(np.random.PCG64(77))
    jobs, cases = [], []
    for k, (shape, nsub, order) in enumerate([((1024, 1024), 1200, 1), ((1024, 1024), 1000, 2),
                                              ((256, 256), 60, 2), ((256, 256), 80, 1)]):
        sp = workloads.Speckle(shape, int(round(0.012 * shape[0] * shape[1])), 2.0, 300 + k)
        ref = sp.reference(2.0)
        g = rng.uniform(-0.02, 0.02, 4)
        field = workloads.Affine(((shape[1] - 1) / 2, (shape[0] - 1) / 2), g[0], g[1], g[2], g[3])
        tar = sp.target(field, 2.0, rng=np.random.Generator(np.random.PCG64(400 + k)))
        rmax = 40 if shape[0] >= 1024 else 25
        radii = rng.integers(15, rmax + 1, nsub)
        margin = rmax + 0.02 * shape[0] / 2 + 8
        coords = np.column_stack([rng.uniform(margin, shape[1] - 1 - margin, nsub),
                                  rng.uniform(margin, shape[0] - 1 - margin, nsub)])
        p0 = np.zeros((nsub, 12))
        p0[:, :2] = np.rint(field.warp(coords[:, 0], coords[:, 1])[:, :2])
        jobs.append(serving.Job(k, ("orc", k), ref, tar, coords, radii, order=order, p0=p0))
        cases.append((ref, tar, coords, radii, order, p0))
    srv = serving.Server(n_slots=3)
    print("serving", [(len(c[2]), c[4], c[0].shape) for c in cases], flush=True)
    mix = srv.run(jobs, arrivals=[0.0, 0.001, 0.002, 0.003])
    srv.close()
    for j, (ref, tar, coords, radii, order, p0) in zip(jobs, cases):
        r = mix[j.job_id][0]
        u = np.unique(radii)
        lists = [O.mask_local("circle", int(x))[0] for x in u]
        start = np.concatenate([[0], np.cumsum([len(a) for a in lists])]).astype(np.int64)
        index = np.searchsorted(u, radii).astype(np.int32)
        orc = O.batch_solve(O.get_c(ref.astype(np.float64), 20), O.get_c(tar.astype(np.float64), 20), ref.shape,
                            20, coords, index, np.concatenate(lists), start, np.full(len(coords), order, np.int32),
                            p0, max_norm=1e-5, max_iterations=50)
        assert np.array_equal(r["status"], orc["status"])
        assert np.array_equal(r["n_px"], orc["n_px"])
        both = (r["converged"] == 1) & (orc["converged"] == 1)
        agree = np.mean(r["converged"] == orc["converged"])
        du = np.abs(r["p"][both, :2] - orc["p"][both, :2]).max()
        ds = np.abs(r["p"][both, 2:6] - orc["p"][both, 2:6]).max()
        print(f"job {j.job_id}: {len(coords)} subsets order {order} {ref.shape}: max |du| {du:.2e} px, "
              f"|dstrain| {ds:.2e}, convergence agreement {agree:.4f}, converged {orc['converged'].mean():.4f}",
              flush=True)
        assert agree >= 0.99 and both.mean() > 0.9
        assert du < 1e-4 and ds < 1e-5
