"""Benchmark: ICGN subsets/s of the FP32-interpolation batched engine against
the plain FP64 drop-in, on one MI355X.

Workload (one step = one frame): a 2048x2048 8-bit speckle pair (50 000
Gaussian speckles, sigma 2.5 px, noise sigma 2 grey levels) under an affine
field (u_x = 1 %, v_y = -0.3 %, shear u_y = 0.5 %); the reference image's
B-spline coefficients stay resident, and each frame prefilters the target
(GPU spectral deconvolution, circular-FIR form) and solves 40 000 circle
subsets of radius 20 on a
10 px grid (first-order warp, ||dp|| < 1e-5, at most 50 iterations) from
p0 = the true displacement rounded to whole pixels.  Pixels are resident in
HBM before the timed region.

Prints one JSON line (rank 0).  The CPU baseline is the FP64 oracle
(a C++ restatement of the reference's CPU path) timed on a bounded sample of
the same subsets with all the host's cores and with one thread, and on the
target's prefilter; the same sample checks the GPU results against it.
"""

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "geopyv-dev_amd"))

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu():
    """Cores this process may use, and the CPU model (lscpu's model name)."""
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return cores, model


def measured_counters():
    """Counters of the batched kernel on this workload as the summary under
    profiles/ records them (VALU issue fraction, LDS bank-conflict cycles
    per LDS instruction, VALU instructions per subset), or {}."""
    path = os.path.join(ROOT, "profiles", "icgn_batch_kernel_uniform_mesh_pmc.txt")
    keys = {"valu_issue": "valu_issue", "lds_conflict_cycles_per_instr": "lds_conflict_cycles_per_instr",
            "valu_per_subset": "valu_per_subset", "commit": "commit"}
    out = {}
    try:
        with open(path) as fh:
            for line in fh:
                k = line.split("=")[0].strip()
                if k in keys and "=" in line:
                    v = line.split("=", 1)[1].strip()
                    try:
                        out[keys[k]] = float(v)
                    except ValueError:
                        out[keys[k]] = v
    except OSError:
        pass
    if out:
        out["source"] = os.path.relpath(path, ROOT)
    return out


def measured_traffic():
    """HBM bytes per launch of the batched kernel on this workload, as the
    FETCH_SIZE / WRITE_SIZE counter summary under profiles/ records them
    (scaled by a copy of known size); None when no summary is present."""
    path = os.path.join(ROOT, "profiles", "icgn_batch_kernel_uniform_mesh_pmc.txt")
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("hbm_bytes_per_launch"):
                    return float(line.split("=")[1])
    except OSError:
        pass
    return None


def truth_errors(res32, res64, truth, coords, radius, h, w):
    """|p - truth| of the displacements over the converged subsets (p50, p99,
    max) and the number of subsets whose warped template reaches past the
    image edge, where positions are clamped (image.rs:130-132) and the truth
    is not recoverable."""
    def stats(res):
        ok = res["converged"] == 1
        e = np.abs(res["p"][:, :2] - truth[:, :2]).max(axis=1)[ok]
        return {"p50": float(np.percentile(e, 50)), "p99": float(np.percentile(e, 99)), "max": float(e.max())}
    reach = radius + np.abs(truth[:, :2]).max(axis=1) + radius * np.abs(truth[:, 2:6]).sum(axis=1)
    clamped = ((coords[:, 0] - reach < 0) | (coords[:, 0] + reach > w - 1) | (coords[:, 1] - reach < 0) |
               (coords[:, 1] + reach > h - 1))
    inside = (res32["converged"] == 1) & ~clamped
    e_in = np.abs(res32["p"][:, :2] - truth[:, :2]).max(axis=1)[inside]
    return {"f32_converged": stats(res32), "fp64_plain_converged": stats(res64),
            "f32_converged_unclamped_max": float(e_in.max()), "clamped_subsets": int(clamped.sum()),
            "not_converged_f32": int(np.sum(res32["converged"] == 0)),
            "not_converged_fp64": int(np.sum(res64["converged"] == 0))}


def run_shear_band(args, torch, gp, _native, workloads):
    """Batched uneven subsets, second-order warp: 4096 x 3000 shear-band
    pair, ~150 000 subsets with radii 15..40 on adaptive nodes, all in one
    frame of concurrent size-class launches under the autotuned schedule."""
    t0 = time.time()
    w = workloads.shear_band_pair(seed=2)
    log(f"workload generated in {time.time() - t0:.1f}s")
    ref, tar, coords, radii, p0, truth = w["ref"], w["tar"], w["coords"], w["radii"], w["p0"], w["truth"]
    n = len(coords)
    border = 20
    eng = gp._engine()
    f_img = _native.DeviceImage(eng, pixels=ref, border=border)
    dev_tar = torch.from_numpy(tar).cuda()
    torch.cuda.synchronize()
    g_img = _native.DeviceImage(eng, device_ptr=dev_tar.data_ptr(), shape=tar.shape, pix_type=_native.PIX_U8,
                                border=border)
    masks = {int(r): gp.Mask(mask_type="local", shape="circle", size=int(r)) for r in np.unique(radii)}
    index, start, xy = gp.templates_from_masks([masks[int(r)].coords for r in radii])
    batch = _native.Batch(eng, coords, 2, index, start, np.rint(xy).astype(np.int16))
    batch.set_p0(p0)
    if not args.no_autotune:
        t = time.time()
        batch.autotune(f_img, g_img, max_iterations=args.max_iterations, max_norm=1e-5)
        log(f"autotuned in {time.time() - t:.1f}s")

    def frame():
        g_img.update(device_ptr=dev_tar.data_ptr(), pix_type=_native.PIX_U8)
        batch.solve(f_img, g_img, method="icgn", precision=_native.PREC_COMP_F32,
                    max_iterations=args.max_iterations, tolerance=0.75, max_norm=1e-5)
        eng.sync()

    for _ in range(args.warmup):
        frame()
    lat, sol = [], []
    torch.cuda.synchronize()
    tb = time.perf_counter()
    for _ in range(args.steps):
        t = time.perf_counter()
        frame()
        lat.append((time.perf_counter() - t) * 1e3)
        sol.append(eng.timing()[1])
    torch.cuda.synchronize()
    total_s = time.perf_counter() - tb
    res = batch.results()
    # one frame of the plain FP64 drop-in on the same batch, for the ratio
    batch.set_p0(p0)
    t = time.perf_counter()
    batch.solve(f_img, g_img, method="icgn", precision=_native.PREC_PLAIN_F64, max_iterations=args.max_iterations,
                tolerance=0.75, max_norm=1e-5)
    eng.sync()
    plain_ms = eng.timing()[1]
    log(f"plain FP64 frame: kernel {plain_ms:.1f} ms, wall {(time.perf_counter() - t) * 1e3:.1f} ms")
    res64 = batch.results()
    out = {
        "metric": "ICGN subsets/sec, uneven radii 15-40 px, second-order warp (FP32 engine)",
        "value": round(n / (total_s / args.steps), 1), "unit": "subsets/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(total_s * 1e3 / args.steps, 4), "higher_is_better": True,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "shear band: 4096x3000 pair, 150000 speckles sigma U[1.5,3], band width 60 px at "
                               "30 deg, slip 4 px, v += 0.8 sin(2 pi x/512); adaptive nodes, radii U{15..40}, "
                               "order 2", "subsets": n, "image": list(ref.shape), "order": 2},
        "frame_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4),
                             "p99": round(float(np.percentile(lat, 99)), 4)},
        "kernel_ms": round(float(np.mean(sol)), 4),
        "fp64_plain_kernel_ms": round(plain_ms, 4),
        "speedup_vs_fp64_plain_kernel": round(plain_ms / float(np.mean(sol)), 3),
        "iterations_mean": round(float(np.mean(res["iterations"])), 3),
        "converged_fraction": round(float(np.mean(res["converged"])), 5),
        "max_abs_du_f32_vs_f64_px": float(np.abs(res["p"][:, :2] - res64["p"][:, :2])[
            (res["converged"] == 1) & (res64["converged"] == 1)].max(initial=0.0)),
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        threads, model = host_cpu()
        Cf = O.get_c(ref.astype(np.float64), border, threads)
        Cg = O.get_c(tar.astype(np.float64), border, threads)

        def cpu_sample(k):
            idx = np.linspace(0, n - 1, min(k, n)).astype(np.int64)
            t = O.wall()
            orc = O.batch_solve(Cf, Cg, ref.shape, border, coords[idx], index[idx], xy.astype(np.float64), start,
                                np.full(len(idx), 2, np.int32), p0[idx], max_norm=1e-5, max_iterations=50,
                                threads=threads)
            return idx, orc, O.wall() - t

        # a probe of 1000 subsets sizes the sample to about --cpu-seconds
        idx, orc, cpu_s = cpu_sample(1000)
        k = int(len(idx) / max(cpu_s, 1e-6) * args.cpu_seconds)
        if k > len(idx):
            idx, orc, cpu_s = cpu_sample(k)
        d = res[idx]
        both = (d["converged"] == 1) & (orc["converged"] == 1)
        out["cpu_baseline"] = {"value": round(len(idx) / cpu_s, 2), "unit": "subsets/s", "cores": threads,
                               "cpu_model": model, "kind": "port",
                               "sample": f"{len(idx)} evenly strided subsets, {cpu_s:.1f} s"}
        out["error_vs_oracle"] = {"max_du_px": float(np.abs(d["p"][both, :2] - orc["p"][both, :2]).max()),
                                  "max_dstrain": float(np.abs(d["p"][both, 2:6] - orc["p"][both, 2:6]).max()),
                                  "converged_match": float(np.mean(d["converged"] == orc["converged"])),
                                  "sample": len(idx)}
    print(json.dumps(out))


def run_sequence(args, torch, gp, _native, workloads):
    """Incremental sequence: 2048² frames (80 000 speckles, σ = 2 px) under a
    shear band whose slip grows linearly to 8 px over 200 frames, 30 000
    subsets r = 25, reference updated every 20 frames (device prefilter),
    warps warm-started from the previous frame, each frame one replay of a
    captured HIP graph (prefilter + warm start + solve).  Frames are staged
    device-to-device from HBM before each replay; the timed region is
    staging + replay + synchronisation, per frame."""
    n_total = 200
    nf = max(2, min(args.frames, n_total))
    t0 = time.time()
    # the first nf frames of the 200-frame sequence (same slip per frame)
    frames, _ = workloads.sequence_frames(seed=3, shape=(2048, 2048), n_frames=nf, n_speckles=80000,
                                          slip_max=8.0 * (nf - 1) / (n_total - 1),
                                          progress=lambda k: k % 10 == 9 and log(f"frame {k + 1} generated"),
                                          workers=min(16, host_cpu()[0]) if nf > 8 else 1)
    log(f"{nf} frames generated in {time.time() - t0:.1f}s")
    coords = workloads.grid_coords((2048, 2048), 11, 66, n_side=None)
    coords = coords[:30000]
    n = len(coords)
    mask = gp.Mask(mask_type="local", shape="circle", size=25)
    xy = np.rint(mask.coords).astype(np.int16)
    eng = gp._engine()
    dev = [torch.from_numpy(fr).cuda() for fr in frames]  # frames resident in HBM
    torch.cuda.synchronize()
    f_img = _native.DeviceImage(eng, device_ptr=dev[0].data_ptr(), shape=frames[0].shape, pix_type=_native.PIX_U8,
                                border=20)
    g_img = _native.DeviceImage(eng, device_ptr=dev[1].data_ptr(), shape=frames[0].shape, pix_type=_native.PIX_U8,
                                border=20)
    batch = _native.Batch(eng, coords, 1, np.zeros(n, np.int32), np.array([0, len(xy)], np.int64), xy)
    batch.reset_warps(np.zeros((n, 12)))
    g_img.stage(device_ptr=dev[1].data_ptr(), pix_type=_native.PIX_U8)
    graph = _native.FrameGraph(batch, f_img, g_img, warm_start=True, max_iterations=50, max_norm=1e-5)
    eng.sync()
    lat, ref_ms, its = [], [], []
    for k in range(2, nf):
        t = time.perf_counter()
        g_img.stage(device_ptr=dev[k].data_ptr(), pix_type=_native.PIX_U8)
        graph.launch()
        eng.sync()
        lat.append((time.perf_counter() - t) * 1e3)
        if k % 20 == 0 and k + 1 < nf:
            t = time.perf_counter()
            f_img.update(device_ptr=dev[k].data_ptr(), pix_type=_native.PIX_U8)
            batch.reset_warps(np.zeros((n, 12)))
            eng.sync()
            ref_ms.append((time.perf_counter() - t) * 1e3)
        if k % 10 == 0:
            its.append(float(np.mean(batch.results()["iterations"])))
    lat = np.array(lat)
    out = {
        "metric": "incremental sequence: frames/s and subsets/s through a captured frame graph (FP32 engine)",
        "value": round(n / (lat.mean() * 1e-3), 1), "unit": "subsets/s", "n_gpus": 1, "steps": len(lat),
        "frames_per_s": round(1e3 / lat.mean(), 2), "higher_is_better": True, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "sequence: 2048x2048 frames, 80000 speckles sigma 2, shear band slip 0->8 px over "
                               "200 frames, 30000 subsets r=25, reference update every 20 frames, warm start",
                   "frames_timed": len(lat), "subsets": n},
        "frame_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4),
                             "p99": round(float(np.percentile(lat, 99)), 4)},
        "reference_update_ms": round(float(np.mean(ref_ms)), 4) if ref_ms else None,
        "iterations_mean_sampled": round(float(np.mean(its)), 3) if its else None,
    }
    print(json.dumps(out))


def run_mesh(args, torch, gp, _native, workloads):
    """Mesh mode: the 2048² pair of the headline workload, its 200 x 200
    grid of nodes (r = 20, 10 px pitch) as a triangular mesh, solved by
    layer-RG from the seed node nearest the image centre with a zero seed
    warp: no initial warps are given, the traversal preconditions every node
    from its solved neighbours (mesh.rs:628-857).  One step = one whole mesh
    solve (seed, rounds, corrections, strains) with the images resident."""
    t0 = time.time()
    w = workloads.uniform_mesh_pair(seed=1)
    log(f"workload generated in {time.time() - t0:.1f}s")
    ref, tar, coords, truth = w["ref"], w["tar"], w["coords"], w["truth"]
    nodes, el = gp.structured_mesh(29.0, 29.0, 2019.0, 2019.0, 10.0)
    assert np.array_equal(nodes, coords)
    n = len(nodes)
    mask = gp.Mask(mask_type="local", shape="circle", size=w["radius"])
    xy = np.rint(mask.coords).astype(np.int16)
    tpl_index, tpl_start = np.zeros(n, np.int32), np.array([0, len(xy)], np.int64)
    eng = gp._engine()
    dref = torch.from_numpy(ref).cuda()
    dtar = torch.from_numpy(tar).cuda()
    torch.cuda.synchronize()
    f_img = _native.DeviceImage(eng, device_ptr=dref.data_ptr(), shape=ref.shape, pix_type=_native.PIX_U8, border=20)
    g_img = _native.DeviceImage(eng, device_ptr=dtar.data_ptr(), shape=ref.shape, pix_type=_native.PIX_U8, border=20)
    eng.sync()
    seed = [1024.0, 1024.0]
    # override: corner nodes whose target window leaves the image (up to
    # 15 px of displacement at the corners) stay short of quality_ok, as
    # they do in the batch mode; they are counted, not fatal
    common = dict(subset_order=1, max_iterations=args.max_iterations, tolerance=0.75, max_norm=1e-5,
                  seed_tolerance=0.9, preconditioning="layer-RG", override=True)

    def solve(precision):
        t = time.perf_counter()
        st, msg, r = _native.mesh_solve(eng, f_img, g_img, nodes, el, 1, tpl_index, tpl_start, xy, seed,
                                        np.zeros(6), precision=precision, **common)
        dt = time.perf_counter() - t
        if st:
            raise RuntimeError(f"mesh solve failed ({st}): {msg}")
        return r, dt

    for _ in range(max(args.warmup, 1)):
        solve(_native.PREC_COMP_F32)
    steps = max(1, min(args.steps, 10))
    times = []
    for k in range(steps):
        r32, dt = solve(_native.PREC_COMP_F32)
        times.append(dt)
        log(f"mesh step {k}: {dt * 1e3:.1f} ms, {r32['n_lists']} lists ({r32['list_seconds'] * 1e3:.1f} ms in list "
            f"solves), {r32['n_node_solves']} node solves")
    r64, dt64 = solve(_native.PREC_PLAIN_F64)
    log(f"plain FP64 mesh solve: {dt64 * 1e3:.1f} ms")
    ms = float(np.mean(times)) * 1e3
    ok = r32["quality_ok"]
    e32 = np.abs(r32["p"][ok, :2] - truth[ok, :2]).max(axis=1)
    d32_64 = float(np.abs(r32["p"][:, :2] - r64["p"][:, :2]).max())
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        threads, model = host_cpu()
        Cf, Cg = O.get_c(ref.astype(np.float64), 20, threads=threads), O.get_c(tar.astype(np.float64), 20,
                                                                                threads=threads)
        t = time.perf_counter()
        st, ro = O.mesh_solve(Cf, Cg, ref.shape, 20, nodes, el, tpl_index, mask.coords, tpl_start, seed,
                              np.zeros(6), preconditioning="layer-RG", threads=threads, workers=threads,
                              **{k: v for k, v in common.items() if k != "preconditioning"})
        cpu_s = time.perf_counter() - t
        log(f"oracle layer-RG mesh solve: {cpu_s:.1f}s on {threads} threads, status {st}")
        cpu = {"value": round(n / cpu_s, 1), "unit": "nodes/s", "cores": threads, "kind": "port",
               "sample": f"the whole {n}-node mesh, layer-RG with {threads} workers (the reference's rayon "
                         "decomposition: one node per task within a colour)",
               "max_abs_du_gpu_f64_vs_oracle_px": float(np.abs(r64["p"][:, :2] - ro["p"][:, :2]).max()),
               "max_abs_du_gpu_f32_vs_oracle_px": float(np.abs(r32["p"][:, :2] - ro["p"][:, :2]).max())}
    out = {
        "metric": "mesh solve: nodes/s of a layer-RG mesh solve from a seed (FP32 engine)",
        "value": round(n / (ms * 1e-3), 1), "unit": "nodes/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "mesh: 2048x2048 pair, 50000 speckles sigma 2.5, affine u_x=1% v_y=-0.3% u_y=0.5%, "
                               "200x200 nodes 10 px pitch as 79202 triangles, circle r=20, order 1, layer-RG "
                               "from the centre node, zero seed warp",
                   "nodes": n, "elements": int(len(el)),
                   "workers": int(torch.cuda.get_device_properties(0).multi_processor_count)},
        "lists_per_solve": r32["n_lists"], "node_solves_per_solve": r32["n_node_solves"],
        "list_solve_ms": round(r32["list_seconds"] * 1e3, 3),
        "nodes_without_quality_ok": r32["n_unsolvable"],
        "plain_f64_ms": round(dt64 * 1e3, 3), "f32_vs_f64_speedup": round(dt64 * 1e3 / ms, 3),
        # the largest errors sit at the image edge, where 13-15 px of
        # displacement carries part of the subset out of the image
        "abs_du_vs_truth_px_quality_ok_nodes": {q: round(float(np.percentile(e32, v)), 5) for q, v in
                                                (("p50", 50), ("p99", 99), ("p99.9", 99.9), ("max", 100))}, "max_abs_du_f32_vs_f64_px": d32_64,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def run_lists(args, torch, gp, _native, workloads):
    """List mode: short list solves (dic_batch_solve_list) on the headline
    pair, as a mesh traversal issues them (mesh.rs:1224-1233 solves one
    colour's nodes at a time): 1, 16, 100 and 400 subsets drawn at random
    from the 40 000, each from its rounded-truth warp, timed one list at a
    time with the images resident; a list's wall time includes its launch,
    the device work and the wait for it."""
    w = workloads.uniform_mesh_pair(seed=1)
    ref, tar, coords = w["ref"], w["tar"], w["coords"]
    n = len(coords)
    mask = gp.Mask(mask_type="local", shape="circle", size=w["radius"])
    xy = np.rint(mask.coords).astype(np.int16)
    eng = gp._engine()
    dref = torch.from_numpy(ref).cuda()
    dtar = torch.from_numpy(tar).cuda()
    torch.cuda.synchronize()
    f_img = _native.DeviceImage(eng, device_ptr=dref.data_ptr(), shape=ref.shape, pix_type=_native.PIX_U8, border=20)
    g_img = _native.DeviceImage(eng, device_ptr=dtar.data_ptr(), shape=ref.shape, pix_type=_native.PIX_U8, border=20)
    b = _native.Batch(eng, coords, 1, np.zeros(n, np.int32), np.array([0, len(xy)], np.int64), xy)
    rng = np.random.default_rng(0)
    reps = max(args.steps, 5)
    sizes = {}
    for size in (1, 16, 100, 400):
        idx = np.sort(rng.choice(n, size, replace=False)).astype(np.int32)
        p0 = w["p0"][idx]
        kw = dict(precision=_native.PREC_COMP_F32, max_iterations=args.max_iterations, tolerance=0.75,
                  max_norm=1e-5)
        for _ in range(max(args.warmup, 1)):
            b.solve_list(f_img, g_img, idx, p0, **kw)
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            r = b.solve_list(f_img, g_img, idx, p0, **kw)
            ts.append(time.perf_counter() - t)
        us = np.array(ts) * 1e6
        sizes[str(size)] = {"p50_us": round(float(np.median(us)), 1), "p99_us": round(float(np.percentile(us, 99)), 1),
                            "subsets_per_s": round(size / (float(np.median(us)) * 1e-6), 1),
                            "iterations_mean": round(float(r["iterations"].mean()), 2),
                            "iterations_max": int(r["iterations"].max())}
        log(f"list of {size}: p50 {sizes[str(size)]['p50_us']} us, iterations mean "
            f"{sizes[str(size)]['iterations_mean']} max {sizes[str(size)]['iterations_max']}")
    out = {
        "metric": "list solve latency: wall time of one short list solve (FP32 engine), 100-subset list",
        "value": sizes["100"]["p50_us"], "unit": "us", "n_gpus": 1, "steps": reps, "warmup": args.warmup,
        "higher_is_better": False, "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "lists: subsets of the 2048x2048 uniform-mesh pair (40000 circle subsets r=20), "
                               "lists of 1/16/100/400 drawn at random, rounded-truth initial warps, order 1",
                   "max_norm": 1e-5},
        "lists": sizes,
    }
    print(json.dumps(out), flush=True)


def run_serving(args, gp, workloads):
    """Serving loop: Poisson arrivals of correlation jobs, image sizes drawn
    from {1024², 2048², 4096x3000} (speckle density 0.012/px², σ = 2 px,
    affine fields with strains U[-2 %, 2 %]), subset counts log-uniform in
    1k-150k, radii U{15..40}, warp order 1 or 2; references cached on the
    device, admission by HBM budget, 4 slots.  Image pairs come from a pool
    (one reference and two targets per size) so generation stays short."""
    from geopyv_dev import serving

    rng = np.random.Generator(np.random.PCG64(44))
    sizes = [(1024, 1024), (2048, 2048), (3000, 4096)]
    pool = {}
    t0 = time.time()
    for si, shape in enumerate(sizes):
        h, w = shape
        sp = workloads.Speckle(shape, int(round(0.012 * h * w)), 2.0, 50 + si)
        ref = sp.reference(2.0)
        tars = []
        for k in range(2):
            g = rng.uniform(-0.02, 0.02, 4)
            field = workloads.Affine(((w - 1) / 2.0, (h - 1) / 2.0), g[0], g[1], g[2], g[3])
            tars.append((field, sp.target(field, 2.0, rng=np.random.Generator(np.random.PCG64(200 + 10 * si + k)))))
        pool[shape] = (ref, tars)
        log(f"pool {shape} generated ({time.time() - t0:.1f}s)")
    rates = [float(r) for r in args.rates.split(",")]
    out = {"metric": "serving loop: ICGN subsets/s and job latency under Poisson arrivals (FP32 engine)",
           "unit": "subsets/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "serving: jobs of 1024^2 / 2048^2 / 4096x3000 pairs, density 0.012/px^2, affine "
                                  "strains U[-2%,2%], subsets log-uniform 1k-150k, radii U{15..40}, order 1 or 2; "
                                  "4 slots, reference cache, HBM budget 64 GiB",
                      "jobs_per_rate": args.jobs},
           "rates": {}}
    def job_stream(lam, n_jobs):
        # jobs are made as they arrive, so a long stream never holds more
        # than the queue
        t = 0.0
        for j in range(n_jobs):
            t += rng.exponential(1.0 / lam)
            shape = sizes[int(rng.integers(0, 3))]
            h, w = shape
            ref, tars = pool[shape]
            field, tar = tars[int(rng.integers(0, 2))]
            nsub = int(round(math.exp(rng.uniform(math.log(1000), math.log(150000)))))
            coords = np.column_stack([rng.uniform(46, w - 47, nsub), rng.uniform(46, h - 47, nsub)])
            coords = np.rint(coords * 4.0) / 4.0
            truth = field.warp(coords[:, 0], coords[:, 1])
            p0 = np.zeros((nsub, 12))
            p0[:, :2] = np.rint(truth[:, :2])
            yield t, serving.Job(j, shape, ref, tar, coords, rng.integers(15, 41, nsub), order=int(rng.integers(1, 3)),
                                 p0=p0)

    def summary(job, res, lat_ms):
        return len(job.coords), int(np.sum(res["converged"])), lat_ms

    for lam in rates:
        srv = serving.Server(n_slots=4, hbm_budget=64 << 30)
        tb = time.perf_counter()
        last = [0.0]

        def progress(done, submitted):
            now = time.perf_counter()
            if now - last[0] > 20.0:
                last[0] = now
                log(f"rate {lam}: {done} of {submitted} submitted jobs done, {now - tb:.0f}s")

        res = srv.stream(job_stream(lam, args.jobs), summary, progress)
        wall = time.perf_counter() - tb
        srv.close()
        lat = np.array([v[2] for v in res.values()])
        nsub_total = sum(v[0] for v in res.values())
        conv = sum(v[1] for v in res.values()) / max(nsub_total, 1)
        out["rates"][str(lam)] = {"subsets_per_s": round(nsub_total / wall, 1), "jobs": len(res),
                                  "subsets": int(nsub_total), "wall_s": round(wall, 3), "latency_ms": {
                                      "p50": round(float(np.percentile(lat, 50)), 3),
                                      "p99": round(float(np.percentile(lat, 99)), 3),
                                      "max": round(float(lat.max()), 3)},
                                  "converged_fraction": round(float(conv), 5)}
        log(f"rate {lam}: {out['rates'][str(lam)]}")
    out["value"] = out["rates"][str(rates[0])]["subsets_per_s"]
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--cpu-sample", type=int, default=40000)
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="repeat the CPU baseline's frame until this much wall time has passed")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--waves-per-subset", type=int, default=0, help="batched-engine schedule (0 = automatic)")
    ap.add_argument("--win-slack", type=int, default=3, help="batched-engine LDS window slack in pixels")
    ap.add_argument("--max-iterations", type=int, default=50,
                    help="iteration cap (the workload's is 50; smaller values split the frame time by phase)")
    ap.add_argument("--max-norm", type=float, default=1e-5,
                    help="convergence limit (the workload's is 1e-5; 0 runs every subset to the iteration cap, "
                         "which splits the kernel's work into a fixed part and a per-iteration part)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 512 MiB device-to-device copy (known HBM bytes) for counter scaling")
    ap.add_argument("--engine", choices=["both", "f32"], default="both",
                    help="f32: time only the batched engine (for profiler runs)")
    ap.add_argument("--workload", choices=["uniform_mesh", "shear_band", "sequence", "serving", "mesh", "lists"],
                    default="uniform_mesh",
                    help="shear_band: uneven radii, second-order warp; sequence: incremental frames through a "
                         "captured frame graph (each its own line, not the headline)")
    ap.add_argument("--frames", type=int, default=24, help="sequence: frames (the full sequence has 200)")
    ap.add_argument("--jobs", type=int, default=120, help="serving: jobs per arrival rate (the full stream has 5000)")
    ap.add_argument("--rates", default="20,50,100", help="serving: Poisson arrival rates, jobs/s")
    ap.add_argument("--no-autotune", action="store_true", help="shear_band: keep the default schedule")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the engine has no multi-GPU "
              "path in this build", file=sys.stderr)
        sys.exit(2)

    import torch  # device buffers for the resident input pixels

    if args.copy_calibration:
        a = torch.ones(512 * 2 ** 20 // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log("copy calibration: 3 copies of 512 MiB")
        return
    import geopyv_dev as gp
    from geopyv_dev import _native, workloads

    if args.workload == "shear_band":
        run_shear_band(args, torch, gp, _native, workloads)
        return
    if args.workload == "sequence":
        run_sequence(args, torch, gp, _native, workloads)
        return
    if args.workload == "serving":
        run_serving(args, gp, workloads)
        return
    if args.workload == "mesh":
        run_mesh(args, torch, gp, _native, workloads)
        return
    if args.workload == "lists":
        run_lists(args, torch, gp, _native, workloads)
        return

    t0 = time.time()
    w = workloads.uniform_mesh_pair(seed=1)
    log(f"workload generated in {time.time() - t0:.1f}s")
    ref, tar = w["ref"], w["tar"]
    coords, p0, truth = w["coords"], w["p0"], w["truth"]
    n = len(coords)
    radius = w["radius"]
    h, wd = ref.shape
    border = 20

    eng = gp._engine()
    f_img = _native.DeviceImage(eng, pixels=ref, border=border)
    dev_tar = torch.from_numpy(tar).cuda()
    torch.cuda.synchronize()
    g_img = _native.DeviceImage(eng, device_ptr=dev_tar.data_ptr(), shape=tar.shape, pix_type=_native.PIX_U8,
                                border=border)
    mask = gp.Mask(mask_type="local", shape="circle", size=radius)
    xy = np.rint(mask.coords).astype(np.int16)
    batch = _native.Batch(eng, coords, 1, np.zeros(n, np.int32), np.array([0, len(xy)], np.int64), xy)
    batch.set_p0(p0)
    batch.set_schedule(args.win_slack, args.waves_per_subset)

    def frame(precision):
        g_img.update(device_ptr=dev_tar.data_ptr(), pix_type=_native.PIX_U8)
        batch.solve(f_img, g_img, method="icgn", precision=precision, max_iterations=args.max_iterations,
                    tolerance=0.75, max_norm=args.max_norm)
        eng.sync()

    # ---- FP32-interpolation batched engine (the measured value) ----
    for _ in range(args.warmup):
        frame(_native.PREC_COMP_F32)
    lat, pre_ms, sol_ms = [], [], []
    torch.cuda.synchronize()
    tb = time.perf_counter()
    for _ in range(args.steps):
        t = time.perf_counter()
        frame(_native.PREC_COMP_F32)
        lat.append((time.perf_counter() - t) * 1e3)
        a, b = eng.timing()
        pre_ms.append(a)
        sol_ms.append(b)
    torch.cuda.synchronize()
    total_s = time.perf_counter() - tb
    res32 = batch.results()
    log(f"f32 engine: {np.mean(lat):.3f} ms/frame, solve {np.mean(sol_ms):.3f} ms, prefilter {np.mean(pre_ms):.3f} ms")

    # ---- plain FP64 drop-in on the same frames ----
    plain_lat, plain_sol = [], []
    if args.engine == "f32" or args.max_iterations != 50 or args.max_norm != 1e-5:
        # profiler and phase-timing runs: no headline line
        args.plain_steps = 0
        args.no_cpu = True
    else:
        frame(_native.PREC_PLAIN_F64)
    for _ in range(args.plain_steps):
        t = time.perf_counter()
        frame(_native.PREC_PLAIN_F64)
        plain_lat.append((time.perf_counter() - t) * 1e3)
        plain_sol.append(eng.timing()[1])
    if not plain_lat:
        log(json.dumps({"f32_ms_per_frame": float(np.mean(lat)), "f32_solve_ms": float(np.mean(sol_ms)),
                        "prefilter_ms": float(np.mean(pre_ms))}))
        return
    res64 = batch.results()
    log(f"fp64 plain: {np.mean(plain_lat):.3f} ms/frame, solve {np.mean(plain_sol):.3f} ms")

    ms_per_step = total_s * 1e3 / args.steps
    value = n / (total_s / args.steps)
    it_mean = float(np.mean(res32["iterations"]))
    # algorithmic bytes per subset: the coefficient window once for setup and
    # once per iteration (FP32), p0 in and the result out
    m = 6
    wx = 2 * int(np.ceil(radius * (1 + 0.010 + 0.005))) + 6
    wy = 2 * int(np.ceil(radius * (1 + 0.0 + 0.003))) + 6
    bytes_per_subset = 4 * wx * wy * (1 + it_mean) + 24 * m + 64
    sol_avg = float(np.mean(sol_ms))
    achieved = n * bytes_per_subset / (sol_avg * 1e-3) / 1e9
    traffic = measured_traffic()

    out = {
        "metric": "ICGN subsets/sec (FP32-comp vs FP64 plain, <=1e-4 px); p50/p99 frame latency",
        "value": round(value, 1),
        "unit": "subsets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "uniform mesh: 2048x2048 speckle pair, 40000 circle subsets r=20 on a 10 px grid, "
                               "order-1 warp, affine field, target prefiltered per frame",
                   "subsets": n, "radius": radius, "image": [h, wd], "order": 1, "max_norm": 1e-5},
        "frame_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4),
                             "p99": round(float(np.percentile(lat, 99)), 4)},
        "frames_per_s": round(1e3 / ms_per_step, 3),
        "kernel_ms": {"prefilter": round(float(np.mean(pre_ms)), 4), "icgn_batch": round(sol_avg, 4),
                      "icgn_plain_f64": round(float(np.mean(plain_sol)), 4)},
        "fp64_plain": {"subsets_per_s": round(n / (np.mean(plain_lat) * 1e-3), 1),
                       "ms_per_frame": round(float(np.mean(plain_lat)), 4)},
        "speedup_vs_fp64_plain": round(float(np.mean(plain_lat) / np.mean(lat)), 3),
        "iterations_mean": round(it_mean, 3),
        "converged_fraction": round(float(np.mean(res32["converged"])), 5),
        "roofline": dict({"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                          "bytes_per_subset": round(bytes_per_subset, 1), "kernel": "icgn_batch_kernel (order 1)",
                          "limiter": "VALU issue (the window bytes are served from L2: traffic << achieved x time)"},
                         **{k: v for k, v in measured_counters().items()}),
        "error_vs_truth_px": truth_errors(res32, res64, truth, coords, radius, h, wd),
    }

    # ---- CPU baseline + parity on a bounded sample ----
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        threads, model = host_cpu()
        log(f"CPU baseline on {threads} cores ({model})")
        # the target's prefilter, as every frame runs it (image.rs:252-327)
        t = O.wall()
        Cf = O.get_c(ref.astype(np.float64), border, threads)
        Cg = O.get_c(tar.astype(np.float64), border, threads)
        pre_s = (O.wall() - t) / 2.0
        t = O.wall()
        O.get_c(tar.astype(np.float64), border, 1)
        pre_s1 = O.wall() - t

        def solve_sample(idx, nthreads):
            return O.batch_solve(Cf, Cg, ref.shape, border, coords[idx], np.zeros(len(idx), np.int32), mask.coords,
                                 np.array([0, len(xy)]), np.full(len(idx), 1, np.int32), p0[idx], max_norm=1e-5,
                                 max_iterations=50, threads=nthreads)

        idx = np.linspace(0, n - 1, min(args.cpu_sample, n)).astype(np.int64)
        t = O.wall()
        reps = 0
        while True:
            orc = solve_sample(idx, threads)
            reps += 1
            cpu_s = O.wall() - t
            if cpu_s >= args.cpu_seconds:
                break
        rate = reps * len(idx) / cpu_s
        # one thread: a strided sample sized to about a quarter of the budget
        k1 = max(64, int(rate / threads * args.cpu_seconds / 4))
        idx1 = np.linspace(0, n - 1, min(k1, n)).astype(np.int64)
        t = O.wall()
        solve_sample(idx1, 1)
        cpu_s1 = O.wall() - t
        rate1 = len(idx1) / cpu_s1
        frame_s = pre_s + n / rate
        out["cpu_baseline"] = {
            "value": round(n / frame_s, 2), "unit": "subsets/s", "cores": threads, "kind": "port",
            "cpu_model": model,
            "solve_only_subsets_per_s": round(rate, 2),
            "prefilter_ms": round(pre_s * 1e3, 2),
            "single_thread": {"subsets_per_s": round(rate1, 2), "prefilter_ms": round(pre_s1 * 1e3, 2),
                              "sample": len(idx1)},
            "sample": f"{reps} x {len(idx)} of the {n} subsets of one frame (evenly strided), FP64 oracle with "
                      f"OpenMP over subsets on {threads} cores, {cpu_s:.1f} s, plus the target's prefilter "
                      f"({pre_s * 1e3:.1f} ms); value = subsets per second of whole frames "
                      f"(prefilter + {n} subsets at the sampled rate)"}
        d32 = res32[idx]
        d64 = res64[idx]
        c32 = (d32["converged"] == 1) & (orc["converged"] == 1)
        c64 = (d64["converged"] == 1) & (orc["converged"] == 1)
        du32 = np.abs(d32["p"][:, :2] - orc["p"][:, :2]).max(axis=1)[c32]
        ds32 = np.abs(d32["p"][:, 2:6] - orc["p"][:, 2:6]).max(axis=1)[c32]
        out["error_vs_oracle"] = {
            "f32_max_du_px": float(du32.max()),
            "f32_p99_du_px": float(np.percentile(du32, 99)),
            "f32_max_dstrain": float(ds32.max()),
            "f32_p99_dstrain": float(np.percentile(ds32, 99)),
            "f32_iterations_match": float(np.mean(d32["iterations"] == orc["iterations"])),
            "fp64_max_du_px": float(np.abs(d64["p"][c64, :2] - orc["p"][c64, :2]).max()),
            "fp64_max_dstrain": float(np.abs(d64["p"][c64, 2:6] - orc["p"][c64, 2:6]).max()),
            "fp64_iterations_match": float(np.mean(d64["iterations"] == orc["iterations"])),
            "converged_match": float(np.mean(d64["converged"] == orc["converged"])),
            "f32_converged_match": float(np.mean(d32["converged"] == orc["converged"])),
            "scope": "subsets that converged in both (the displacement bar applies to converged solutions)",
            "sample": len(idx),
        }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
