"""Serving loop: continuous batching of correlation jobs on one GPU.

Jobs (a reference/target image pair, subset centres with per-subset radii,
warp order, initial warps) arrive over time.  `Server` hands each to the
library's serving loop (dic_submit / dic_poll, csrc/serving.cpp): worker
threads there, one engine and stream per slot, admit the queued jobs first
come first served as soon as a slot is free and the job's device bytes fit
the HBM budget beside the jobs in flight and the cached reference images,
so jobs that arrive while others are on the GPU overlap with them on the
device.  Reference B-spline coefficients stay resident in the library's
cache keyed by the job's reference key (inside the same budget, evicted only
when unused), so a stream of jobs against one reference prefilters it once.
Latency runs from a job's arrival to its records being on the host.

Results do not depend on what else runs: each job's kernels read only its
own buffers and the shared, read-only reference coefficients.
"""

import collections
import threading
import time

import numpy as np

from . import _native


class Job:
    """One correlation request.  ref_key names the reference image for the
    cache (jobs with equal keys must carry equal reference pixels)."""

    def __init__(self, job_id, ref_key, ref, tar, coords, radii, order=1, p0=None, arrival=0.0, border=20,
                 max_iterations=50, max_norm=1e-5, tolerance=0.75):
        self.job_id = job_id
        self.ref_key = ref_key
        self.ref = ref
        self.tar = tar
        self.coords = np.asarray(coords, dtype=np.float64).reshape(-1, 2)
        self.radii = np.broadcast_to(np.asarray(radii, dtype=np.int64), (len(self.coords),)).copy()
        self.order = int(order)
        self.p0 = None if p0 is None else np.asarray(p0, dtype=np.float64).reshape(len(self.coords), 12)
        self.arrival = float(arrival)
        self.border = int(border)
        self.max_iterations = int(max_iterations)
        self.max_norm = float(max_norm)
        self.tolerance = float(tolerance)

    def image_bytes(self):
        h, w = np.asarray(self.tar).shape
        hp, wp = h + 2 * self.border, w + 2 * self.border
        return hp * wp * 12  # FP64 + FP32 coefficients

    def device_bytes(self):
        """Estimate of what the job allocates besides a cached reference:
        the target's coefficients, the row-pass buffer of its slot, the
        staged pixels and the batch (per subset: centre, warps, result,
        dispatch order; templates)."""
        if getattr(self, "_bytes", None) is None:
            h, w = np.asarray(self.tar).shape
            hp, wp = h + 2 * self.border, w + 2 * self.border
            ax = np.arange(-int(self.radii.max()), int(self.radii.max()) + 1)
            d2 = ax[:, None] ** 2 + ax[None, :] ** 2
            npx = sum(int(np.count_nonzero(d2 <= r * r)) for r in np.unique(self.radii))
            self._bytes = self.image_bytes() + hp * wp * 8 + h * w * 8 + len(self.coords) * 512 + npx * 8 + (1 << 20)
        return self._bytes


class Server:
    """n_slots concurrent jobs, admitted first-come first-served within
    hbm_budget bytes (jobs in flight plus cached references), by the
    library's serving loop; this class builds each job's templates,
    submits it at its arrival and collects the records in submission
    order."""

    def __init__(self, n_slots=4, hbm_budget=64 << 30, device=0):
        self.device = device
        self.budget = int(hbm_budget)
        self.native = _native.NativeServer(n_slots=n_slots, hbm_budget=hbm_budget, device=device)
        self.keys = {}        # reference key -> the library's 64-bit cache key
        self.templates = {}   # radius -> circle template offsets
        self.results = {}
        self.errors = {}
        self.reduce = None    # job, results, latency_ms -> what results[] keeps (all of it when None)
        self.closed = False
        self.cv = threading.Condition()
        self.tickets = collections.deque()
        self.collector = threading.Thread(target=self._collect, daemon=True)
        self.collector.start()

    def _template(self, r):
        t = self.templates.get(r)
        if t is None:
            from . import Mask
            t = np.rint(Mask(mask_type="local", shape="circle", size=int(r)).coords).astype(np.int16)
            self.templates[r] = t
        return t

    # ---- admission: the library queues the job, this thread returns ----
    def submit(self, job, arrival=None):
        job.arrival = time.perf_counter() if arrival is None else arrival
        radii = np.unique(job.radii)
        lists = [self._template(int(r)) for r in radii]
        start = np.concatenate([[0], np.cumsum([len(a) for a in lists])]).astype(np.int64)
        index = np.searchsorted(radii, job.radii).astype(np.int32)
        key = self.keys.setdefault(job.ref_key, len(self.keys) + 1)
        submitted = time.perf_counter()
        ticket = self.native.submit(key, job.ref, job.tar, job.coords, job.order, index, start,
                                    np.concatenate(lists), p0=job.p0, border=job.border,
                                    precision=_native.PREC_COMP_F32, max_iterations=job.max_iterations,
                                    tolerance=job.tolerance, max_norm=job.max_norm)
        with self.cv:
            self.tickets.append((job, ticket, submitted))
            self.cv.notify_all()

    def _collect(self):
        while True:
            with self.cv:
                while not self.tickets and not self.closed:
                    self.cv.wait()
                if not self.tickets:
                    return
                job, ticket, submitted = self.tickets.popleft()
            try:
                res, lat = self.native.poll(ticket, -1)
                # the library times from dic_submit; add the wait from arrival
                lat += (submitted - job.arrival) * 1e3
                kept = (res, lat) if self.reduce is None else self.reduce(job, res, lat)
                with self.cv:
                    self.results[job.job_id] = kept
                    self.cv.notify_all()
            except Exception as exc:  # reported per job, the server keeps going
                with self.cv:
                    self.errors[job.job_id] = exc
                    self.cv.notify_all()

    def run(self, jobs, arrivals=None):
        """Submit jobs at their arrival offsets (seconds from now; all at
        once if None) and wait for all of them.  Returns {job_id: (results,
        latency_ms)}."""
        t0 = time.perf_counter()
        order = sorted(range(len(jobs)), key=lambda i: 0.0 if arrivals is None else arrivals[i])
        for i in order:
            if arrivals is not None:
                delay = t0 + arrivals[i] - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
            self.submit(jobs[i], arrival=t0 + (0.0 if arrivals is None else arrivals[i]))
        ids = [j.job_id for j in jobs]
        with self.cv:
            while not all(i in self.results or i in self.errors for i in ids):
                self.cv.wait()
            failed = {i: self.errors[i] for i in ids if i in self.errors}
        if failed:
            raise RuntimeError(f"jobs failed: {failed}")
        return {i: self.results[i] for i in ids}

    def stream(self, jobs, reduce, progress=None):
        """Serve a stream of (arrival offset in seconds, Job) pairs, each
        submitted at its arrival; jobs may be created lazily by the
        iterator.  Only reduce(job, results, latency_ms) of each finished
        job is kept, so a long stream holds no more than the jobs in the
        queue.  progress(done, submitted) is called as jobs finish.
        Returns {job_id: reduce(...)}."""
        self.reduce = reduce
        t0 = time.perf_counter()
        ids = []
        try:
            for offset, job in jobs:
                delay = t0 + offset - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                ids.append(job.job_id)
                self.submit(job, arrival=t0 + offset)
                if progress is not None and len(ids) % 100 == 0:
                    with self.cv:
                        progress(len(self.results), len(ids))
            with self.cv:
                last = -1
                while True:
                    done = sum(1 for i in ids if i in self.results or i in self.errors)
                    if progress is not None and done != last:
                        progress(done, len(ids))
                        last = done
                    if done == len(ids):
                        break
                    self.cv.wait(timeout=5.0)
                failed = {i: self.errors[i] for i in ids if i in self.errors}
        finally:
            self.reduce = None
        if failed:
            raise RuntimeError(f"jobs failed: {failed}")
        return {i: self.results.pop(i) for i in ids}

    def close(self):
        with self.cv:
            self.closed = True
            self.cv.notify_all()
        self.collector.join()
        self.native.close()
