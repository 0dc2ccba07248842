"""Serving loop: continuous batching of correlation jobs on one GPU.

Jobs (a reference/target image pair, subset centres with per-subset radii,
warp order, initial warps) arrive over time.  The server admits a job as
soon as a slot is free and the job's device-memory estimate fits the HBM
budget beside the jobs in flight and the cached reference images; each
admitted job runs on its slot's engine (its own HIP stream), so jobs that
arrive while others are on the GPU overlap with them on the device.
Reference B-spline coefficients stay resident in an LRU cache keyed by the
job's reference key (inside the same budget, evicted only when unused), so
a stream of jobs against one reference prefilters it once.  Latency is
measured from a job's arrival to its results being on the host.

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
    hbm_budget bytes (jobs in flight plus cached references)."""

    def __init__(self, n_slots=4, hbm_budget=64 << 30, device=0):
        self.device = device
        self.budget = int(hbm_budget)
        self.engines = [_native.Engine(device) for _ in range(n_slots)]
        self.cache_engine = _native.Engine(device)
        self.cache = collections.OrderedDict()  # key -> [DeviceImage, bytes, users]
        self.cache_bytes = 0
        self.in_flight = 0
        self.pending = collections.deque()
        self.cv = threading.Condition()
        self.results = {}
        self.errors = {}
        self.reduce = None  # job, results, latency_ms -> what results[] keeps (all of it when None)
        self.closed = False
        self.free_slots = list(range(n_slots))
        self.workers = [threading.Thread(target=self._worker, daemon=True) for _ in range(n_slots)]
        for t in self.workers:
            t.start()

    # ---- admission ----
    def submit(self, job, arrival=None):
        with self.cv:
            job.arrival = time.perf_counter() if arrival is None else arrival
            self.pending.append(job)
            self.cv.notify_all()

    def _fits(self, job):
        need = job.device_bytes() + (0 if job.ref_key in self.cache else job.image_bytes())
        free = self.budget - self.in_flight - self.cache_bytes
        if need <= free:
            return True
        # evict unused references, least recently used first
        for key in list(self.cache):
            if need <= free:
                break
            entry = self.cache[key]
            if entry[2] == 0 and key != job.ref_key:
                del self.cache[key]
                self.cache_bytes -= entry[1]
                free += entry[1]
        return need <= free

    def _next(self):
        """Under the lock: the oldest pending job, if a slot is free and it
        fits (first come, first served: no job overtakes the oldest).  A job
        too large for the budget even alone runs when nothing else does."""
        if not self.free_slots or not self.pending:
            return None
        job = self.pending[0]
        if self._fits(job) or (self.in_flight == 0 and all(e[2] == 0 for e in self.cache.values())):
            self.pending.popleft()
            return job
        return None

    def _reference(self, job):
        with self.cv:
            entry = self.cache.get(job.ref_key)
            if entry is not None:
                self.cache.move_to_end(job.ref_key)
                entry[2] += 1
                return entry[0]
        img = _native.DeviceImage(self.cache_engine, pixels=np.asarray(job.ref), border=job.border)
        self.cache_engine.sync()
        with self.cv:
            entry = self.cache.get(job.ref_key)
            if entry is None:
                entry = [img, job.image_bytes(), 0]
                self.cache[job.ref_key] = entry
                self.cache_bytes += entry[1]
            entry[2] += 1
            return entry[0]

    def _release(self, job):
        with self.cv:
            entry = self.cache.get(job.ref_key)
            if entry is not None:
                entry[2] -= 1

    # ---- execution ----
    def _worker(self):
        while True:
            with self.cv:
                job = None
                while job is None:
                    if self.closed and not self.pending:
                        return
                    job = self._next()
                    if job is None:
                        self.cv.wait()
                slot = self.free_slots.pop()
                cost = job.device_bytes()
                self.in_flight += cost
            try:
                res = self._run(job, self.engines[slot])
                done = time.perf_counter()
                lat = (done - job.arrival) * 1e3
                kept = (res, lat) if self.reduce is None else self.reduce(job, res, lat)
                with self.cv:
                    self.results[job.job_id] = kept
            except Exception as exc:  # reported per job, the server keeps going
                with self.cv:
                    self.errors[job.job_id] = exc
            finally:
                with self.cv:
                    self.in_flight -= cost
                    self.free_slots.append(slot)
                    self.cv.notify_all()

    def _run(self, job, eng):
        from . import Mask

        f = self._reference(job)
        try:
            g = _native.DeviceImage(eng, pixels=np.asarray(job.tar), border=job.border)
            # one circle template per distinct radius
            radii = np.unique(job.radii)
            lists = [np.rint(Mask(mask_type="local", shape="circle", size=int(r)).coords).astype(np.int16)
                     for r in radii]
            start = np.concatenate([[0], np.cumsum([len(a) for a in lists])]).astype(np.int64)
            index = np.searchsorted(radii, job.radii).astype(np.int32)
            b = _native.Batch(eng, job.coords, job.order, index, start, np.concatenate(lists))
            if job.p0 is not None:
                b.set_p0(job.p0)
            b.solve(f, g, method="icgn", precision=_native.PREC_COMP_F32, max_iterations=job.max_iterations,
                    tolerance=job.tolerance, max_norm=job.max_norm)
            eng.sync()
            out = b.results()
            del b, g
            return out
        finally:
            self._release(job)

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
        for t in self.workers:
            t.join()
