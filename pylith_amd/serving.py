"""Serving loop: a stream of forward-model solve requests with continuous batching.

Each request is a quasi-static elasticity solve with prescribed fault slip on one of several
mesh sizes.  Per mesh size the engine keeps the mesh, the assembled operator and its
preconditioner resident on the device (a Green's-function handle over the C ABI), admitted
against an HBM budget.  One worker per mesh size takes whatever requests of its size are
waiting when it becomes free (up to the block width) and solves them as one block; workers of
different sizes run concurrently on their own HIP streams, so their work overlaps on the GPU.
Requests that arrive meanwhile wait for the next block of their size.  Latency is measured from
arrival to the completion of the request's block.
"""
import collections
import threading
import time

import numpy as np
import torch

from . import workloads

GB = 1e9


class Request:
    __slots__ = ("rid", "size", "slip_values", "arrival", "start", "done", "u", "stats", "batch")

    def __init__(self, rid, size, slip_values, arrival=0.0):
        self.rid, self.size, self.slip_values, self.arrival = rid, size, slip_values, arrival
        self.start = self.done = None
        self.batch = 0
        self.u = None
        self.stats = None


class MeshState:
    """Resident device state of one mesh size: DM, fault, operator, block vectors."""

    def __init__(self, size, seed, block, rtol, n_patch=(8, 8), precond="amg"):
        from .greensfns import GreensFunctions
        self.size = size
        self.prob = workloads.faulted_box(size, seed=seed, vs_sigma=0.2)
        self.gf = GreensFunctions(self.prob.dm)
        self.gf.set_options(block=block)
        self.gf.set_preconditioner(precond)
        self.block = block
        self.rtol = rtol
        self.patch = workloads.patch_index(self.prob.fault, *n_patch)
        self.n_patches = n_patch[0] * n_patch[1]
        n = self.prob.mesh.num_nodes
        nnzb = 27 * n
        # operator (FP64 + FP32), block vectors (3 FP64 + 5 FP32 per column), DM and plan
        self.bytes = nnzb * 9 * 12 + 3 * n * block * (3 * 8 + 5 * 4) + 3 * n * 8 * 8 + self.prob.mesh.num_cells * 64
        if precond == "amg":  # fine-level V-cycle workspace (2 x 64 FP32 columns) + transfer operators
            self.bytes += 2 * 3 * n * 64 * 4 + 2 * n * 5 * 18 * 4

    def solve(self, reqs):
        slip = workloads.patch_slip(self.prob.fault, self.patch, np.stack([r.slip_values for r in reqs]))
        s = torch.from_numpy(slip).cuda()
        u = torch.zeros(len(reqs), self.prob.local_size, dtype=torch.float64, device="cuda")
        # the narrowest block width that holds the batch: no work on empty columns
        width = next(k for k in (8, 16, 32, 64) if k >= min(len(reqs), self.block))
        self.gf.set_options(block=width)
        st = self.gf.solve(s, u, rtol=self.rtol, maxit=30)
        return u, st


class ServingEngine:
    def __init__(self, sizes, budget_bytes=200 * GB, block=16, rtol=1e-10, seed=20261202, keep_results=False,
                 precond="amg"):
        self.sizes = list(sizes)
        self.precond = precond
        self.budget = budget_bytes
        self.block = block
        self.rtol = rtol
        self.seed = seed
        self.keep = keep_results
        self.used = 0.0
        self.cv = threading.Condition()
        self.states = {}
        self.queues = {s: collections.deque() for s in self.sizes}
        self.workers = {}
        self.completed = []
        self.errors = []
        self.stop = False
        self.t0 = time.perf_counter()

    # --- resident state, admitted by HBM budget ------------------------------------------------
    def state(self, size):
        if size not in self.states:
            st = MeshState(size, self.seed + size, self.block, self.rtol, precond=self.precond)
            with self.cv:
                if self.used + st.bytes > self.budget:
                    raise MemoryError(f"resident state of mesh size {size} exceeds the HBM budget")
                self.used += st.bytes
            self.states[size] = st
        return self.states[size]

    # --- workers ----------------------------------------------------------------------------------
    def _take(self, size, per_request):
        """Wait for requests of `size`; take as many as are waiting, up to the block width and
        to what the HBM budget admits (their outputs are reserved until the block is done)."""
        q = self.queues[size]
        with self.cv:
            while True:
                if q:
                    room = int((self.budget - self.used) // per_request)
                    n = min(len(q), self.block, room)
                    if n > 0:
                        batch = [q.popleft() for _ in range(n)]
                        self.used += n * per_request
                        return batch
                elif self.stop:
                    return None
                self.cv.wait(0.01)

    def _release(self, nbytes):
        with self.cv:
            self.used -= nbytes
            self.cv.notify_all()

    def _worker(self, size):
        try:
            st = self.state(size)
            per = st.prob.local_size * 8
            while True:
                batch = self._take(size, per)
                if batch is None:
                    return
                now = time.perf_counter() - self.t0
                for r in batch:
                    r.start = now
                    r.batch = len(batch)
                u, stats = st.solve(batch)
                done = time.perf_counter() - self.t0
                for i, r in enumerate(batch):
                    r.done = done
                    r.stats = stats
                    r.u = u[i].cpu().numpy() if self.keep else None
                del u
                self._release(len(batch) * per)
                with self.cv:
                    self.completed.extend(batch)
                    self.cv.notify_all()
        except Exception as e:  # surfaced by run()
            self.errors.append(e)

    def warm(self):
        """Create the resident state of every size (set-up is not part of request latency)."""
        for s in self.sizes:
            self.state(s)

    def run(self, requests):
        """Replay `requests` (arrival times in seconds from now) and return them completed."""
        self.stop = False
        self.completed = []
        self.t0 = time.perf_counter()
        for s in self.sizes:
            t = threading.Thread(target=self._worker, args=(s,), daemon=True)
            self.workers[s] = t
            t.start()
        for r in sorted(requests, key=lambda r: r.arrival):
            delay = r.arrival - (time.perf_counter() - self.t0)
            if delay > 0:
                time.sleep(delay)
            with self.cv:
                self.queues[r.size].append(r)
                self.cv.notify_all()
        while True:
            with self.cv:
                n = len(self.completed)
            if n >= len(requests) or self.errors:
                break
            time.sleep(0.005)
        with self.cv:
            self.stop = True
            self.cv.notify_all()
        for t in self.workers.values():
            t.join()
        if self.errors:
            raise self.errors[0]
        return sorted(self.completed, key=lambda r: r.rid)


def make_requests(n, sizes, probs, rate, n_patches=64, sigma=0.5, seed=7):
    """Poisson arrivals at `rate` per second (rate None: all at t = 0), mesh size drawn with
    `probs`, Gaussian patch slip (sigma m)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        if rate:
            t += rng.exponential(1.0 / rate)
        size = sizes[rng.choice(len(sizes), p=probs)]
        out.append(Request(i, size, rng.normal(0.0, sigma, n_patches), t if rate else 0.0))
    return out


def latency_stats(done):
    lat = np.array([r.done - r.arrival for r in done])
    span = max(r.done for r in done) - min(r.arrival for r in done)
    return {"p50_ms": 1e3 * float(np.percentile(lat, 50)), "p99_ms": 1e3 * float(np.percentile(lat, 99)),
            "solves_per_s": len(done) / span, "n": len(done)}
