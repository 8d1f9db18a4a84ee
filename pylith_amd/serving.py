"""Serving loop: a stream of forward-model solve requests with continuous batching.

Each request is a quasi-static elasticity solve with prescribed fault slip on one of several
mesh sizes.  Per mesh size the engine keeps the mesh, the assembled operator, its multigrid
preconditioner and a block of k solve slots resident on the device (a Green's-function handle
over the C ABI, plx_gf_slots_*).  A worker per mesh size loops:

  1. admit: queued requests of its size enter the free slots, while the HBM budget has room for
     their results; they join the columns already iterating (continuous batching: a request does
     not wait for the running block to finish);
  2. step: one captured chunk of the FP32 block PCG for every running slot; slots whose inner
     solve converged take their FP64 refinement step and retire when the FP64 relative residual
     reaches the tolerance;
  3. retire: finished slots hand their displacement to their request and free the slot.

Workers of different sizes run concurrently on their own HIP streams, so their work overlaps on
the GPU.  Errors are per request, as PETSc's -ksp_error_if_not_converged (problems: Options.cc
:301-303) and the material queries' input checks (materials/Query.cc:223-231) make them per solve
in PyLith: an invalid request (unknown mesh size, slip of the wrong length or not finite, a result
larger than the whole budget), a solve that does not converge, or a mesh size whose resident state
cannot be built fails on its own; the engine continues with the others.  Latency is measured from
arrival to retirement.  The resident bytes of a mesh size are measured (device free memory before
and after building its state), not estimated.
"""
import collections
import threading
import time

import numpy as np
import torch

from . import _lib as L
from . import workloads

GB = 1e9


class Request:
    __slots__ = ("rid", "size", "slip_values", "arrival", "start", "done", "u", "stats", "error",
                 "running_at_admission", "batch")

    def __init__(self, rid, size, slip_values, arrival=0.0):
        self.rid, self.size, self.slip_values, self.arrival = rid, size, slip_values, arrival
        self.start = self.done = None
        self.u = None
        self.stats = None
        self.error = None
        self.running_at_admission = 0  # other slots of its mesh size iterating when it entered
        self.batch = 0                  # slots running after its admission (it included)


class MeshState:
    """Resident device state of one mesh size: DM, fault, operator, preconditioner, k slots."""

    def __init__(self, size, seed, block, rtol, maxit=30, n_patch=(8, 8), precond="amg", vs=3000.0,
                 vs_sigma=0.2, inner_rtol=1e-4):
        from .greensfns import GreensFunctions
        if not (np.isfinite(vs) and vs > 0.0):  # materials/Query.cc:223-231: Vs must be positive
            raise ValueError(f"mesh size {size}: shear-wave speed {vs} is not positive")
        free0 = torch.cuda.mem_get_info()[0]
        self.size = size
        self.prob = workloads.faulted_box(size, seed=seed, vs_sigma=vs_sigma, vs=vs)
        # a stream of its own (default priority): mesh sizes served by different workers do not
        # wait for each other's work on a shared stream
        self.stream = torch.cuda.Stream()
        self.prob.dm.set_stream(self.stream)
        self.gf = GreensFunctions(self.prob.dm)
        self.gf.set_preconditioner(precond)
        self.gf.set_options(block=block, inner_rtol=inner_rtol)
        self.gf.slots_open(block, rtol, maxit)
        self.block = block
        self.rtol = rtol
        self.patch = workloads.patch_index(self.prob.fault, *n_patch)
        self.n_patches = n_patch[0] * n_patch[1]
        self.slots = [None] * block
        torch.cuda.synchronize()
        self.bytes = max(free0 - torch.cuda.mem_get_info()[0], 0)
        self.out_bytes = self.prob.local_size * 8

    def free_slots(self):
        return [i for i, r in enumerate(self.slots) if r is None]

    def running(self):
        return sum(r is not None for r in self.slots)

    def admit(self, reqs):
        ids = self.free_slots()[:len(reqs)]
        slip = workloads.patch_slip(self.prob.fault, self.patch, np.stack([r.slip_values for r in reqs]))
        running = self.running()
        with torch.cuda.stream(self.stream):
            self.gf.slots_load(ids, torch.from_numpy(slip).cuda())
        for i, r in zip(ids, reqs):
            self.slots[i] = r
            r.running_at_admission = running
        for r in reqs:
            r.batch = self.running()

    def step(self, keep):
        """One step; returns the requests that finished (with their result or error)."""
        states = self.gf.slots_step()
        out = []
        for i, st in enumerate(states):
            if st in (L.SLOT_CONVERGED, L.SLOT_FAILED):
                r = self.slots[i]
                with torch.cuda.stream(self.stream):
                    u = torch.empty(self.prob.local_size, dtype=torch.float64, device="cuda") if keep else None
                    r.stats = self.gf.slot_take(i, u)
                if st == L.SLOT_FAILED:
                    r.error = (f"not converged after {r.stats['outer_iterations']} refinement steps "
                               f"(relative residual {r.stats['relative_residual']:.3e})")
                elif keep:
                    with torch.cuda.stream(self.stream):
                        r.u = u.cpu().numpy()
                self.slots[i] = None
                out.append(r)
        return out

    def close(self):
        self.gf.close()
        self.prob.dm.close()


class ServingEngine:
    def __init__(self, sizes, budget_bytes=200 * GB, block=16, rtol=1e-9, maxit=30, seed=20261202,
                 keep_results=False, precond="amg", material=None, state_factory=None, inner_rtol=1e-4):
        self.state_factory = state_factory or MeshState  # the resident state class of a mesh size
        self.inner_rtol = inner_rtol
        self.sizes = list(sizes)
        self.precond = precond
        self.budget = budget_bytes
        self.block = block
        self.rtol = rtol
        self.maxit = maxit
        self.seed = seed
        self.keep = keep_results
        self.material = material or {}  # size -> dict(vs=..., vs_sigma=...) for the resident state
        self.used = 0.0
        self.cv = threading.Condition()
        self.states = {}
        self.state_errors = {}
        self.queues = {s: collections.deque() for s in self.sizes}
        self.workers = {}
        self.completed = []
        self.errors = []
        self.stop = False
        self.t0 = time.perf_counter()

    # --- resident state, admitted by HBM budget ------------------------------------------------
    def state(self, size):
        if size in self.state_errors:
            raise self.state_errors[size]
        if size not in self.states:
            try:
                st = self.state_factory(size, self.seed + size, self.block, self.rtol, self.maxit,
                                        precond=self.precond, inner_rtol=self.inner_rtol,
                                        **self.material.get(size, {}))
            except Exception as e:
                self.state_errors[size] = e
                raise
            with self.cv:
                if self.used + st.bytes > self.budget:
                    st.close()
                    err = MemoryError(f"resident state of mesh size {size} ({st.bytes / GB:.2f} GB) exceeds the "
                                      f"HBM budget")
                    self.state_errors[size] = err
                    raise err
                self.used += st.bytes
            self.states[size] = st
        return self.states[size]

    def warm(self):
        """Create the resident state of every size (set-up is not part of request latency); a size
        whose state cannot be built is remembered, and its requests fail with that error."""
        for s in self.sizes:
            try:
                self.state(s)
            except Exception:
                pass

    # --- requests ---------------------------------------------------------------------------------
    def _finish(self, r, error=None):
        if error is not None:
            r.error = error
        r.done = time.perf_counter() - self.t0
        with self.cv:
            self.completed.append(r)
            self.cv.notify_all()

    def _validate(self, r):
        if r.size not in self.queues:
            return f"unknown mesh size {r.size}"
        v = np.asarray(r.slip_values, float)
        if v.ndim != 1 or not np.all(np.isfinite(v)):
            return "slip values must be a finite vector"
        if r.size in self.states and v.size != self.states[r.size].n_patches:
            return f"expected {self.states[r.size].n_patches} patch slips, got {v.size}"
        return None

    def _worker(self, size):
        try:
            st = self.state(size)
        except Exception as e:  # every request of this size fails with the set-up error
            st, err = None, f"mesh size {size}: {e}"
        q = self.queues[size]
        while True:
            admit = []
            with self.cv:
                while True:
                    if st is None:
                        failed = list(q)
                        q.clear()
                        break
                    free = len(st.free_slots())
                    resident = sum(x.bytes for x in self.states.values())
                    while q and len(admit) < free:
                        if st.out_bytes > self.budget - resident:  # can never fit
                            admit.append(("error", q.popleft()))
                            continue
                        if self.used + st.out_bytes > self.budget:
                            break
                        self.used += st.out_bytes
                        admit.append(("ok", q.popleft()))
                    if admit or st.running():
                        break
                    if self.stop and not q:
                        return
                    self.cv.wait(0.01)
            if st is None:
                for r in failed:
                    self._finish(r, err)
                with self.cv:
                    if self.stop and not q:
                        return
                    self.cv.wait(0.01)
                continue
            ok = [r for tag, r in admit if tag == "ok"]
            for tag, r in admit:
                if tag == "error":
                    self._finish(r, f"the result ({st.out_bytes / GB:.3f} GB) cannot fit the HBM budget")
            done = []
            if ok:
                now = time.perf_counter() - self.t0
                for r in ok:
                    r.start = now
                try:
                    st.admit(ok)
                except Exception as e:  # these requests fail; the running slots go on
                    for r in ok:
                        r.error = f"admission failed: {e}"
                    done += ok
            try:
                done += st.step(self.keep) if st.running() else []
            except Exception as e:  # the handle failed: fail what it holds, keep serving the rest
                held = [r for r in st.slots if r is not None]
                st.slots = [None] * st.block
                for r in held:
                    r.error = f"solve failed: {e}"
                done += held
                self.errors.append(e)
            for r in done:
                with self.cv:
                    self.used -= st.out_bytes
                self._finish(r)

    def run(self, requests):
        """Replay `requests` (arrival times in seconds from now) and return them completed (each
        with its result or its error)."""
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
            err = self._validate(r)
            if err:
                self._finish(r, err)
                continue
            with self.cv:
                self.queues[r.size].append(r)
                self.cv.notify_all()
        with self.cv:
            while len(self.completed) < len(requests):
                self.cv.wait(0.05)
            self.stop = True
            self.cv.notify_all()
        for t in self.workers.values():
            t.join()
        return sorted(self.completed, key=lambda r: r.rid)

    def close(self):
        for st in self.states.values():
            st.close()
        self.states.clear()


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
    ok = [r for r in done if r.error is None]
    lat = np.array([r.done - r.arrival for r in ok])
    span = max(r.done for r in ok) - min(r.arrival for r in ok)
    return {"p50_ms": 1e3 * float(np.percentile(lat, 50)), "p99_ms": 1e3 * float(np.percentile(lat, 99)),
            "solves_per_s": len(ok) / span, "n": len(ok), "errors": len(done) - len(ok)}
