"""Host logic of the serving engine (admission, HBM budget, per-request errors, shutdown) with a
stand-in for the device state: every outcome is an error or a result, never a hang.  No GPU."""
import threading
import time

import numpy as np
import pytest

from pylith_amd import serving


class FakeState:
    """Resident state of one mesh size that finishes a request after `steps` steps; requests whose
    first slip value is 99 fail as a non-converged solve would."""
    bytes = 1000
    out_bytes = 100
    n_patches = 4
    steps = 3

    def __init__(self, size, seed, block, rtol, maxit=30, precond="amg", vs=3000.0, vs_sigma=0.2, inner_rtol=1e-4):
        if not vs > 0.0:
            raise ValueError(f"mesh size {size}: shear-wave speed {vs} is not positive")
        self.size, self.block = size, block
        self.slots = [None] * block
        self.age = [0] * block

    def free_slots(self):
        return [i for i, r in enumerate(self.slots) if r is None]

    def running(self):
        return sum(r is not None for r in self.slots)

    def admit(self, reqs):
        ids = self.free_slots()[:len(reqs)]
        running = self.running()
        for i, r in zip(ids, reqs):
            self.slots[i] = r
            self.age[i] = 0
            r.running_at_admission = running

    def step(self, keep):
        time.sleep(0.002)  # a step takes time, so later arrivals find running slots
        out = []
        for i, r in enumerate(self.slots):
            if r is None:
                continue
            self.age[i] += 1
            if self.age[i] >= self.steps:
                if r.slip_values[0] == 99:
                    r.error = "not converged after 30 refinement steps (relative residual 1.0e-02)"
                else:
                    r.u = np.full(3, float(r.rid))
                self.slots[i] = None
                out.append(r)
        return out

    def close(self):
        pass


def engine(**kw):
    return serving.ServingEngine(kw.pop("sizes", (8, 12)), block=4, keep_results=True, state_factory=FakeState, **kw)


def run_with_timeout(eng, reqs, seconds=20.0):
    out = {}
    t = threading.Thread(target=lambda: out.setdefault("done", eng.run(reqs)), daemon=True)
    t.start()
    t.join(seconds)
    assert not t.is_alive(), "the serving loop hung"
    return out["done"]


def req(rid, size, first=0.5, n=4):
    v = np.full(n, 0.1)
    v[0] = first
    return serving.Request(rid, size, v)


def test_budget_smaller_than_one_output_is_an_error_not_a_hang():
    eng = engine(budget_bytes=2 * FakeState.bytes + 50)  # resident states fit, one output does not
    eng.warm()
    done = run_with_timeout(eng, [req(0, 8), req(1, 12)])
    assert all(r.error and "cannot fit" in r.error for r in done)
    assert eng.used == 2 * FakeState.bytes


def test_bad_requests_fail_alone_and_the_rest_complete():
    eng = engine(material={12: {"vs": -5.0}})
    eng.warm()
    assert 12 in eng.state_errors and 8 in eng.states
    reqs = [req(0, 8), req(1, 8, first=np.nan), req(2, 8, n=3), req(3, 10), req(4, 12), req(5, 8, first=99),
            req(6, 8), req(7, 8)]
    done = run_with_timeout(eng, reqs)
    errs = {r.rid: r.error for r in done}
    assert errs[0] is None and errs[6] is None and errs[7] is None
    assert "finite" in errs[1]
    assert "expected 4 patch slips" in errs[2]
    assert "unknown mesh size" in errs[3]
    assert "shear-wave speed" in errs[4]
    assert "not converged" in errs[5]
    assert all(np.array_equal(r.u, np.full(3, float(r.rid))) for r in done if r.error is None)
    assert eng.used == FakeState.bytes  # every reservation released


def test_requests_join_a_running_block_and_budget_serialises():
    eng = engine(sizes=(8,))
    eng.warm()
    reqs = [serving.Request(i, 8, np.full(4, 0.1), arrival=0.002 * i) for i in range(12)]
    done = run_with_timeout(eng, reqs)
    assert all(r.error is None for r in done)
    assert any(r.running_at_admission > 0 for r in done)
    # room for one output at a time: never more than one slot in use
    eng = engine(sizes=(8,), budget_bytes=FakeState.bytes + 150)
    eng.warm()
    done = run_with_timeout(eng, [req(i, 8) for i in range(5)])
    assert all(r.error is None and r.running_at_admission == 0 for r in done)


def test_state_build_failure_of_every_size_still_returns():
    eng = engine(material={8: {"vs": 0.0}, 12: {"vs": -1.0}})
    done = run_with_timeout(eng, [req(0, 8), req(1, 12)])  # states built lazily by the workers
    assert all(r.error and "shear-wave speed" in r.error for r in done)
