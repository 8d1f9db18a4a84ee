"""Serving with continuous batching on the device: a request admitted into a block that is
already iterating gets the same result as alone, a bad or non-converging request fails on its
own while the others complete, requests are admitted by the HBM budget, and latencies are
recorded."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from pylith_amd import _lib as L  # noqa: E402
from pylith_amd import serving  # noqa: E402

pytestmark = pytest.mark.gpu


def drain(st, keep=True, limit=2000):
    done = []
    for _ in range(limit):
        if not st.running():
            return done
        done += st.step(keep)
    raise AssertionError("slots did not finish")


def test_request_admitted_into_a_running_block():
    """Slot-level: B enters while A is mid-solve; both equal their solo results (a column's
    iterates do not depend on the other columns)."""
    st = serving.MeshState(8, 5, 8, 1e-11)
    rng = np.random.default_rng(2)
    A = serving.Request(0, 8, rng.normal(0, 0.5, 64))
    B = serving.Request(1, 8, rng.normal(0, 0.5, 64))
    alone = {}
    for r in (A, B):
        st.admit([r])
        (d,) = drain(st)
        alone[r.rid] = d.u.copy()
        assert d.error is None and d.stats["converged"], d.stats
    A2, B2 = serving.Request(0, 8, A.slip_values), serving.Request(1, 8, B.slip_values)
    st.admit([A2])
    for _ in range(3):
        assert not st.step(True)  # A2 still running
    st.admit([B2])
    assert B2.running_at_admission == 1
    done = drain(st)
    assert {r.rid for r in done} == {0, 1}
    for r in (A2, B2):
        d = np.abs(r.u - alone[r.rid]).max() / np.abs(alone[r.rid]).max()
        print(f"[serving] request {r.rid}: joined vs alone {d:.2e}, stats {r.stats}", flush=True)
        assert d <= 1e-12, (r.rid, d)
    st.close()


def test_results_same_alone_and_in_mix():
    sizes = (8, 12)
    eng = serving.ServingEngine(sizes, block=8, rtol=1e-11, keep_results=True)
    eng.warm()
    reqs = serving.make_requests(14, sizes, (0.5, 0.5), rate=None, seed=3)
    print(f"[serving] mix of {len(reqs)} requests", flush=True)
    mixed = eng.run(reqs)
    assert [r.rid for r in mixed] == list(range(len(reqs)))
    assert all(r.error is None for r in mixed)
    for r in mixed[:6]:
        (a,) = eng.run([serving.Request(r.rid, r.size, r.slip_values)])
        scale = np.abs(a.u).max()
        assert scale > 0
        assert np.abs(a.u - r.u).max() <= 1e-8 * scale, r.rid
    eng.close()


def test_bad_and_nonconverging_requests_fail_alone():
    # size 12 has an invalid material (the resident state cannot be built); refinement capped at
    # 2 steps with an unreachable tolerance for the size-8 requests of one engine
    eng = serving.ServingEngine((8, 12), block=8, rtol=1e-10, keep_results=True, material={12: {"vs": -100.0}})
    eng.warm()
    good = serving.make_requests(4, (8,), (1.0,), rate=None, seed=4)
    bad_slip = serving.Request(10, 8, np.full(64, np.nan))
    short = serving.Request(11, 8, np.zeros(10))
    bad_mat = serving.Request(12, 12, np.zeros(64))
    done = eng.run(good + [bad_slip, short, bad_mat])
    errs = {r.rid: r.error for r in done}
    assert all(errs[r.rid] is None for r in good)
    assert "finite" in errs[10] and "patch slips" in errs[11] and "shear-wave speed" in errs[12]
    assert abs(eng.used - eng.states[8].bytes) < 1
    eng.close()
    strict = serving.ServingEngine((8,), block=8, rtol=1e-30, maxit=2, keep_results=True)
    strict.warm()
    done = strict.run(serving.make_requests(3, (8,), (1.0,), rate=None, seed=5))
    assert all(r.error and "not converged after 2" in r.error for r in done)
    strict.close()


def test_budget_admission_serializes_outputs():
    sizes = (8,)
    eng = serving.ServingEngine(sizes, block=8, rtol=1e-10, keep_results=True)
    eng.warm()
    st = eng.states[8]
    assert st.bytes > 0  # measured from the device's free memory
    # room for the resident state and one request's output at a time
    eng.budget = eng.used + 1.5 * st.out_bytes
    reqs = serving.make_requests(5, sizes, (1.0,), rate=None, seed=4)
    done = eng.run(reqs)
    assert len(done) == 5 and all(r.error is None for r in done)
    assert all(r.running_at_admission == 0 for r in done)
    assert abs(eng.used - st.bytes) < 1
    ref = serving.ServingEngine(sizes, block=8, rtol=1e-10, keep_results=True)
    ref.warm()
    again = ref.run([serving.Request(r.rid, r.size, r.slip_values) for r in reqs])
    for a, b in zip(done, again):
        assert np.abs(a.u - b.u).max() <= 1e-8 * np.abs(b.u).max()
    tiny = serving.ServingEngine((8,), budget_bytes=1e3)
    tiny.warm()
    assert isinstance(tiny.state_errors.get(8), MemoryError)
    (r,) = tiny.run([serving.Request(0, 8, np.zeros(64))])
    assert "HBM budget" in r.error
    eng.close()
    ref.close()


def test_poisson_stream_latencies():
    sizes = (8, 12)
    eng = serving.ServingEngine(sizes, block=8, rtol=1e-10)
    eng.warm()
    reqs = serving.make_requests(40, sizes, (0.6, 0.4), rate=400.0, seed=5)
    done = eng.run(reqs)
    stats = serving.latency_stats(done)
    print(f"[serving] {stats}, joined a running block: {sum(r.running_at_admission > 0 for r in done)}",
          flush=True)
    assert stats["n"] == 40 and stats["p99_ms"] >= stats["p50_ms"] > 0
    eng.close()
