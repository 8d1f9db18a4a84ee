"""Serving loop with continuous batching: every request's result is the same alone and in the
mix (within the refinement tolerance), requests are admitted by the HBM budget, and latencies
are recorded."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from pylith_amd import serving  # noqa: E402

pytestmark = pytest.mark.gpu


def test_results_same_alone_and_in_mix():
    sizes = (8, 12)
    eng = serving.ServingEngine(sizes, block=8, rtol=1e-11, keep_results=True)
    eng.warm()
    reqs = serving.make_requests(14, sizes, (0.5, 0.5), rate=None, seed=3)
    print(f"[serving] mix of {len(reqs)} requests", flush=True)
    mixed = eng.run(reqs)
    assert [r.rid for r in mixed] == list(range(len(reqs)))
    # batches formed: more requests than blocks
    assert any(r.stats["num_blocks"] >= 1 for r in mixed)
    for r in mixed[:6]:
        alone = serving.Request(r.rid, r.size, r.slip_values)
        (a,) = eng.run([alone])
        scale = np.abs(a.u).max()
        assert scale > 0
        assert np.abs(a.u - r.u).max() <= 1e-8 * scale, r.rid


def test_budget_admission_serializes_outputs():
    sizes = (8,)
    eng = serving.ServingEngine(sizes, block=8, rtol=1e-10, keep_results=True)
    eng.warm()
    st = eng.states[8]
    # room for the resident state and one request's output at a time
    eng.budget = eng.used + 1.5 * st.prob.local_size * 8
    reqs = serving.make_requests(5, sizes, (1.0,), rate=None, seed=4)
    done = eng.run(reqs)
    assert len(done) == 5
    # the budget admits one request's output at a time: blocks of one
    assert all(r.batch == 1 for r in done)
    assert abs(eng.used - st.bytes) < 1
    ref = serving.ServingEngine(sizes, block=8, rtol=1e-10, keep_results=True)
    ref.warm()
    again = ref.run([serving.Request(r.rid, r.size, r.slip_values) for r in reqs])
    for a, b in zip(done, again):
        assert np.abs(a.u - b.u).max() <= 1e-8 * np.abs(b.u).max()
    with pytest.raises(MemoryError):
        serving.ServingEngine((8,), budget_bytes=1e3).warm()


def test_poisson_stream_latencies():
    sizes = (8, 12)
    eng = serving.ServingEngine(sizes, block=8, rtol=1e-10)
    eng.warm()
    reqs = serving.make_requests(20, sizes, (0.6, 0.4), rate=200.0, seed=5)
    done = eng.run(reqs)
    stats = serving.latency_stats(done)
    print(f"[serving] {stats}", flush=True)
    assert stats["n"] == 20 and stats["p99_ms"] >= stats["p50_ms"] > 0
