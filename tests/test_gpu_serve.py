"""Ensemble serving on one GPU.

Every case must give the same bits alone and when served (the step is
deterministic per case), every request must finish, and the latency
statistics must be consistent with the arrivals.  The concurrent mode is
switched off; its test stays skipped with the reason.
"""
import numpy as np
import pytest

from picmcc import serve

pytestmark = pytest.mark.gpu


def test_sequential_serving_is_fifo_and_exact():
    reqs = serve.ar_ensemble(4, seed=5, rate_hz=50.0, grid=(16, 40), ppc=(10, 30), cycles=(1, 1),
                             steps_per_cycle=400)
    res = serve.SequentialServer().run(reqs, keep_state=lambda rid: rid < 2)
    assert [r["rid"] for r in res] == sorted(range(4), key=lambda i: reqs[i].arrival)
    prev_finish = 0.0
    for r in res:
        assert r["admit"] >= r["arrival"] - 1e-3 and r["admit"] >= prev_finish
        assert r["latency"] >= r["service"] > 0.0
        prev_finish = r["finish"]
    s = serve.summarize(res)
    assert s["cases"] == 4 and s["p99_latency_s"] >= s["p50_latency_s"] > 0.0
    assert 0.0 < s["busy_fraction"] <= 1.0 + 1e-9
    by_id = {r["rid"]: r for r in res}
    for req in reqs[:2]:
        alone = serve.run_alone(req, keep_state=True)
        got = by_id[req.rid]["state"]
        assert np.array_equal(alone["state"]["den"], got["den"]), f"case {req.rid}: density differs"
        for s_ in range(2):
            for k in ("x", "vx", "id"):
                assert np.array_equal(alone["state"]["particles"][s_][k], got["particles"][s_][k])


@pytest.mark.skipif(not serve.CONCURRENT, reason=serve.DISABLED_REASON)
def test_results_same_alone_and_in_the_mix():
    # dt = T_RF/400: these decks stay within their particle capacity (checked
    # with the CPU oracle); a much coarser dt makes them run away
    reqs = serve.ar_ensemble(6, seed=3, rate_hz=200.0, grid=(16, 40), ppc=(10, 30), cycles=(1, 1),
                             steps_per_cycle=400)
    srv = serve.EnsembleServer(slice_steps=100, max_active=4)
    res = srv.run(reqs, keep_state=lambda rid: True)
    assert sorted(r["rid"] for r in res) == list(range(6))
    for r in res:
        assert r["finish"] >= r["admit"] >= 0.0
        assert r["latency"] > 0.0
    by_id = {r["rid"]: r for r in res}
    for req in reqs[:3]:
        alone = serve.run_alone(req, keep_state=True)
        mix = by_id[req.rid]["state"]
        assert np.array_equal(alone["state"]["den"], mix["den"]), f"case {req.rid}: density differs"
        for s, (a, b) in enumerate(zip(alone["state"]["particles"], mix["particles"])):
            for k in ("x", "y", "vx", "vy", "vz", "id", "cell_start"):
                assert np.array_equal(a[k], b[k]), f"case {req.rid}: species {s} {k} differs"
    summ = serve.summarize(res)
    assert summ["p99_latency_s"] >= summ["p50_latency_s"] > 0.0
