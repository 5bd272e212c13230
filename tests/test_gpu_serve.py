"""Ensemble serving on one GPU.

Every case must give the same bits alone and when served (the step is
deterministic per case), every request must finish, and the latency
statistics must be consistent with the arrivals.  The served ensembles are
argon-only and mixed argon and Ar/O2 (the serving workload's generator at a
small grid size), served concurrently (each case on its own stream, or the
argon cases stepped together through a pic_batch) and first come first
served.

A served case is not compared with the oracle's orc_run directly: the two
deposits sum the charge in different orders, so the potentials differ at the
solver tolerance and the trajectories separate.  The chain that ties a
served case to the oracle is: served = alone (here, bit for bit); the fused
step of a case alone = its staged stages (test_gpu_stages.py); each staged
stage = the oracle's stage from the same state (test_gpu_stages.py,
test_gpu_baseline_decks.py, test_gpu_aro2.py).
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
@pytest.mark.parametrize("batch", [False, True], ids=["streams", "batched"])
def test_results_same_alone_and_in_the_mix(batch):
    # dt = T_RF/400: these decks stay within their particle capacity (checked
    # with the CPU oracle); a much coarser dt makes them run away
    reqs = serve.ar_ensemble(6, seed=3, rate_hz=200.0, grid=(16, 40), ppc=(10, 30), cycles=(1, 1),
                             steps_per_cycle=400)
    srv = serve.EnsembleServer(slice_steps=100, max_active=4, batch=batch)
    res = srv.run(reqs, keep_state=lambda rid: True)
    if batch:
        assert srv.rebuilds >= 2, "the batched server never batched"
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


@pytest.mark.parametrize("batch", [False, True], ids=["streams", "batched"])
def test_mixed_argon_and_aro2_cases_same_alone_and_in_the_mix(batch):
    """Argon and Ar/O2 cases of the serving workload's generator (grids of
    24-40 cells, one RF cycle at dt = T/400), served four at a time: every
    species of every case has the same bits as the case run alone."""
    reqs = serve.config5_requests(6, seed=11, rate_hz=100.0, grid=(24, 40), ppc=(10, 20), cycles=(1, 1),
                                  dx=5e-4)
    assert {serve.case_gas(r.deck) for r in reqs} == {0, 2}
    res = serve.EnsembleServer(slice_steps=100, max_active=4, batch=batch).run(reqs, keep_state=lambda rid: True)
    assert sorted(r["rid"] for r in res) == list(range(6))
    assert all(r["error"] == 0 for r in res)
    by_id = {r["rid"]: r for r in res}
    for req in reqs:
        alone = serve.run_alone(req, keep_state=True)
        assert alone["error"] == 0
        mix = by_id[req.rid]["state"]
        assert np.array_equal(alone["state"]["den"], mix["den"]), f"case {req.rid}: density differs"
        for s, (a, b) in enumerate(zip(alone["state"]["particles"], mix["particles"])):
            for k in ("x", "y", "vx", "vy", "vz", "id", "cell_start"):
                assert np.array_equal(a[k], b[k]), f"case {req.rid} (gas {serve.case_gas(req.deck)}): species {s} {k}"
