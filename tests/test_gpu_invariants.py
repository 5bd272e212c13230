"""Size-independent checks of the fused step at the full 256x256 argon
workload (26 M particles), where the CPU oracle is too slow to follow, on
both paths (FP32 with the FP64-refined field solve; the FP64 plain path
with the reference's Jacobi PCG in FP64).

After 20 (FP32) or 8 (FP64) fused steps (graph replays):
- the particle balance closes exactly for both species:
  N = N0 + ionisations - absorbed - deleted at walls;
- the buckets are well formed: cell offsets non-decreasing, the last equal to
  the particle count, every in-cell offset in [0, 1);
- particle ids are unique per species and below the species' id counter;
- the same steps run kernel by kernel (no graph) give the same bits.
"""
import numpy as np
import pytest

from picmcc import Case, FP32_REFINED, FP64_PLAIN, decks

pytestmark = pytest.mark.gpu


def counts(c):
    inf = c.info()
    m = c.array("mccr").reshape(-1, inf.TnRct).astype(np.int64)
    pairs = int(m[:, 3].sum() + m[:, 4].sum())
    absorbed = c.array("CondCharge").reshape(inf.nsp, -1).astype(np.int64).sum(axis=1)
    lost = c.array("Lost").astype(np.int64)
    return [int(inf.np[s]) for s in range(inf.nsp)], pairs, absorbed, lost


@pytest.mark.parametrize("prec,STEPS", [(FP32_REFINED, 20), (FP64_PLAIN, 8)], ids=["fp32_refined", "fp64_plain"])
def test_full_size_invariants_and_graph_replay(prec, STEPS):
    d = decks.config2()
    g = Case(d, prec)
    n0, _, _, _ = counts(g)
    g.step(STEPS)
    g.sync()
    n1, pairs, absorbed, lost = counts(g)
    assert pairs > 0
    for s in range(2):
        assert n1[s] == n0[s] + pairs - absorbed[s] - lost[s], f"species {s} balance"
    nid = g.array("next_id")
    parts = [g.particles(s) for s in range(2)]
    for s, p in enumerate(parts):
        cs = p["cell_start"]
        assert cs[0] == 0 and cs[-1] == n1[s] and np.all(np.diff(cs) >= 0)
        for k in ("x", "y"):
            v = p[k]
            assert np.all((v >= 0) & (v < 1)), f"species {s}: in-cell offset {k} outside [0, 1)"
        ids = p["id"]
        assert len(np.unique(ids)) == len(ids), f"species {s}: duplicate ids"
        assert int(ids.max()) < int(nid[s])
    g.close()
    # the same steps launched kernel by kernel (stage timers force the eager path)
    e = Case(d, prec)
    e.stage_timing(True)
    e.step(STEPS)
    e.sync()
    for s in range(2):
        q = e.particles(s)
        for k in ("x", "y", "vx", "vy", "vz", "id", "cell_start"):
            assert np.array_equal(parts[s][k], q[k]), f"species {s} {k}: graph replay and eager steps differ"
    e.close()
