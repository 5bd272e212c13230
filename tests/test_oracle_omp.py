"""The OpenMP build of the FP64 oracle (the CPU baseline's step) gives the
sequential oracle's result bit for bit: particles, ids, buckets, counters,
surface charge and potential after steps with wall losses, dielectric hits
and ionisation, at several thread counts."""
import numpy as np
import pytest

from oracle import Oracle
from picmcc import decks


def state(o):
    out = {}
    for s in range(2):
        for k, v in o.species(s).items():
            out[f"{s}.{k}"] = v.copy()
    for name in ("mccr", "CondCharge", "Lost", "sigma_sp", "phi", "den", "sum_den"):
        out[name] = o.arr(name).copy()
    return out


@pytest.mark.parametrize("threads", [2, 5])
def test_parallel_step_equals_the_sequential_step(threads):
    # high pressure and voltage: many ionisations within a few steps
    d = decks.small(ncx=20, ncy=16, ppc=40, pres_torr=0.5, volts=300.0)
    a = Oracle(d, "f64")
    b = Oracle(d, "f64_omp")
    for o in (a, b):
        o.orc_deposit()
    assert a.orc_run(12) == 0
    assert b.orc_run_par(12, threads) == 0
    sa, sb = state(a), state(b)
    assert int(sa["mccr"].reshape(-1, int(a.scalar("TnRct")))[:, 3:5].sum()) > 0, "no ionisation exercised"
    assert int(np.sum(sa["Lost"])) + int(np.sum(sa["CondCharge"])) > 0, "no wall loss exercised"
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    a.close()
    b.close()
