"""Cases stepped together in a batch (pic_batch_*: each kernel of the fused
step launched once for all cases over a concatenated grid) give the same
bits as the same cases stepped alone: particles, ids, buckets, potential and
collision counts, for cases of different grids, particle counts, voltages
and pressures, in both precisions, over steps that cover both particle
buffer parities, with a case that joins the batch at the other parity.

Also: a case stepped outside its batch invalidates the batch (-71); a case
in a batch cannot be destroyed (-72); unsuitable cases are refused (-70)."""
import numpy as np
import pytest

from picmcc import Batch, Case, FP32_REFINED, FP64_PLAIN, PicError, decks

pytestmark = pytest.mark.gpu

FIELDS = ("x", "y", "vx", "vy", "vz", "id", "cell_start")


def mix():
    return [decks.small(ncx=16, ncy=12, ppc=40, volts=150.0),
            decks.small(ncx=24, ncy=20, ppc=25, volts=250.0, pres_torr=0.3),
            decks.small(ncx=9, ncy=30, ppc=60, volts=100.0, pres_torr=0.05),
            decks.ccp_deck(ncx=40, ncy=32, ppc=20, volts=200.0, cycles=1)]


def state(c):
    out = {}
    for s in range(2):
        p = c.particles(s)
        for k in FIELDS:
            out[f"{s}.{k}"] = p[k]
    for k in ("phi", "mccr", "CondCharge", "Lost", "den", "sum_den"):
        out[k] = c.array(k)
    out["tstep"] = c.info().tstep
    return out


@pytest.mark.parametrize("prec", [FP32_REFINED, FP64_PLAIN], ids=["fp32", "fp64"])
def test_batch_steps_equal_steps_alone(prec):
    opts = {"field_solver": 2} if prec == FP64_PLAIN else {}
    ds = mix()
    alone = [Case(d, prec, **opts) for d in ds]
    batched = [Case(d, prec, **opts) for d in ds]
    for c in alone + batched:
        assert c.field_solver_kind() == 1
    # the last case joins at the other buffer parity
    alone[-1].step(1)
    batched[-1].step(1)
    b = Batch(batched)
    b.step(5)
    b.step(2)
    b.sync()
    for c in alone[:-1]:
        c.step(7)
    alone[-1].step(7)
    for c in alone:
        c.sync()
    for i, (a, c) in enumerate(zip(alone, batched)):
        sa, sc = state(a), state(c)
        for k in sa:
            assert np.array_equal(sa[k], sc[k]), f"case {i}: {k} differs between the batch and the case alone"
    # the cases stay usable: more batch steps after reading them
    b.step(3)
    b.sync()
    for c in alone:
        c.step(3)
        c.sync()
    for i, (a, c) in enumerate(zip(alone, batched)):
        for s in range(2):
            assert np.array_equal(a.particles(s)["x"], c.particles(s)["x"]), f"case {i} after more steps"
    # a case in a batch is not destroyed; stepping one outside invalidates the batch
    with pytest.raises(PicError, match="belongs to a batch"):
        batched[0].close()
    batched[1].step(1)
    with pytest.raises(PicError, match="outside the batch"):
        b.step(1)
    b.close()
    for c in alone + batched:
        c.close()


def test_unsuitable_cases_are_refused():
    a = Case(decks.small(), FP32_REFINED)
    o2 = Case(decks.oxygen_deck(gas=1, ncx=8, ncy=6, ppc=10), FP32_REFINED)
    with pytest.raises(PicError, match="not an argon case"):
        Batch([a, o2])
    d = Case(decks.small(), FP64_PLAIN)   # the Jacobi PCG solve
    with pytest.raises(PicError, match="precision|separable"):
        Batch([a, d])
    with pytest.raises(PicError, match="listed twice"):
        Batch([a, a])
    for c in (a, o2, d):
        c.close()
