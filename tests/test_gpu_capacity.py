"""A case that outgrows its particle capacity reports it and stays in bounds.

dt = T_RF/60 at 0.36 Torr and 170 V makes this small argon discharge run
away (the CPU oracle overflows its capacity at step 19).  The library must
report the overflow (-30) at the next synchronisation, keep every kernel
inside the particle arrays (cell offsets past the capacity are clamped),
and still return at most `capacity` particles.
"""
import pytest

from picmcc import Case, FP32_REFINED, PicError, decks

pytestmark = pytest.mark.gpu


def test_capacity_overflow_is_reported_and_contained():
    d = decks.ccp_deck(ncx=24, ncy=32, dx=5e-4, volts=168.5, pres_torr=0.36, ppc=28, steps_per_cycle=60,
                       cycles=1, seed=1000, philox_seed=5000)
    c = Case(d, FP32_REFINED)
    c.step(40)
    with pytest.raises(PicError) as e:
        c.sync()
    assert e.value.code == -30
    inf = c.info()
    for s in range(inf.nsp):
        assert inf.np[s] <= inf.cap[s]
        p = c.particles(s)
        assert len(p["x"]) == inf.np[s]
    c.close()
