"""A case that outgrows its particle capacity reports it and stays in bounds.

dt = T_RF/60 at 0.36 Torr and 170 V makes this small argon discharge run
away (the CPU oracle overflows its capacity at step 19).  The library must
report the overflow (-30) at the next synchronisation, keep every kernel
inside the particle arrays (cell offsets past the capacity are clamped),
and still return at most `capacity` particles.  A field solve abandoned at a
timed-out grid barrier is reported the same way (-34).
"""
import pytest

from picmcc import Case, FP32_REFINED, FP64_PLAIN, PicError, decks

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


def test_abandoned_field_solve_is_reported_and_cleared(monkeypatch):
    """The persistent solver's grid barrier waits are bounded.  With the
    limit forced to zero from the host (PIC_PCG_SPIN_US, read when the case
    is created), a multi-block FP64 PCG solve abandons at its first wait:
    every block leaves the kernel, pic_sync reports -34 and clears the
    barrier, and the next solve on the same case ends the same way instead of
    hanging.  A case created without the override solves normally."""
    d = decks.small(ncx=24, ncy=24, ppc=10, volts=150.0)
    monkeypatch.setenv("PIC_PCG_SPIN_US", "0")
    c = Case(d, FP64_PLAIN)
    monkeypatch.delenv("PIC_PCG_SPIN_US")
    try:
        for _ in range(2):
            c.advance_time()
            c.field_solve()
            with pytest.raises(PicError) as e:
                c.sync()
            assert e.value.code == -34
    finally:
        c.close()
    ok = Case(d, FP64_PLAIN)
    try:
        ok.advance_time()
        ok.field_solve()
        ok.sync()
        it, outer, rel = ok.last_field_solve()
        assert it > 0 and rel <= 1e-10
    finally:
        ok.close()


def test_field_solver_launch_shape_is_tuned_and_kept(monkeypatch):
    """The persistent PCG's launch shape is tuned at case creation on the
    case's own operator: threads per block 128, 256 or 512, a grid that the
    device holds at once and no larger than the unknowns need; a second case
    of the same size reuses it; PIC_PCG_BLOCKS fixes the grid instead.  The
    tuned solver converges to the requested tolerance."""
    d = decks.small(ncx=48, ncy=40, ppc=10, volts=150.0)
    monkeypatch.delenv("PIC_PCG_BLOCKS", raising=False)
    a = Case(d, FP64_PLAIN)
    b = Case(d, FP64_PLAIN)
    try:
        nb, nt = a.field_solver_shape()
        assert nt in (128, 256, 512) and 1 <= nb <= min(nt, (a.info().A_size + nt - 1) // nt)
        assert b.field_solver_shape() == (nb, nt)
        for c in (a, b):
            c.advance_time()
            c.field_solve()
            c.sync()
            it, outer, rel = c.last_field_solve()
            assert it > 0 and rel <= 1e-10
    finally:
        a.close()
        b.close()
    monkeypatch.setenv("PIC_PCG_BLOCKS", "7")
    f = Case(d, FP64_PLAIN)
    try:
        assert f.field_solver_shape() == (7, 256)
    finally:
        f.close()
    s = Case(d, FP32_REFINED)   # separable operator: no PCG
    try:
        assert s.field_solver_kind() == 1 and s.field_solver_shape() == (0, 0)
    finally:
        s.close()
