"""Ar/O2 chemistry (gas type 2) collision step: HIP path against the oracle.

The oracle's orc_mcc_o2 with gas type 2 (oracle/orc_step.c) restates
src/cuda_mcc.cu:21-71 and src/cuda_mccArO2.cu:51-1265 on the CPU.  Both
sides take the same setup arrays (the tables are pinned against the
reference's compiled fits in test_xsec_oxygen.py, the reaction table against
the reference text in test_aro2_oracle.py) and the same particle snapshot;
the rebuilt buckets, the new ids, the reaction counts and the O- removed by
electron detachment must agree bit for bit.  Velocities are spread over
0.01-200 eV (electrons) and 1 meV-50 eV (ions) and the particle densities
raised, so that every reaction group fires.  Whole staged steps keep the
particle balance of all five species exactly, and the 512x256 workload
(44.6 M particles) runs with the same invariants.
"""
import numpy as np
import pytest

import oracle
from picmcc import FP32_REFINED, FP64_PLAIN, Case, HostSetup, decks

pytestmark = pytest.mark.gpu

PRECS = [(FP32_REFINED, "f32"), (FP64_PLAIN, "f64")]
FIELDS = ("x", "y", "vx", "vy", "vz", "id")
NSP = 5
T = 88
EMAX = (200.0, 50.0, 50.0, 50.0, 20.0)
EMIN = (0.01, 1e-3, 1e-3, 1e-3, 1e-3)


def _deck(**kw):
    # fluid densities near the gas density, so that collisions with Ar*,
    # O2(a), O2(b), O(3P) and O(1D) are as frequent as with the background
    dense = {"Ar*": 2e21, "O2A": 2e21, "O2B": 2e21, "OP": 2e21, "OD": 2e21}
    base = dict(gas=2, o2_ratio=0.3, ncx=10, ncy=8, ppc=150, pres_torr=0.1, volts=100.0, neutral_densities=dense)
    base.update(kw)
    return decks.oxygen_deck(**base)


def _spread(c, hs, real, rng, attach=False):
    """Replaces every velocity by an isotropic one with a log-uniform energy
    (electrons at 4-10 eV, the dissociative attachment peak, if attach)."""
    for s in range(NSP):
        p = c.particles(s)
        n = len(p["x"])
        lo, hi = (4.0, 10.0) if (attach and s == 0) else (EMIN[s], EMAX[s])
        e = np.exp(rng.uniform(np.log(lo), np.log(hi), n))
        sp = np.sqrt(e / hs.scalar("Escale%d" % s))
        mu = rng.uniform(-1, 1, n)
        ph = rng.uniform(0, 2 * np.pi, n)
        st = np.sqrt(1 - mu * mu)
        p["vx"] = (sp * st * np.cos(ph)).astype(real)
        p["vy"] = (sp * st * np.sin(ph)).astype(real)
        p["vz"] = (sp * mu).astype(real)
        c.set_particles(s, p, int(c.array("next_id")[s]))


def _consts(hs, deck, c):
    k = {f: hs.scalar(f) for f in ("TnRct", "nvel", "MCCn", "idLOGX", "dt", "idx", "idy")}
    k["gas"] = 2
    k["dtm"] = hs.scalar("dt_mcc")
    k["G"] = c.info().ngx * c.info().ngy
    k["nlogx"] = 81
    k["np2c"] = [hs.scalar("np2c%d" % s) for s in range(NSP)]
    k["Escale"] = [hs.scalar("Escale%d" % s) for s in range(NSP)]
    k["sigv"] = hs.array("sigv")
    for f in ("Th_e", "xe", "cx", "vsave", "BackDen1", "BackVel1", "BackDen2", "BackVel2", "FluidN"):
        k[f] = hs.array(f)
    k["seed"] = int(deck["SimulationMethod"]["PhiloxSeed"])
    return k


def _check_mcc(c, hs, deck, prec_name, den_scale):
    info = c.info()
    den = c.array("den")
    c.set_array("den", (den * den_scale).astype(den.dtype))
    k = _consts(hs, deck, c)
    k["den"] = c.array("den")
    k["step"] = info.tstep
    snap = [dict(c.particles(s), cap=info.cap[s]) for s in range(NSP)]
    mccr0, nid0 = c.array("mccr"), c.array("next_id")[:NSP]
    dc0 = c.array("detach_removed").astype(np.int64)
    c.mcc()
    c.sync()
    got = [c.particles(s) for s in range(NSP)]
    want, nid, mccr, rc, dc = oracle.mcc_o2(prec_name, k, snap, nid0, mccr0.astype(np.int64), delcnt=dc0)
    assert rc == 0
    for s in range(NSP):
        assert np.array_equal(got[s]["cell_start"], want[s]["cell_start"]), s
        for f in FIELDS:
            a, b = got[s][f], want[s][f]
            assert a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8)), (s, f)
    assert np.array_equal(c.array("next_id")[:NSP], nid)
    assert np.array_equal(c.array("mccr").astype(np.int64), mccr)
    assert np.array_equal(c.array("detach_removed").astype(np.int64), dc)
    d = mccr - mccr0.astype(np.int64)
    return [len(p["x"]) for p in snap], [len(p["x"]) for p in got], d.reshape(-1, int(k["TnRct"]))


@pytest.mark.parametrize("prec,name", PRECS)
def test_aro2_collision_step_matches_oracle(prec, name):
    deck = _deck()
    hs = HostSetup(deck, prec, load=False)
    c = Case(deck, prec)
    real = c.real
    rng = np.random.default_rng(11)
    fired = np.zeros(T, dtype=np.int64)
    try:
        for it in range(12):
            print("Ar/O2 collision step", it, name, flush=True)
            c.set_step(21 + it)
            _spread(c, hs, real, rng, attach=it % 3 == 2)
            n0, n1, d = _check_mcc(c, hs, deck, name, den_scale=6.0)
            fired += d.sum(axis=0)
            print("  counts", n0, "->", n1, flush=True)
    finally:
        c.close()
        hs.close()
    print("events per reaction", fired.tolist(), flush=True)
    assert fired[68:].sum() == 0   # rate-constant reactions are not collisions
    kinds = {"elastic": [0, 5], "loss": [1, 2, 6, 7, 8, 9, 10, 12, 13, 14, 23, 24, 25, 31, 32, 33, 37, 38],
             "superelastic": [21, 22, 29, 30, 45], "ionisation Ar": [3, 4], "ionisation O2+": [15, 19, 27],
             "ionisation O+": [18, 26, 34, 42, 44], "dissociative ionisation O+ + O-": [17],
             "attachment": [11, 20, 28], "detachment e+O-": [35], "recombination e": [36],
             "Ar+ charge exchange": [65], "Ar+ scattering": [66, 67], "O2+ to O+": [52, 55],
             "O2+ kept": [53, 56, 57, 59], "O+ to O2+": [60, 63, 64], "O+ kept": [62],
             "O- detachment": [47, 48, 51], "O- recombination": [49, 50]}
    for kind, rs in kinds.items():
        assert fired[rs].sum() > 0, kind


def _keV_o2_plus(c, escale, rng, per_cell):
    """A synthetic O2+ bucket: per_cell ions in every cell, energies
    log-uniform over 1-10 keV, isotropic, ids 0..n-1."""
    inf = c.info()
    ngy = inf.ngy
    counts = np.zeros(inf.ngx * ngy, dtype=np.int64)
    for i in range(inf.ngx - 1):
        counts[i * ngy: i * ngy + ngy - 1] = per_cell
    cs = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    n = int(cs[-1])
    e = np.exp(rng.uniform(np.log(1e3), np.log(1e4), n))
    sp = np.sqrt(e / escale)
    mu = rng.uniform(-1, 1, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    st = np.sqrt(1 - mu * mu)
    real = c.real
    p = dict(x=rng.random(n).astype(real), y=rng.random(n).astype(real), vx=(sp * st * np.cos(ph)).astype(real),
             vy=(sp * st * np.sin(ph)).astype(real), vz=(sp * mu).astype(real), id=np.arange(n, dtype=np.uint32),
             cell_start=cs)
    c.set_particles(2, p, n)


@pytest.mark.parametrize("prec,name", PRECS)
def test_aro2_o2_ion_charge_transfer_to_argon(prec, name):
    """O2+ + Ar -> Ar+ + O2 (reaction 58) has a thousandth of the cross
    section of its partner 59 at every energy, so the mixed test above sees
    it about once in ten runs.  Here 19 200 O2+ of 1-10 keV (where the
    collision rate is near its null-collision maximum) cross 1 Torr of 95 %
    argon with dt = T_RF/40 for three steps: reaction 58 fires (14 times in
    FP32 for this seed) and every bucket, id and count still equals the
    oracle's."""
    deck = _deck(o2_ratio=0.05, pres_torr=1.0, steps_per_cycle=40, ncx=16, ncy=12, ppc=100)
    for sp in ("O+", "Ar+"):   # room for the charge-exchange products of three steps
        deck["GasSpecies"]["ChargeSpecies"][sp]["Max_np"] = 60000
    deck["SimulationMethod"]["ParticleLimit"] = 60000
    hs = HostSetup(deck, prec, load=False)
    c = Case(deck, prec)
    rng = np.random.default_rng(58)
    fired = np.zeros(T, dtype=np.int64)
    try:
        for it in range(3):
            c.set_step(100 + it)
            _keV_o2_plus(c, hs.scalar("Escale2"), rng, 100)
            _, _, d = _check_mcc(c, hs, deck, name, den_scale=1.0)
            fired += d.sum(axis=0)
    finally:
        c.close()
        hs.close()
    print("O2+ events 52..59", fired[52:60].tolist(), flush=True)
    assert fired[58] > 0 and fired[59] > 1000


@pytest.mark.parametrize("prec,name", PRECS)
def test_aro2_staged_steps_keep_the_bookkeeping(prec, name):
    """Whole Ar/O2 steps (pic_step runs the stages one by one for gas type
    2): the buckets stay sorted and consistent, the ids unique, and the step
    after the field solve and push still matches the oracle's collision
    step."""
    deck = _deck(ncx=12, ncy=10, ppc=30)
    hs = HostSetup(deck, prec, load=False)
    c = Case(deck, prec)
    try:
        c.step(4)
        c.sync()
        for s in range(NSP):
            p = c.particles(s)
            cs = p["cell_start"]
            assert cs[0] == 0 and cs[-1] == len(p["x"]) and np.all(np.diff(cs) >= 0)
            assert len(np.unique(p["id"])) == len(p["id"])
            assert np.all((p["x"] >= 0) & (p["x"] < 1) & (p["y"] >= 0) & (p["y"] < 1))
        c.advance_time(); c.field_solve(); c.efield(); c.move(); c.sort_boundary()
        _check_mcc(c, hs, deck, name, den_scale=1.0)
    finally:
        c.close()
        hs.close()


def _stoich():
    """net change of (e, Ar+, O2+, O+, O-) per event of each reaction; the O-
    also lose the electron-detachment removals, counted apart"""
    d = {}
    for r in (3, 4):
        d[r] = (1, 1, 0, 0, 0)
    for r in (15, 19, 27):
        d[r] = (1, 0, 1, 0, 0)
    for r in (18, 26, 34, 42, 44):
        d[r] = (1, 0, 0, 1, 0)
    d[17] = (0, 0, 0, 1, 1)
    for r in (11, 20, 28):
        d[r] = (-1, 0, 0, 0, 1)
    d[35] = (1, 0, 0, 0, 0)
    d[36] = (-1, 0, 0, 0, 0)
    for r in (52, 55):
        d[r] = (0, 0, -1, 1, 0)
    d[58] = (0, 1, -1, 0, 0)
    for r in (60, 63, 64):
        d[r] = (0, 0, 1, -1, 0)
    for r in (47, 48, 51):
        d[r] = (1, 0, 0, 0, -1)
    for r in (49, 50):
        d[r] = (0, 0, 0, 0, -1)
    return d


def _balance(c, steps):
    nsp = c.info().nsp
    ncond = c.info().CondNUMR

    def state():
        inf = c.info()
        cc = c.array("CondCharge").astype(np.int64).reshape(nsp, ncond).sum(axis=1)
        return (np.array([inf.np[s] for s in range(nsp)], dtype=np.int64),
                c.array("mccr").astype(np.int64).reshape(-1, T).sum(axis=0),
                c.array("Lost").astype(np.int64) + cc, int(c.array("detach_removed").astype(np.int64).sum()))
    n0, m0, w0, r0 = state()
    c.step(steps)
    c.sync()
    n1, m1, w1, r1 = state()
    dm = m1 - m0
    assert dm.sum() > 0
    made = np.zeros(nsp, dtype=np.int64)
    for r, v in _stoich().items():
        made += dm[r] * np.array(v, dtype=np.int64)
    made[4] -= r1 - r0
    assert np.array_equal(n1 - n0, made - (w1 - w0)), (n1 - n0, made, w1 - w0, r1 - r0)
    return n1, dm


@pytest.mark.parametrize("prec", [FP32_REFINED, FP64_PLAIN])
def test_aro2_particle_balance_closes(prec):
    """Whole Ar/O2 steps at 96x96 cells: the change of every species' count
    equals what the counted reactions make and remove, less the electron
    detachment removals of O- and the particles taken by the walls and
    electrodes."""
    c = Case(_deck(ncx=96, ncy=96, ppc=40, pres_torr=0.05), prec)
    try:
        _balance(c, 6)
    finally:
        c.close()


def test_aro2_512x256_workload_invariants():
    """The 512x256 Ar/O2 workload (O2 20 %, 30 mTorr, 150 V, 100 per cell for
    e, Ar+ and O2+ and 20 for O+ and O-: 44.6 M particles), FP32 with the
    FP64-refined field solve: a few whole steps keep the balance of every
    species, sorted buckets, offsets in [0, 1) and unique ids."""
    deck = decks.config3()
    c = Case(deck, FP32_REFINED)
    try:
        inf = c.info()
        print("512x256 Ar/O2 particles", list(inf.np), flush=True)
        assert sum(inf.np) > 44_000_000
        n1, dm = _balance(c, 5)
        print("after 5 steps", n1.tolist(), "events", int(dm.sum()), flush=True)
        for s in range(NSP):
            p = c.particles(s)
            cs = p["cell_start"]
            assert cs[0] == 0 and cs[-1] == len(p["x"]) and np.all(np.diff(cs) >= 0)
            assert np.all((p["x"] >= 0) & (p["x"] < 1) & (p["y"] >= 0) & (p["y"] < 1))
            assert len(np.unique(p["id"])) == len(p["id"])
        it, outer, rel = c.last_field_solve()
        assert rel < 1e-9
    finally:
        c.close()
