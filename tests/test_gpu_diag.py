"""GPU diagnostics records (the reference's Diagnostic, src/cuda_Diagnostic.cu:37-267)
checked against the case's own arrays, read back and summed in numpy:

- particle counts equal the case's counts;
- kinetic energy per species equals m v^2/2 x np2c summed over the particles,
  with m from the physical constants and np2c from the deck;
- the charge of each conductor equals the sum of Sigma x Area over the nodes
  whose CondID names it (the reference's Surf_charge);
- the absorbed counts are the differences of the running CondCharge between
  consecutive records (the first record counts from the start of the case);
- the conductor voltages equal V_cond; t and step follow the step counter;
- the history keeps records oldest first and survives further steps.
"""
import numpy as np
import pytest

from picmcc import Case, FP32_REFINED, FP64_PLAIN, decks

pytestmark = pytest.mark.gpu

AMU = 1.66053e-27
MASS = (5.485799486e-4 * AMU, 39.95 * AMU)   # electron, Ar+


def find(d, key):
    if isinstance(d, dict):
        for k, v in d.items():
            if k == key:
                return v
            r = find(v, key)
            if r is not None:
                return r
    return None


def snapshot(c):
    inf = c.info()
    nsp, nc = inf.nsp, inf.CondNUMR
    sig = c.array("Sigma").astype(np.float64)
    area = c.array("Area").astype(np.float64)
    cid = c.array("CondID").astype(np.int64)
    q = np.array([np.sum(sig[cid == i + 1] * area[cid == i + 1]) for i in range(nc)])
    v2 = []
    for s in range(nsp):
        p = c.particles(s)
        vx, vy, vz = (p[k].astype(np.float64) for k in ("vx", "vy", "vz"))
        v2.append(float(np.sum(vx * vx + vy * vy + vz * vz)))
    return {"np": [int(inf.np[s]) for s in range(nsp)], "v2": v2, "q": q,
            "cc": c.array("CondCharge").reshape(nsp, nc).astype(np.int64),
            "v": c.array("V_cond")[:nc].astype(np.float64), "tstep": int(inf.tstep), "dt": float(inf.dt)}


def deck():
    """Thick electrodes (conductor cells, which absorb) and a dielectric block
    (surface charge on the conductors' neighbours): every record field moves."""
    from test_setup_parity import deck_variants
    d = deck_variants()["thick_electrodes_dielectric"]
    d["Geometry"]["Source"][0]["AC(V)"] = 150.0
    return d


@pytest.mark.parametrize("prec,rtol", [(FP32_REFINED, 2e-6), (FP64_PLAIN, 1e-11)], ids=["fp32", "fp64"])
def test_diagnostics_records_match_the_case_arrays(prec, rtol):
    d = deck()
    np2c = find(d, "np2c")
    c = Case(d, prec)
    snaps = []
    for k in (5, 9, 6):
        c.step(k)
        c.diag_sample()
        snaps.append(snapshot(c))
    h = c.diag_history()
    assert len(h["t"]) == 3
    cc_prev = np.zeros_like(snaps[0]["cc"])
    for r, sn in enumerate(snaps):
        assert int(h["step"][r]) == sn["tstep"]
        assert h["t"][r] == pytest.approx(sn["tstep"] * sn["dt"], rel=1e-12)
        assert list(h["np"][r]) == sn["np"], f"record {r}: counts"
        for s in range(2):
            want = 0.5 * MASS[s] * np2c * sn["v2"][s]
            assert h["ke"][r][s] == pytest.approx(want, rel=rtol), f"record {r} species {s}: kinetic energy"
        scale = max(np.max(np.abs(sn["q"])), 1e-300)
        assert np.all(np.abs(h["q_cond"][r] - sn["q"]) <= rtol * scale), f"record {r}: conductor charge"
        assert np.array_equal(h["absorbed"][r].astype(np.int64), sn["cc"] - cc_prev), f"record {r}: absorbed"
        cc_prev = sn["cc"]
        assert np.array_equal(h["v_cond"][r], sn["v"]), f"record {r}: voltages"
    assert h["absorbed"].sum() > 0, "no particle reached a conductor"
    assert np.any(h["q_cond"] != 0)
    # more steps: the history grows at its end and keeps the earlier records
    c.step(2)
    c.diag_sample()
    h2 = c.diag_history()
    assert len(h2["t"]) == 4
    for k in ("t", "ke", "q_cond", "v_cond"):
        assert np.array_equal(h2[k][:3], h[k])
    c.close()
