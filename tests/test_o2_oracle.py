"""Pins of the oracle's O2 collision step (oracle/orc_step.c, orc_mcc_o2).

The reaction table (hybrid-final_amd/csrc/o2_chem.h) is checked against the
reference's handler text (src/cuda_mccO2.cu) where that tree is present,
with planted slips.  The oracle's use of it is pinned on synthetic inputs
where one reaction has all the cross section and the collision probability
is 1, so the outcome of every particle has a closed form: the energy after a
loss, the products of an ionisation and their velocities, a charge exchange
into another species, and the O- removed by electron detachment (checked
against Philox keys computed here through the oracle's own Philox export,
which test_oracle_pins.py pins against the Random123 known answers).
"""
import ctypes as C
import importlib.util
import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"
T = 67
ESC = (2.843e-12, 1.0e-7, 5.0e-8, 5.0e-8)   # eV per (m/s)^2, synthetic


def _k(G=1, rows=(), sig=None, bg=0.0, fluid=(0, 0, 0, 0), den=(0, 0, 0, 0), th=None):
    cx = np.zeros((81, T))
    for r, v in rows:
        cx[:, r] = v
    k = dict(G=G, TnRct=T, nlogx=81, nvel=16, MCCn=1, idLOGX=10.0, dt=1e-9, dtm=1e-9, idx=1e3, idy=1e3,
             np2c=[1.0] * 4, Escale=list(ESC), sigv=np.zeros(20) if sig is None else sig,
             Th_e=np.zeros(T) if th is None else th, xe=np.arange(81) * 0.1 - 4.0, cx=cx.ravel(),
             vsave=np.linspace(0.5, 2.0, 16), BackDen1=np.full(G, bg), BackVel1=np.full(G, 300.0),
             FluidN=np.repeat(np.asarray(fluid, dtype=float), G), den=np.repeat(np.asarray(den, dtype=float), G),
             seed=12345, step=3, ions=1)
    return k


def _bucket(n, esc, e_lo, e_hi, rng, cap=None, id0=0):
    e = rng.uniform(e_lo, e_hi, n)
    sp = np.sqrt(e / esc)
    mu = rng.uniform(-1, 1, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    st = np.sqrt(1 - mu * mu)
    return dict(x=rng.uniform(0, 1, n), y=rng.uniform(0, 1, n), vx=sp * st * np.cos(ph), vy=sp * st * np.sin(ph),
                vz=sp * mu, id=np.arange(id0, id0 + n, dtype=np.uint32), cell_start=np.array([0, n], dtype=np.int32),
                cap=cap or 4 * n + 16), e


def _empty():
    return dict(x=[], y=[], vx=[], vy=[], vz=[], id=np.zeros(0, np.uint32), cell_start=np.array([0, 0], np.int32),
                cap=1024)


def _speed(b):
    return np.sqrt(np.asarray(b["vx"]) ** 2 + np.asarray(b["vy"]) ** 2 + np.asarray(b["vz"]) ** 2)


def _run(k, bs, prec="f64"):
    return oracle.mcc_o2(prec, k, bs, np.zeros(4, np.uint32) + 1000, np.zeros(T, np.int64))


def test_energy_loss_takes_the_threshold():
    rng = np.random.default_rng(1)
    th = np.zeros(T); th[1] = 2.5
    sig = np.zeros(20); sig[0] = 1e-30      # R * sigv / v is far below the cross section
    k = _k(rows=[(1, 1e-19)], sig=sig, bg=1e40, th=th)
    b, e = _bucket(200, ESC[0], 3.0, 40.0, rng)
    out, nid, mccr, rc = _run(k, [b, _empty(), _empty(), _empty()])
    assert rc == 0 and mccr[1] == 200 and mccr.sum() == 200
    np.testing.assert_allclose(ESC[0] * _speed(out[0]) ** 2, e - 2.5, rtol=1e-12)
    assert np.array_equal(out[0]["id"], b["id"])


def test_ionisation_products():
    """e + O2 -> 2e + O2+ (reaction 10): the new electron carries the unit
    direction (speed 1, as the reference writes it), the ion a Maxwellian
    speed vsave[n] * BackVel1, and the electron loses the threshold."""
    rng = np.random.default_rng(2)
    th = np.zeros(T); th[10] = 12.06
    sig = np.zeros(20); sig[0] = 1e-30
    k = _k(rows=[(10, 1e-19)], sig=sig, bg=1e40, th=th)
    b, e = _bucket(100, ESC[0], 15.0, 60.0, rng)
    out, nid, mccr, rc = _run(k, [b, _empty(), _empty(), _empty()])
    assert rc == 0 and mccr[10] == 100
    assert len(out[0]["x"]) == 200 and len(out[1]["x"]) == 100
    np.testing.assert_allclose(ESC[0] * _speed(out[0])[:100] ** 2, e - 12.06, rtol=1e-12)
    np.testing.assert_allclose(_speed(out[0])[100:], 1.0, rtol=1e-12)
    ion = _speed(out[1]) / 300.0
    assert np.all(np.min(np.abs(ion[:, None] - k["vsave"][None, :]), axis=1) < 1e-9)
    # products sit at their parent's position; new ids count up from next_id
    assert np.array_equal(out[0]["x"][100:], b["x"]) and np.array_equal(out[1]["y"], b["y"])
    assert list(out[1]["id"]) == list(range(1000, 1100)) and list(nid) == [1100, 1100, 1000, 1000]


def test_charge_exchange_changes_species():
    """O2+ + O(3P) -> O+ (reaction 47): every O2+ becomes an O+ with the
    neutral's velocity."""
    rng = np.random.default_rng(3)
    sig = np.zeros(20); sig[12] = 1e-30
    k = _k(rows=[(47, 1e-19)], sig=sig, fluid=(0, 0, 1e40, 0))
    b, _ = _bucket(80, ESC[1], 0.1, 5.0, rng)
    out, nid, mccr, rc = _run(k, [_empty(), b, _empty(), _empty()])
    assert rc == 0 and mccr[47] == 80
    assert len(out[1]["x"]) == 0 and len(out[2]["x"]) == 80
    sp = _speed(out[2]) / 300.0
    assert np.all(np.min(np.abs(sp[:, None] - k["vsave"][None, :]), axis=1) < 1e-9)


def _keys(ids, step, seed):
    lib = C.CDLL(os.path.join(os.path.dirname(oracle.__file__), "liborc_f64.so"))
    out = []
    for i in ids:
        ctr = (C.c_uint32 * 4)(int(i), step, (3 << 8) | 2, 0)
        key = (C.c_uint32 * 2)(seed & 0xFFFFFFFF, seed >> 32)
        o = (C.c_uint32 * 4)()
        lib.orc_philox(ctr, key, o)
        out.append(o[0])
    return np.array(out, dtype=np.uint64)


@pytest.mark.parametrize("n_e,n_neg", [(5, 12), (12, 5)])
def test_detachment_removes_the_smallest_key_o_minus(n_e, n_neg):
    """e + O- -> 2e + O(3P) (reaction 30): one electron more per event and,
    per event, one O- of the cell removed while any is left, smallest Philox
    key (seed, id, step) first."""
    rng = np.random.default_rng(4)
    th = np.zeros(T); th[30] = 1.0
    sig = np.zeros(20); sig[3] = 1e-30
    k = _k(rows=[(30, 1e-19)], sig=sig, den=(0, 0, 0, 1e40), th=th)
    be, _ = _bucket(n_e, ESC[0], 2.0, 10.0, rng)
    bn, _ = _bucket(n_neg, ESC[3], 0.1, 1.0, rng, id0=500)
    out, nid, mccr, rc = _run(k, [be, _empty(), _empty(), bn])
    assert rc == 0 and mccr[30] == n_e
    assert len(out[0]["x"]) == 2 * n_e
    keys = _keys(bn["id"], 3, 12345)
    keep = np.sort(bn["id"][np.argsort(keys, kind="stable")[min(n_e, n_neg):]])
    assert np.array_equal(np.sort(out[3]["id"]), keep)


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "src")), reason="reference tree not present")
def test_table_matches_reference_text():
    spec = importlib.util.spec_from_file_location("check_o2_chem", os.path.join(HERE, "golden", "check_o2_chem.py"))
    chk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(chk)
    assert chk.main(REF) == []
    with open(chk.HDR) as f:
        hdr = f.read()
    with open(os.path.join(REF, "src", "cuda_mccO2.cu")) as f:
        src = f.read()
    slips = [("X(12, 1, O2T_ION, O2E_LOSS, 0.0f, 2, 3)", "X(12, 1, O2T_ION, O2E_LOSS, 0.0f, 2, 1)"),
             ("X(49, 0, O2T_SCAT", "X(49, 1, O2T_SCAT"),
             ("X(16, 1, O2T_LOSS, O2E_GAIN, 0.977", "X(16, 1, O2T_LOSS, O2E_LOSS, 0.977"),
             ("X(53, 0, O2T_CXTO, O2E_NONE, 0.0f, 1, -1)", "X(53, 0, O2T_CXKEEP, O2E_NONE, 0.0f, -1, -1)"),
             ("G(3, 0, 9, 52, 2)", "G(3, 0, 7, 52, 2)"),
             ("T(0, 3, 3, 2, 3)", "T(0, 3, 3, 2, 1)"),
             ("T(2, 0, 16, 0, 0)", "T(2, 0, 16, 1, 0)")]
    for a, b in slips:
        assert a in hdr, a
        assert chk.check(src, hdr.replace(a, b, 1)), b
