"""O2 and Ar/O2 cross-section tables and null-collision constants (host setup).

The tables are pinned bit for bit (FP32) against fixtures written from
oracle/_ref, where the reference's own src/PhysicsData.c fit functions,
CrossSection() data and FUNC_decay are compiled and driven with the table
loops of Oxygen_CrossSectionSET / ArO2_CrossSectionSET.  The row, rate and
pair specification is checked against the reference source text where that
tree is present.  The collision steps that use them are tested on the GPU
in test_gpu_o2.py and test_gpu_aro2.py.
"""
import base64
import json
import os

import numpy as np
import pytest

import picmcc
from picmcc import FP32_REFINED, FP64_PLAIN, HostSetup, PicError, decks

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"
AMU, QE = 1.66053e-27, 1.602e-19


def _fixture(name):
    with open(os.path.join(HERE, "golden", name)) as f:
        g = json.load(f)
    tab = np.frombuffer(base64.b64decode(g["table_f32_b64"]), dtype="<f4").reshape(g["rows"], g["cols"])
    th = np.frombuffer(base64.b64decode(g["th_e_f32_b64"]), dtype="<f4")
    return tab, th


def _setup(gas, prec=FP32_REFINED, **kw):
    base = dict(gas=gas, ncx=8, ncy=6, ppc=5)
    base.update(kw)
    return HostSetup(decks.oxygen_deck(**base), prec, load=False)


@pytest.mark.parametrize("gas,name,ncol,nrct", [(1, "o2_xsec_table.json", 58, 67),
                                                (2, "aro2_xsec_table.json", 68, 88)])
def test_tables_match_reference_build(gas, name, ncol, nrct):
    tab, th = _fixture(name)
    s = _setup(gas)
    assert int(s.scalar("TnRct")) == nrct
    cx = s.array("cx").reshape(81, nrct)
    assert np.array_equal(cx[:, :ncol].view(np.uint32), tab.view(np.uint32))
    assert not np.any(cx[:, ncol:]), "rate-constant reactions have no table column"
    assert np.array_equal(s.array("Th_e")[:ncol].view(np.uint32), th.view(np.uint32))
    # a plausible slip (one reaction on the wrong fit or threshold) moves a column
    assert not np.array_equal(cx[:, 10], tab[:, 11])


def test_fp64_tables_track_fp32():
    for gas, name in ((1, "o2_xsec_table.json"), (2, "aro2_xsec_table.json")):
        tab, _ = _fixture(name)
        s = _setup(gas, FP64_PLAIN)
        cx = s.array("cx").reshape(81, -1)[:, :tab.shape[1]]
        # same forms without the float roundings.  The polynomial fits in E
        # keep E^k in float in FP32; where their terms cancel (vibrational
        # and rotational fits near 1.6 and 16 eV) that moves the FP32 value
        # by up to ~1e-4 relative.  Entries next to a threshold may switch on
        # or off differently, so compare where both are nonzero.
        both = (cx != 0) & (tab != 0)
        rel = np.abs(cx[both] - tab[both]) / np.abs(tab[both])
        assert np.max(rel) < 2e-4
        assert np.median(rel) < 1e-7
        assert np.count_nonzero((cx != 0) != (tab != 0)) <= 2


def test_rate_constants_and_mass_ratios():
    s1, s2 = _setup(1), _setup(2)
    rr1, rr2 = s1.array("RR"), s2.array("RR")
    assert rr1[58] == np.float32(6.90e-16) and rr1[66] == np.float32(2.20e-24)
    assert rr2[68] == np.float32(2.00e-13) and rr2[87] == np.float32(1.12e-15)
    assert not np.any(rr1[:58]) and not np.any(rr2[:68])
    m1, m2 = s1.array("mofM"), s2.array("mofM")
    me = np.float32(5.485799486e-4 * AMU)
    f32 = np.float32
    # electron on O2, O- on O2 (16/32), O2(a) + O2(a) (1), O2+ on Ar (32/39.95)
    assert m1[0] == me / f32(32 * AMU)
    assert m1[41] == np.float32(0.5) and m1[65] == np.float32(1.0)
    assert m2[58] == f32(32 * AMU) / f32(39.95 * AMU)
    assert m2[65] == f32(39.95 * AMU) / f32(39.95 * AMU)


def test_null_collision_constants():
    """Each pair's sigma*v maximum bounds its own table sum at every grid
    energy and is reached on the grid; the electron collision probability
    sets the MCC sub-cycle count."""
    s = _setup(2, volts=150.0, pres_torr=0.03)
    sigv = s.array("sigv")
    assert len(sigv) == 25
    T = int(s.scalar("TnRct"))
    cx = s.array("cx").reshape(81, T).astype(np.float64)
    E = s.array("xee").astype(np.float64)
    v_e = np.sqrt(2 * QE * E / np.float32(5.485799486e-4 * AMU))
    # e + Ar uses columns 0-3, e + O2 columns 5-18 (the pair list is checked
    # against the reference text below)
    for k, cols in ((0, [0, 1, 2, 3]), (2, list(range(5, 19)))):
        sv = v_e * cx[:, cols].sum(axis=1)
        assert np.all(sv <= sigv[k] * (1 + 1e-6))
        assert np.max(sv) == pytest.approx(float(sigv[k]), rel=1e-6)
    cp = s.array("colprob")
    assert len(cp) == 5 and np.all(cp > 0) and np.all(cp < 1)
    mccn = int(s.scalar("MCCn"))
    assert mccn == (max(1, int(np.ceil(np.float32(cp[0]) / np.float32(0.08)))) if cp[0] > 0.08 else 1)
    assert s.scalar("dt_mcc") == pytest.approx(s.scalar("dt") / mccn, rel=1e-6)


def test_species_and_backgrounds():
    s = _setup(2, pres_torr=0.03, o2_ratio=0.2)
    assert int(s.scalar("nsp")) == 5 and int(s.scalar("nbg")) == 2
    n_ar, n_o2 = s.scalar("bg_dens0"), s.scalar("bg_dens1")
    assert n_o2 / n_ar == pytest.approx(0.2 / 0.8, rel=1e-6)
    s1 = _setup(1)
    assert int(s1.scalar("nsp")) == 4 and int(s1.scalar("nbg")) == 1
    # the loader is species-generic: four and five species load
    full = HostSetup(decks.oxygen_deck(gas=1, ncx=8, ncy=6, ppc=5), FP32_REFINED, load=True)
    assert [int(full.scalar("np%d" % i)) for i in range(4)] == [240, 240, 48, 48]


def test_reaction_names_follow_the_deck_format():
    obj, keys = decks.reaction_keys(1)
    assert obj == "OxygenCase" and len(keys) == 67 and keys[0] == "0.e+O2>e+O2"
    obj, keys = decks.reaction_keys(2)
    assert obj == "Argon/OxygenCase" and len(keys) == 88 and keys[87] == "87.Ar*+O2>O2+Ar"
    obj, keys = decks.reaction_keys(0)
    assert obj == "ArgonCase" and keys == list(decks.AR_COLLISIONS)
    # a switch set to 0 removes its column
    d = decks.oxygen_deck(gas=1, ncx=8, ncy=6, ppc=5)
    d["Collision_On_Off"]["OxygenCase"]["10.e+O2>2e+O2^"] = 0.0
    cx = HostSetup(d, FP32_REFINED, load=False).array("cx").reshape(81, 67)
    assert not np.any(cx[:, 10]) and np.any(cx[:, 9])


@pytest.mark.parametrize("prec", [FP32_REFINED, FP64_PLAIN])
def test_ar_o2_second_background_gas(prec):
    """Calculate_2GasPara (src/cuda_Field.cu:670-678): the O2 background of an
    Ar/O2 deck has the O2 share of the pressure at the same temperature, so
    BackDen2 / BackDen1 = ratio / (1 - ratio) and BackVel2 / BackVel1 =
    sqrt(m_Ar / m_O2) at every gas node, and both are 0 elsewhere."""
    hs = HostSetup(decks.oxygen_deck(gas=2, o2_ratio=0.25, ncx=8, ncy=6, ppc=5), prec, load=False)
    d1, d2 = hs.array("BackDen1"), hs.array("BackDen2")
    v1, v2 = hs.array("BackVel1"), hs.array("BackVel2")
    gas = hs.array("DensRegion") != 0
    assert gas.any() and not np.any(d2[~gas]) and not np.any(v2[~gas])
    tol = 1e-6 if prec == FP32_REFINED else 1e-13
    np.testing.assert_allclose(d2[gas] / d1[gas], 0.25 / 0.75, rtol=tol)
    np.testing.assert_allclose(v2[gas] / v1[gas], np.sqrt(39.950 / 32.0), rtol=tol)
    # the argon-only and O2-only decks have no second gas
    assert HostSetup(decks.oxygen_deck(gas=1, ncx=8, ncy=6, ppc=5), prec, load=False).array("BackDen2").size == 0


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "src")), reason="reference tree not present")
def test_spec_matches_reference_text():
    import importlib.util
    spec = importlib.util.spec_from_file_location("check_xsec_spec", os.path.join(HERE, "golden", "check_xsec_spec.py"))
    chk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(chk)
    assert chk.main(REF) == []
    # planted slips in the specification text are found
    with open(chk.HDR) as f:
        hdr = f.read()
    with open(os.path.join(REF, "src", "PhysicsData.c")) as f:
        phys = f.read()
    with open(os.path.join(REF, "src", "cuda_mcc.cu")) as f:
        mcc = f.read()
    slip = hdr.replace("{12.06, 1, K_FIT, F_ION, 0}", "{12.06, 2, K_FIT, F_ION, 0}", 1)
    assert chk.check_rows(phys, slip, "Oxygen_CrossSectionSET", "O2_Data", "kO2Rows", 58)
    slip = hdr.replace("{4.2, 1, K_FIT_X35, F_DA, 0}", "{4.2, 1, K_FIT, F_DA, 0}", 1)
    assert chk.check_rows(phys, slip, "Oxygen_CrossSectionSET", "O2_Data", "kO2Rows", 58)
    slip = hdr.replace("{1, T_BG, 0, {48, 49, 50, -2}}", "{1, T_BG, 0, {48, 49, -1, 50, -2}}", 1)
    assert chk.check_pairs(mcc, slip, "MainGas == OXYGEN", "O2_Data", "kO2Pairs", 20)
