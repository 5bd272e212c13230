"""Checkpoint and resume in the reference's V000 dump format
(src/cuda_SaveDump.cu:5-155, src/load.c:5-120), both precisions.

- pic_dump_save: the file parses with the layout checker to the last byte;
  header, np2c and counts match the case; every particle record is the cell's
  node index, corner + in-cell offset rounded as the reference rounds it, and
  the velocity as float; per-node densities, surface charge, potentials and E
  are the case's arrays as float; the history section holds the diagnostics
  records.
- pic_dump_load into a fresh case of the same deck: the particles are the
  file's, bucketed as the reference buckets them (checker's bucket()), ids
  0..n-1 per species; the surface charge, potentials, E and step are the
  file's; the deposit that starts a run gives the density of the loaded
  particles.
- two cases resumed from one file step identically (bit for bit).
- the error paths: another gas, another np2c, a truncated file, a particle
  outside the grid, and a far-wall position kept in the last cell.
"""
import os

import numpy as np
import pytest

import dump_v000 as V
from picmcc import Case, FP32_REFINED, FP64_PLAIN, PicError, decks

pytestmark = pytest.mark.gpu

PRECS = [(FP32_REFINED, np.float32), (FP64_PLAIN, np.float64)]
IDS = ["fp32", "fp64"]


def dims(c, gas=0):
    inf = c.info()
    return V.Dims(inf.nsp, V.NFSP[gas], inf.ngx - 1, inf.ngy - 1, inf.CondNUMR, inf.TnRct)


def deck():
    return decks.small(ncx=20, ncy=14, ppc=30)


@pytest.mark.parametrize("prec,dt", PRECS, ids=IDS)
def test_save_writes_the_case_state(prec, dt, tmp_path):
    c = Case(deck(), prec)
    c.step(4)
    c.diag_sample()
    c.step(3)
    c.diag_sample()
    path = str(tmp_path / "a.dmp")
    c.dump_save(path)
    D = dims(c)
    f = V.read(path, D)
    inf = c.info()
    assert f["trailing_bytes"] == 0 and f["key"] == [0, 0, 0]
    assert f["tstep"] == inf.tstep == 7 and f["t"] == inf.t and f["gas"] == 0
    ngy = inf.ngy
    for s in range(inf.nsp):
        p = c.particles(s)
        r = f["particles"][s]
        assert len(r["x"]) == inf.np[s]
        cs = p["cell_start"]
        node = np.repeat(np.arange(len(cs) - 1), np.diff(cs))
        assert np.array_equal(r["cell"], node.astype(np.int32))
        xi, yi = (node // ngy), (node % ngy)
        if dt == np.float32:
            xa = xi.astype(np.float32) + p["x"]
            ya = yi.astype(np.float32) + p["y"]
        else:
            xa = (xi + p["x"]).astype(np.float32)
            ya = (yi + p["y"]).astype(np.float32)
        assert np.array_equal(r["x"], xa) and np.array_equal(r["y"], ya)
        for k in ("vx", "vy", "vz"):
            assert np.array_equal(r[k], p[k].astype(np.float32))
    G = inf.ngx * inf.ngy
    assert np.array_equal(f["den"], c.array("den").reshape(inf.nsp, G).astype(np.float32))
    assert np.array_equal(f["sigma"], c.array("sigma_sp").reshape(inf.nsp, G).astype(np.float32))
    phi, tot = c.array("phi"), c.array("TotPot")
    assert np.array_equal(f["Pois_Pot"], phi.astype(np.float32))
    assert np.array_equal(f["TotPotential"], tot.astype(np.float32))
    assert np.array_equal(f["Lap_Pot"], (tot - phi).astype(np.float32))
    assert np.array_equal(f["Ex"], c.array("Ex").astype(np.float32))
    assert np.array_equal(f["Ey"], c.array("Ey").astype(np.float32))
    # history: the two diagnostics records
    h = c.diag_history()
    assert f["hist_count"] == 2 and f["dhist"] == 3
    assert np.array_equal(f["hist"]["t"], h["t"].astype(np.float32))
    assert np.array_equal(f["hist"]["np"], h["np"].T.astype(np.float32))
    assert np.array_equal(f["hist"]["surf_charge"], h["q_cond"].T.astype(np.float32))
    assert np.array_equal(f["hist"]["volt"], h["v_cond"].T.astype(np.float32))
    c.close()


def load_checks(path, prec, dt, D):
    f = V.read(path, D)
    b = Case(deck(), prec)
    b.dump_load(path)
    inf = b.info()
    assert inf.tstep == f["tstep"]
    nid = b.array("next_id")
    for s in range(D.nsp):
        r = f["particles"][s]
        order, cs, ox, oy = V.bucket(r["x"], r["y"], D.ncx, D.ncy, dt)
        p = b.particles(s)
        assert np.array_equal(p["cell_start"], cs)
        assert np.array_equal(p["x"], ox) and np.array_equal(p["y"], oy)
        for k in ("vx", "vy", "vz"):
            assert np.array_equal(p[k], r[k][order].astype(dt))
        assert np.array_equal(p["id"], np.arange(len(order), dtype=np.uint32))
        assert int(nid[s]) == len(order)
    G = D.G
    assert np.array_equal(b.array("sigma_sp").reshape(D.nsp, G), f["sigma"].astype(dt))
    assert np.array_equal(b.array("phi"), f["Pois_Pot"].astype(dt))
    assert np.array_equal(b.array("TotPot"), f["TotPotential"].astype(dt))
    assert np.array_equal(b.array("Ex"), f["Ex"].astype(dt))
    return b, f


@pytest.mark.parametrize("prec,dt", PRECS, ids=IDS)
def test_load_restores_the_file_and_resumes_deterministically(prec, dt, tmp_path):
    a = Case(deck(), prec)
    a.step(6)
    path = str(tmp_path / "a.dmp")
    a.dump_save(path)
    D = dims(a)
    den_a = a.array("den")
    a.close()
    b, f = load_checks(path, prec, dt, D)
    # the loaded particles sit within float rounding of the saved ones, so
    # the deposit of the loaded state is the saved density within that
    den_b = b.array("den")
    scale = np.max(np.abs(den_a))
    assert np.max(np.abs(den_b - den_a)) <= 1e-5 * scale
    # resumed twice from one file: identical runs
    b.step(4)
    b.sync()
    c = Case(deck(), prec)
    c.dump_load(path)
    c.step(4)
    c.sync()
    for s in range(D.nsp):
        pb, pc = b.particles(s), c.particles(s)
        for k in ("x", "y", "vx", "vy", "vz", "id", "cell_start"):
            assert np.array_equal(pb[k], pc[k]), f"species {s} {k}"
    assert b.info().tstep == f["tstep"] + 4
    b.close()
    c.close()


def test_load_rejects_what_does_not_fit_and_keeps_far_wall_positions(tmp_path):
    a = Case(deck(), FP32_REFINED)
    a.step(2)
    path = str(tmp_path / "a.dmp")
    a.dump_save(path)
    D = dims(a)
    a.close()
    f = V.read(path, D)
    # a particle on the far wall in x and one at the far corner: kept in the last cell
    p = f["particles"][0]
    p["x"][0] = np.float32(D.ncx)
    p["x"][1] = np.float32(D.ncx)
    p["y"][1] = np.float32(D.ncy)
    edge = str(tmp_path / "edge.dmp")
    with open(edge, "wb") as fh:
        fh.write(V.write(f, D))
    b, _ = load_checks(edge, FP32_REFINED, np.float32, D)
    b.step(1)
    b.sync()
    b.close()
    # the errors the loader reports
    cases = {}
    g = dict(f)
    g["gas"] = 1
    cases["gas"] = (V.write(g, D), "gas type")
    g = dict(f)
    g["np2c"] = [np.float32(x * 2) for x in f["np2c"]]
    cases["np2c"] = (V.write(g, D), "np2c")
    whole = V.write(f, D)
    cases["short"] = (whole[:len(whole) // 2], "ends early")
    cases["short_tail"] = (whole[:-4], "ends early")
    q = V.read(path, D)
    q["particles"][1]["y"][0] = np.float32(-0.5)
    cases["outside"] = (V.write(q, D), "outside the grid")
    for name, (data, msg) in cases.items():
        bad = str(tmp_path / f"{name}.dmp")
        with open(bad, "wb") as fh:
            fh.write(data)
        c = Case(deck(), FP32_REFINED)
        with pytest.raises(PicError, match=msg):
            c.dump_load(bad)
        c.close()
    with pytest.raises(PicError, match="cannot open"):
        c = Case(deck(), FP32_REFINED)
        try:
            c.dump_load(os.path.join(str(tmp_path), "missing.dmp"))
        finally:
            c.close()
