"""Host setup of the product (deck -> geometry, operator, Laplace and gas
fields, tables, initial load) equals the oracle's bit for bit (no GPU)."""
import copy

import numpy as np
import pytest

from oracle import Oracle
from picmcc import FP32_REFINED, FP64_PLAIN, HostSetup, PicError, decks

SETUP_ARRAYS = ["Boundary", "CondID", "Face", "DensRegion", "Area", "Temp", "PlasmaRegion", "eps_r", "eps",
                "StructIdx", "A_idx", "Ai", "Aj", "Aval", "Atemp", "Mdiag", "cond_b", "temp_b", "lap_phi",
                "BackDen1", "BackVel1", "FluidN", "xe", "xee", "cx", "sigv", "P1", "P2", "Pion", "vsave",
                "fl_x1", "fl_x2", "fl_xy", "fl_yx", "fl_y1", "fl_y2", "fl_xg1", "fl_xg2", "fl_yg1", "fl_yg2",
                "fl_D", "fl_den", "fl_ax", "fl_bx", "fl_ay", "fl_by", "fl_flux_x", "fl_flux_y"]


def deck_variants():
    base = decks.small(ncx=12, ncy=9, ppc=10)
    out = {"small": base}
    d = copy.deepcopy(base)   # thick electrodes (conductor cells) and a dielectric block
    g = d["Geometry"]
    g["ConductorSpec"][0].update({"Y0": 7, "Y1": 9})
    g["ConductorSpec"][1].update({"Y0": 0, "Y1": 1})
    g["DielectricSpec"] = [{"M_ID": 3, "X0": 9, "X1": 12, "Y0": 1, "Y1": 7, "Epsilon": 4.0}]
    out["thick_electrodes_dielectric"] = d
    d = copy.deepcopy(base)   # Dirichlet right wall, extra grounded conductor inside
    g = d["Geometry"]
    g["BoundaryCondition"][3]["B.C"] = 1
    g["ConductorSpec"].append({"M_ID": 3, "X0": 5, "X1": 7, "Y0": 3, "Y1": 5, "Temp": 350})
    out["inner_conductor"] = d
    d = copy.deepcopy(base)   # anisotropic cells, phase and DC on the source
    d["Geometry"]["SystemSpec"]["X_length(m)"] = 12 * 7e-4
    d["Geometry"]["Source"][0].update({"DC(V)": -20.0, "Phase(deg)": 30.0})
    out["anisotropic"] = d
    out["config1"] = decks.config1()
    # taller than wide: the reference's top-right structure-map corner reads
    # past its node array here (src/start.c:1848)
    out["tall"] = decks.small(ncx=6, ncy=20, ppc=4)
    return out


@pytest.mark.parametrize("ncx,ncy", [(6, 20), (12, 9), (10, 10)])
def test_structure_map_top_right_corner(ncx, ncy):
    """The padded map's top-right corner holds minus the boundary code of a
    top-wall node: node (ncy-1, ncy) as the reference reads it while that
    node exists (ncy <= ncx), node (ncx-1, ncy) otherwise."""
    d = decks.small(ncx=ncx, ncy=ncy, ppc=2)
    # the top wall is Dirichlet up to x = ncx - 3 and Neumann beyond, so
    # top-wall nodes differ and the corner shows which one was read
    bc = d["Geometry"]["BoundaryCondition"]
    bc[1]["X1"] = ncx - 3
    bc.append({"X0": ncx - 2, "X1": ncx, "Y0": ncy, "Y1": ncy, "B.C": 2, "Temp": 300.0})
    h = HostSetup(d, FP32_REFINED, load=False)
    ngy = ncy + 1
    B = h.array("Boundary")
    SI = h.array("StructIdx").reshape(ncx + 2, ncy + 2)
    node = (ncy - 1 if ncy <= ncx else ncx - 1) * ngy + ncy
    assert SI[ncx + 1, ncy + 1] == -B[node]
    if ncy < ncx:   # the two candidate nodes really differ here
        assert B[node] != B[(ncx - 1) * ngy + ncy]
    assert np.array_equal(SI, Oracle(d, "f32").arr("StructIdx").reshape(ncx + 2, ncy + 2))


@pytest.mark.parametrize("name", list(deck_variants()))
@pytest.mark.parametrize("prec,pn", [(FP32_REFINED, "f32"), (FP64_PLAIN, "f64")])
def test_setup_matches_oracle(name, prec, pn):
    d = deck_variants()[name]
    h = HostSetup(d, prec)
    o = Oracle(d, pn)
    for a in SETUP_ARRAYS:
        x, y = h.array(a), o.arr(a)
        assert x.shape == y.shape, a
        assert np.array_equal(x, y.astype(x.dtype)), f"{name}/{pn}: {a}"
    for s in range(2):
        for k in ("x", "y", "vx", "vy", "vz", "id", "cell_start"):
            x = h.array(f"sp{s}.{k}")
            y = o.arr(f"sp{s}.{k}")[:len(x)]
            assert np.array_equal(x, y), f"{name}/{pn}: sp{s}.{k}"
    for sc in ("dt", "dx", "dy", "MCCn", "dt_mcc", "nvel", "A_size", "CondNUMR", "dtc", "fl_nx", "fl_ny", "nfsp"):
        assert h.scalar(sc) == o.scalar(sc), sc


def test_missing_keys_read_as_zero():
    """parson semantics: a missing key reads as 0 (src/parson.c:1319-1321); the
    reference's own validation then rejects a zero TimeStep_PIC."""
    d = decks.small()
    del d["SimulationMethod"]["TimeStep_PIC"]
    with pytest.raises(PicError) as e:
        HostSetup(d)
    assert "TimeStep_PIC" in str(e.value)


@pytest.mark.parametrize("mutate,word", [
    (lambda d: d["Geometry"].__setitem__("BoundaryCondition", d["Geometry"]["BoundaryCondition"][:3]), "too small"),
    (lambda d: d["Geometry"]["ConductorSpec"][0].__setitem__("M_ID", 2), "M_ID"),
    (lambda d: d["Geometry"]["Source"][0].__setitem__("M_ID", 9), "Source"),
    (lambda d: d["GasSpecies"].__setitem__("Type(0:Ar,1:O2,2:Ar/O2)", 7), "Type"),
])
def test_deck_validation_errors(mutate, word):
    d = decks.small()
    mutate(d)
    with pytest.raises(PicError) as e:
        HostSetup(d)
    assert word in str(e.value)


def test_malformed_json_is_an_error():
    with pytest.raises(PicError):
        HostSetup('{"Geometry": {')


OXY_FLUID_ARRAYS = ["Temp", "BackDen1", "BackVel1", "FluidN", "fl_x1", "fl_x2", "fl_xy", "fl_yx", "fl_y1", "fl_y2",
                    "fl_xg1", "fl_xg2", "fl_yg1", "fl_yg2", "fl_D", "fl_den", "fl_ax", "fl_bx", "fl_ay", "fl_by",
                    "fl_flux_x", "fl_flux_y"]


@pytest.mark.parametrize("gas", [1, 2], ids=["O2", "ArO2"])
@pytest.mark.parametrize("prec,pn", [(FP32_REFINED, "f32"), (FP64_PLAIN, "f64")])
def test_oxygen_fluid_setup_matches_oracle(gas, prec, pn):
    """O2 and Ar/O2 cases: backgrounds (Ar/O2 with its two-gas split), the
    four or five fluid species' lines, wall factors, diffusion and
    Scharfetter-Gummel coefficients, initial densities and fluxes, and the
    rate constants of the reactions after the tabulated ones."""
    d = decks.oxygen_deck(gas=gas, ncx=10, ncy=7, ppc=2)
    ns = d["GasSpecies"]["NeutralSpecies"]
    for k, name in enumerate([n for n in ("Ar*", "O2A", "O2B", "OP", "OD") if n in ns]):
        ns[name]["Quenching"] = 0.1 * (k + 1)      # a distinct wall factor per species
    h = HostSetup(d, prec, load=False)
    o = Oracle(d, pn)
    names = OXY_FLUID_ARRAYS + (["BackDen2", "BackVel2"] if gas == 2 else [])
    for a in names:
        x, y = h.array(a), o.arr(a)
        assert x.shape == y.shape, a
        assert np.array_equal(x, y.astype(x.dtype)), f"gas {gas}/{pn}: {a}"
    T = int(o.scalar("TnRct"))
    assert h.scalar("TnRct") == T
    ntab = 58 if gas == 1 else 68
    assert np.array_equal(h.array("RR")[ntab:T], o.arr("RR")[ntab:T])
    for sc in ("dtc", "fl_nx", "fl_ny", "nfsp"):
        assert h.scalar(sc) == o.scalar(sc), sc
