"""Pins of the CPU oracle (no GPU).

Each check fixes the oracle's output by something other than its own
formulas: published known answers (Philox4x32-10, Park-Miller), output of
the reference's own functions compiled from its sources (oracle/_ref) or the
fixtures written from it (tests/golden), closed forms of the discrete field
problem, exact invariants of push, walls and deposit, libm for the
deterministic elementary functions, and collision statistics.
"""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest

from oracle import Oracle
from picmcc import decks

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLD = os.path.join(HERE, "golden")


def lib(p="f64"):
    return Oracle.__new__(Oracle) if False else C.CDLL(os.path.join(ROOT, "oracle", f"liborc_{p}.so"))


# ---------------------------------------------------------------- Philox
# Random123 known-answer vectors for philox4x32-10 (Salmon et al., SC'11)
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,expect", PHILOX_KAT)
def test_philox_known_answers(ctr, key, expect):
    L = lib()
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    L.orc_philox(c, k, o)
    assert tuple(o) == expect


# ---------------------------------------------------------------- Park-Miller
def test_park_miller_known_answer():
    """Minimal standard generator: x_{n+1} = 16807 x_n mod (2^31-1); from
    seed 1 the 10000th value is 1043618065 (Park & Miller, CACM 1988)."""
    L = lib()
    out = (C.c_long * 10000)()
    L.orc_lcg_seq(C.c_long(1), 10000, out)
    assert out[0] == 16807 and out[1] == 282475249 and out[2] == 1622650073
    assert out[9999] == 1043618065


# ---------------------------------------------------------------- detmath
@pytest.mark.parametrize("fn,np_fn,lo,hi,rel", [
    (0, np.exp, -60, 60, 3e-15), (2, np.log10, 1e-30, 1e30, 3e-15), (3, np.sin, -40, 40, 1e-15),
    (4, np.cos, -40, 40, 1e-15), (6, np.arctan, -1e4, 1e4, 1e-15), (5, np.tan, -1.5, 1.5, 2e-15)])
def test_detmath_against_libm(fn, np_fn, lo, hi, rel):
    L = lib()
    L.orc_detmath.restype = C.c_double
    L.orc_detmath.argtypes = [C.c_int, C.c_double, C.c_double]
    rng = np.random.default_rng(fn)
    if fn == 2:
        xs = 10.0 ** rng.uniform(math.log10(lo), math.log10(hi), 20000)
    else:
        xs = rng.uniform(lo, hi, 20000)
    got = np.array([L.orc_detmath(fn, x, 0.0) for x in xs])
    want = np_fn(xs)
    scale = np.maximum(np.abs(want), 1e-3 if fn in (3, 4, 5) else 1e-300)
    assert np.max(np.abs(got - want) / scale) < rel


def test_detmath_pow():
    L = lib()
    L.orc_detmath.restype = C.c_double
    L.orc_detmath.argtypes = [C.c_int, C.c_double, C.c_double]
    rng = np.random.default_rng(7)
    a = rng.uniform(1.0, 500.0, 5000)
    b = rng.uniform(0.0, 1.0, 5000)
    got = np.array([L.orc_detmath(7, x, y) for x, y in zip(a, b)])
    assert np.max(np.abs(got / np.power(a, b) - 1)) < 5e-15


def _detmathf(fn, x, y=None):
    L = lib("f32")
    L.orc_detmathf.restype = None
    L.orc_detmathf.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.zeros_like(x) if y is None else np.ascontiguousarray(y, dtype=np.float32)
    out = np.empty_like(x)
    L.orc_detmathf(fn, len(x), x.ctypes.data, y.ctypes.data, out.ctypes.data)
    return out


def _ulps(got, want, floor=0.0):
    """|got - want| in units of the float spacing at |want| (at least at floor)."""
    sp = np.spacing(np.maximum(np.abs(want).astype(np.float32), np.float32(floor))).astype(np.float64)
    return float(np.max(np.abs(got.astype(np.float64) - want) / sp))


@pytest.mark.parametrize("name,fn,make,np_fn,floor,bound", [
    ("log", 1, lambda r: np.float32(10.0) ** r.uniform(-30, 30, 100000).astype(np.float32), np.log, 0.0, 2.0),
    ("log10", 2, lambda r: np.float32(10.0) ** r.uniform(-30, 30, 100000).astype(np.float32), np.log10, 0.0, 3.5),
    ("exp", 0, lambda r: r.uniform(-103, 88, 100000).astype(np.float32), np.exp, 0.0, 1.5),
    ("sin", 3, lambda r: r.uniform(-1e4, 1e4, 100000).astype(np.float32), np.sin, 1.0, 1.5),
    ("cos", 4, lambda r: r.uniform(0, 2 * np.pi, 100000).astype(np.float32), np.cos, 1.0, 1.5),
    ("tan", 5, lambda r: r.uniform(0, 1.56, 100000).astype(np.float32), np.tan, 0.0, 3.0),
    ("atan", 6, lambda r: (np.float32(10.0) ** r.uniform(-6, 4, 100000)).astype(np.float32)
     * np.where(r.random(100000) < 0.5, -1, 1).astype(np.float32), np.arctan, 0.0, 3.0)])
def test_single_precision_detmath_against_libm(name, fn, make, np_fn, floor, bound):
    """The FP32 path's elementary functions (detmath.h dmf_*) against libm in
    double: sin and cos within 1.5 float ulps of 1 (absolute, as the
    collision phases need), the others relative to the result."""
    x = make(np.random.default_rng(fn + 100))
    assert _ulps(_detmathf(fn, x), np_fn(x.astype(np.float64)), floor) <= bound


def test_single_precision_pow():
    """pow(1 + E, R) of the scattering angle: a, b drawn as the collisions
    draw them; exp(b log a) loses |b log a| times the log's error (< 10 ulps
    here)."""
    r = np.random.default_rng(77)
    a = r.uniform(1.0, 1001.0, 100000).astype(np.float32)
    b = r.uniform(0.0, 1.0, 100000).astype(np.float32)
    assert _ulps(_detmathf(7, a, b), np.power(a.astype(np.float64), b.astype(np.float64))) <= 10.0
    assert _detmathf(0, [0.0])[0] == 1.0 and _detmathf(1, [1.0])[0] == 0.0
    assert _detmathf(6, [1.0])[0] == np.float32(np.pi / 4)


# ---------------------------------------------------------------- reference-built tables
def _load_gold(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def test_argon_cross_sections_match_reference_build():
    """cx table (FP32) equals Argon_CrossSectionSET's arithmetic run through the
    reference's own CrossSection/FUNC_decay (fixture from oracle/_ref)."""
    g = _load_gold("ar_xsec_table.json")
    o = Oracle(decks.small(), "f32")
    xe, xee, cx = o.arr("xe"), o.arr("xee"), o.arr("cx").reshape(81, 10)
    assert np.array_equal(xe, np.array(g["xe"], dtype=np.float32))
    assert np.array_equal(xee, np.array(g["xee"], dtype=np.float32))
    for r in range(7):
        assert np.array_equal(cx[:, r], np.array(g["cx"][r], dtype=np.float32)), f"reaction {r}"


def test_maxwell_table_matches_reference_build():
    g = _load_gold("maxwell_vsave.json")
    o = Oracle(decks.small(), "f32")
    vs = o.arr("vsave")
    assert len(vs) == g["nvel"]
    idx = np.array(g["index"])
    assert np.array_equal(vs[idx], np.array(g["value"], dtype=np.float32))
    assert float(np.sum(vs.astype(np.float64))) == pytest.approx(g["sum"], rel=0, abs=0)


def test_raw_cross_section_data_matches_reference_build():
    g = _load_gold("xsec_raw_checksum.json")
    o_sets = {}
    ref = os.path.join(ROOT, "oracle", "_ref", "ref_dump.txt")
    if not os.path.exists(ref):
        pytest.skip("oracle/_ref not built here (needs the reference tree)")
    with open(ref) as f:
        lines = f.read().split("\n")
    i = 0
    while i < len(lines):
        if lines[i].startswith("SET "):
            _, s, n = lines[i].split()
            vals = np.array([[float(v) for v in lines[i + 1 + k].split()] for k in range(int(n))])
            o_sets[int(s)] = vals
            i += int(n) + 1
        else:
            i += 1
    for s, chk in g.items():
        v = o_sets[int(s)]
        assert len(v) == chk["n"]
        assert float(np.sum(v[:, 0])) == pytest.approx(chk["sum_x"], rel=1e-12)
        assert float(np.sum(v[:, 1])) == pytest.approx(chk["sum_y"], rel=1e-12)


# ---------------------------------------------------------------- field closed forms
def test_laplace_solution_is_linear_between_plates():
    """Top row at 1 V, bottom row at 0 V, Neumann sides, uniform eps: the
    discrete solution is exactly phi = j / ncy."""
    d = decks.small(ncx=10, ncy=8)
    o = Oracle(d, "f64")
    ngx, ngy = 11, 9
    lap = o.arr("lap_phi").reshape(2, ngx, ngy)
    jj = np.arange(ngy) / (ngy - 1)
    assert np.allclose(lap[0], np.broadcast_to(jj, (ngx, ngy)), atol=1e-13)
    assert np.allclose(lap[1], 1 - np.broadcast_to(jj, (ngx, ngy)), atol=1e-13)


def test_poisson_uniform_charge_is_parabolic():
    """Uniform rhs c on every unknown with grounded plates: the 5-point
    solution is u_j = c/(2 k) j (n - j) with k = dx/dy (second differences
    of a quadratic are exact)."""
    d = decks.small(ncx=6, ncy=10)
    o = Oracle(d, "f64")
    b = o.arr("b")
    b[:] = 3.0
    o.orc_field()
    phi = o.arr("phi").reshape(7, 11)
    k = 1.0  # dx = dy
    j = np.arange(11)
    want = 3.0 / (2 * k) * j * (10 - j)
    assert np.allclose(phi, np.broadcast_to(want, (7, 11)), rtol=1e-12, atol=1e-12)


def test_temperature_uniform_without_neumann_sink():
    """All-Dirichlet walls at one temperature: the temperature operator has
    zero row sums, so T is that temperature everywhere."""
    d = decks.small(ncx=8, ncy=6)
    for b in d["Geometry"]["BoundaryCondition"]:
        b["B.C"] = 1
    o = Oracle(d, "f64")
    T = o.arr("Temp")
    assert np.allclose(T, 300.0, rtol=1e-12)


# ---------------------------------------------------------------- push / walls / deposit invariants
def _single_cell_state(o, ngy, cell, xs, ys, vxs, vys, vzs, G):
    n = len(xs)
    cs = np.zeros(G + 1, dtype=np.int32)
    cs[cell + 1:] = n
    return {"x": xs, "y": ys, "vx": vxs, "vy": vys, "vz": vzs, "id": np.arange(n, dtype=np.uint32),
            "cell_start": cs}


def test_push_uniform_field_closed_form():
    """Uniform E: v1 = v0 + E*Ascale, x1 = x0 + v1*dt/dx (leapfrog), no crossing."""
    d = decks.small(ncx=8, ncy=8)
    o = Oracle(d, "f64")
    ngy, G = 9, 81
    o.arr("Ex")[:] = 1234.0
    o.arr("Ey")[:] = -987.0
    cell = 3 * ngy + 4
    xs = np.array([0.25, 0.5, 0.75]); ys = np.array([0.5, 0.25, 0.6])
    vx = np.array([10.0, -20.0, 0.0]); vy = np.array([5.0, 0.0, -3.0]); vz = np.array([1.0, 2.0, 3.0])
    o.set_species(0, _single_cell_state(o, ngy, cell, xs, ys, vx, vy, vz, G))
    o.set_species(1, _single_cell_state(o, ngy, cell, xs[:0], ys[:0], vx[:0], vy[:0], vz[:0], G))
    o.orc_move()
    s = o.species(0)
    As = o.scalar("Ascale0")
    dtdx = o.scalar("dt") / o.scalar("dx")
    v1x, v1y = vx + 1234.0 * As, vy - 987.0 * As
    assert np.allclose(s["vx"], v1x, rtol=1e-14) and np.allclose(s["vy"], v1y, rtol=1e-14)
    assert np.allclose(s["x"], xs + v1x * dtdx, rtol=1e-13) and np.allclose(s["y"], ys + v1y * dtdx, rtol=1e-13)
    assert np.array_equal(s["vz"], vz)


def test_push_moves_at_most_one_cell_per_axis():
    """A particle fast enough to cross more than one cell in a step moves one
    cell per axis (the reference's single wrap, src/cuda_Move.cu:231-254),
    and its offset is clamped into that cell instead of staying outside
    [0, 1); a slower crosser keeps its exact wrapped offset."""
    d = decks.small(ncx=8, ncy=8)
    for prec, top in (("f64", np.nextafter(1.0, 0.0)), ("f32", np.nextafter(np.float32(1), np.float32(0)))):
        o = Oracle(d, prec)
        ngy, G = 9, 81
        o.arr("Ex")[:] = 0.0
        o.arr("Ey")[:] = 0.0
        cell = 3 * ngy + 4
        dtdx = o.scalar("dt") / o.scalar("dx")
        # cells per step: +2.5 and -1.7 in x (fast), +0.3 in y (slow crosser: y 0.9 -> 0.2)
        xs = np.array([0.5, 0.5]); ys = np.array([0.9, 0.5])
        vx = np.array([2.5, -1.7]) / dtdx; vy = np.array([0.3, 0.0]) / dtdx; vz = np.zeros(2)
        o.set_species(0, _single_cell_state(o, ngy, cell, xs, ys, vx, vy, vz, G))
        o.set_species(1, _single_cell_state(o, ngy, cell, xs[:0], ys[:0], vx[:0], vy[:0], vz[:0], G))
        o.orc_move()
        s = o.species(0)
        tag = o.arr("sp0.tag")[:2]
        assert list(tag) == [ngy + 1, -ngy], tag
        assert s["x"][0] == top and s["x"][1] == 0.0
        assert abs(s["y"][0] - 0.2) < 1e-6 and s["y"][1] == 0.5


def test_mirrored_offset_stays_inside_the_cell():
    """A fast particle driven into a Neumann wall is clamped onto the wall
    (offset 0) and mirrored; the mirrored offset is kept below 1."""
    d = decks.small(ncx=8, ncy=8)
    for prec, top in (("f64", np.nextafter(1.0, 0.0)), ("f32", np.nextafter(np.float32(1), np.float32(0)))):
        o = Oracle(d, prec)
        ngy, G = 9, 81
        o.arr("Ex")[:] = 0.0
        o.arr("Ey")[:] = 0.0
        cell = 0 * ngy + 4          # first column: the x = 0 wall is Neumann
        dtdx = o.scalar("dt") / o.scalar("dx")
        xs = np.array([0.5]); ys = np.array([0.5])
        vx = np.array([-1.7]) / dtdx; vy = np.zeros(1); vz = np.zeros(1)
        o.set_species(0, _single_cell_state(o, ngy, cell, xs, ys, vx, vy, vz, G))
        o.set_species(1, _single_cell_state(o, ngy, cell, xs[:0], ys[:0], vx[:0], vy[:0], vz[:0], G))
        o.orc_move()
        assert o.orc_sort() == 0
        s = o.species(0)
        assert len(s["x"]) == 1 and s["x"][0] == top and s["vx"][0] > 0 and s["y"][0] == 0.5


def test_deposit_weights_sum_to_particle_count():
    d = decks.small(ncx=8, ncy=8, ppc=30)
    o = Oracle(d, "f64")
    o.orc_deposit()
    den = o.arr("den").reshape(2, 9, 9)
    # undo the wall factors: x2 on Neumann sides, x4 on electrode faces, x8 on
    # corners - except the top-left corner, whose face lookup reads node
    # (1,0) (src/start.c:1765-1769 indexes ID+1 across the column) and
    # finds a conductor, so it keeps NO_FACE and no factor
    fac = np.ones((9, 9))
    fac[0, 1:-1] = fac[-1, 1:-1] = 2
    fac[:, 0] = fac[:, -1] = 4
    fac[0, 0] = fac[-1, 0] = fac[-1, -1] = 8
    fac[0, -1] = 1
    face = o.arr("Face").reshape(9, 9)
    assert np.all(face[1:-1, 0] == 2) and np.all(face[1:-1, -1] == 3)     # UP at bottom, DOWN at top
    assert face[0, -1] == 8 and face[0, 0] == 5 and face[-1, 0] == 4 and face[-1, -1] == 6
    for s in range(2):
        raw = den[s] / fac
        assert raw.sum() == pytest.approx(o.scalar(f"np{s}"), rel=1e-12)


def test_single_particle_cic_weights():
    d = decks.small(ncx=6, ncy=6)
    o = Oracle(d, "f64")
    ngy, G = 7, 49
    cell = 2 * ngy + 3
    one = lambda v: np.array([v])
    o.set_species(0, _single_cell_state(o, ngy, cell, one(0.25), one(0.75), one(0.), one(0.), one(0.), G))
    o.set_species(1, _single_cell_state(o, ngy, cell, one(0.5)[:0], one(0.)[:0], one(0.)[:0], one(0.)[:0],
                                        one(0.)[:0], G))
    o.orc_deposit()
    den = o.arr("den")[:G]
    assert den[cell] == pytest.approx(0.75 * 0.25)
    assert den[cell + 1] == pytest.approx(0.75 * 0.75)
    assert den[cell + ngy] == pytest.approx(0.25 * 0.25)
    assert den[cell + ngy + 1] == pytest.approx(0.25 * 0.75)
    assert np.count_nonzero(den) == 4


def test_walls_conserve_and_reflect():
    """With E = 0 and no collisions: every particle is either kept (its count
    unchanged) or removed through a Dirichlet electrode row; Neumann sides
    reflect (x -> 1-x, vx -> -vx)."""
    d = decks.small(ncx=6, ncy=6)
    for k in d["Collision_On_Off"]["ArgonCase"]:
        d["Collision_On_Off"]["ArgonCase"][k] = 0.0
    o = Oracle(d, "f64")
    ngy, G = 7, 49
    dx = o.scalar("dx")
    dt = o.scalar("dt")
    v = 0.3 * dx / dt   # 0.3 cells per step
    # particle 0: left wall cell moving left; particle 1: bottom row moving down
    st = {"x": np.array([0.1, 0.5]), "y": np.array([0.5, 0.1]), "vx": np.array([-v, 0.0]),
          "vy": np.array([0.0, -v]), "vz": np.array([0.0, 0.0]), "id": np.array([0, 1], dtype=np.uint32)}
    cells = [0 * ngy + 3, 2 * ngy + 0]
    order = np.argsort(cells)
    cs = np.zeros(G + 1, dtype=np.int32)
    for c in cells:
        cs[c + 1:] += 1
    st = {k: v_[order] for k, v_ in st.items()}
    st["cell_start"] = cs
    o.set_species(0, st)
    o.set_species(1, {k: (v_[:0] if k != "cell_start" else np.zeros(G + 1, np.int32)) for k, v_ in st.items()})
    o.arr("Ex")[:] = 0
    o.arr("Ey")[:] = 0
    o.orc_move()
    assert o.orc_sort() == 0
    s = o.species(0)
    assert len(s["x"]) == 1                       # the one heading into the grounded row is gone
    assert s["id"][0] == 0
    assert s["x"][0] == pytest.approx(0.2)        # 0.1 - 0.3 reflected about the wall
    assert s["vx"][0] == pytest.approx(v)
    assert s["cell_start"][3 + 1] - s["cell_start"][3] == 1


# ---------------------------------------------------------------- collision statistics
def test_null_collision_fraction_matches_probability():
    """Electrons in one cell: the fraction that makes any collision draw
    equals 1 - (1 - P2)^MCCn (binomial, 5 sigma)."""
    d = decks.small(ncx=4, ncy=4, capacity_factor=100.0)
    o = Oracle(d, "f64")
    ngy, G = 5, 25
    cell = 2 * ngy + 2
    n = 20000
    rng = np.random.default_rng(3)
    vth = 1.0e6
    st = {"x": rng.random(n), "y": rng.random(n), "vx": rng.normal(0, vth, n), "vy": rng.normal(0, vth, n),
          "vz": rng.normal(0, vth, n), "id": np.arange(n, dtype=np.uint32)}
    cs = np.zeros(G + 1, dtype=np.int32)
    cs[cell + 1:] = n
    st["cell_start"] = cs
    o.set_species(0, st)
    o.set_species(1, {k: (v[:0] if k != "cell_start" else np.zeros(G + 1, np.int32)) for k, v in st.items()})
    o.orc_set_tstep(1)
    assert o.orc_mcc() == 0
    ev = o.arr("mccr").reshape(G, 10)[cell]
    # every drawn collision picks a target; null outcomes are not counted, so
    # compare the count of real events against its expectation per reaction
    P2, MCCn = o.arr("P2")[cell], int(o.scalar("MCCn"))
    p_any = 1 - (1 - P2) ** MCCn
    # expected real events: sum over particles of p_any * sigma_tot(E) v / sigv
    E = o.scalar("Escale0") * (st["vx"] ** 2 + st["vy"] ** 2 + st["vz"] ** 2)
    L = o.lib
    L.orc_ar_sigma.restype = C.c_double
    L.orc_ar_sigma.argtypes = [C.c_void_p, C.c_int, C.c_double]
    sig = np.array([sum(L.orc_ar_sigma(o.h, r, e) * (e > th) if r else L.orc_ar_sigma(o.h, r, e)
                        for r, th in [(0, 0), (1, 11.55), (2, 11.548), (3, 15.76)]) for e in E])
    vel = np.sqrt(st["vx"] ** 2 + st["vy"] ** 2 + st["vz"] ** 2)
    P1 = o.arr("P1")[cell]
    p_target0 = P1 / P2 if P2 > 0 else 0
    mean = np.sum(p_any * p_target0 * np.minimum(sig * vel / o.arr("sigv")[0], 1.0))
    total = ev[0:4].sum()
    assert abs(total - mean) < 5 * math.sqrt(mean) + 5


def test_poisson_oracle_operator_is_the_field_operator():
    """For uniform permittivity and dx = dy, the batched-Poisson oracle's
    operator equals the field operator restated from src/load.c:121-245
    (unknown ordering a*ny + b), up to the operator's constant scale."""
    from picmcc import FP64_PLAIN, HostSetup, decks
    from poisson_oracle import operator
    d = decks.small(ncx=6, ncy=5)
    h = HostSetup(d, FP64_PLAIN, load=False)
    Ai, Aj, Av = h.array("Ai"), h.array("Aj"), h.array("Aval")
    N = len(Ai) - 1
    A = np.zeros((N, N))
    for r in range(N):
        for q in range(Ai[r], Ai[r + 1]):
            A[r, Aj[q]] = Av[q]
    # CCP deck: Neumann sides in x, conductors on the bottom and top rows
    nx, ny = 7, 4
    assert N == nx * ny
    B = operator(nx, ny, (0, 0, 1, 1)).toarray()
    scale = A[0, 0] / B[0, 0]
    np.testing.assert_allclose(A, scale * B, rtol=0, atol=1e-12 * abs(scale))


@pytest.mark.parametrize("pn", ["f64", "f32"])
def test_particle_balance_closes(pn):
    """Every particle that appears or disappears is counted: after any
    number of steps N_s = N_s(0) + ionisations - absorbed_s - deleted_s for
    both species (each ionisation adds one electron and one ion)."""
    d = decks.ccp_deck(ncx=32, ncy=24, dx=5e-4, volts=250.0, pres_torr=0.3, ppc=30, steps_per_cycle=400,
                       cycles=1, seed=11, philox_seed=22)
    o = Oracle(d, pn)
    o.orc_deposit()
    n0 = [int(o.scalar("np0")), int(o.scalar("np1"))]
    assert o.orc_run(60) == 0
    n1 = [int(o.scalar("np0")), int(o.scalar("np1"))]
    m = o.arr("mccr").reshape(-1, int(o.scalar("TnRct"))).astype(np.int64)
    pairs = int(m[:, 3].sum() + m[:, 4].sum())
    absorbed = o.arr("CondCharge").reshape(2, -1).sum(axis=1)
    lost = o.arr("Lost")
    assert pairs > 0 and lost[0] > 0   # the case ionises and loses electrons at the walls
    for s in range(2):
        assert n1[s] == n0[s] + pairs - int(absorbed[s]) - int(lost[s])



# ---------------------------------------------------------------- collision kinematics and channels
def _kin(o, which, vals):
    L = o.lib
    L.orc_kinematics.restype = C.c_int
    L.orc_kinematics.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    inp = (C.c_double * len(vals))(*vals)
    out = (C.c_double * 3)()
    assert L.orc_kinematics(which, inp, out) == 0
    return np.array(out[:])


@pytest.mark.parametrize("pn,rtol", [("f64", 1e-12), ("f32", 2e-5)])
def test_electron_scattering_angle_closed_form(pn, rtol):
    """anewvel keeps the speed, turns the velocity by cos(chi) =
    (2 + E - 2 (1+E)^R1) / E about the incident direction and spreads the
    azimuth over 2 pi R2 (R2 and R2 + 1/2 give opposite transverse parts);
    a zero-energy electron keeps its direction."""
    o = Oracle(decks.small(ncx=4, ncy=4), pn)
    rng = np.random.default_rng(5)
    dirs = [np.array([0.0, 0.0, 1.0]), np.array([1.0, 0.0, 0.0])]
    for _ in range(6):
        u = rng.normal(size=3)
        dirs.append(u / np.linalg.norm(u))
    for E in (0.3, 4.0, 60.0):
        for u in dirs:
            for r1 in (0.05, 0.5, 0.93):
                vel = 1.7e6
                a = _kin(o, 0, [E, vel, *u, r1, 0.2])
                b = _kin(o, 0, [E, vel, *u, r1, 0.7])
                assert abs(np.linalg.norm(a) - vel) <= rtol * vel * 10
                cos_chi = (2 + E - 2 * (1 + E) ** r1) / E
                assert abs(a @ u / vel - cos_chi) <= rtol * 50
                ta, tb = a - (a @ u) * u, b - (b @ u) * u
                assert np.allclose(ta, -tb, atol=rtol * 50 * vel)
    u = dirs[1]
    assert np.allclose(_kin(o, 0, [0.0, 2.0, *u, 0.3, 0.4]), 2.0 * u)


@pytest.mark.parametrize("pn,rtol", [("f64", 1e-12), ("f32", 2e-5)])
def test_ion_scattering_closed_form(pn, rtol):
    """Equal-mass hard spheres (newvel_ionsc): the relative speed is scaled by
    cos(chi) = sqrt(R1) and the velocity turned by chi."""
    o = Oracle(decks.small(ncx=4, ncy=4), pn)
    rng = np.random.default_rng(6)
    for _ in range(20):
        u = rng.normal(size=3)
        vel = 3.1e3
        v = vel * u / np.linalg.norm(u)
        r1, r2 = rng.random(), rng.random()
        w = _kin(o, 1, [*v, vel, r1, r2])
        c = np.sqrt(r1)
        assert abs(np.linalg.norm(w) - vel * c) <= rtol * vel * 10
        if c > 1e-3:
            assert abs(w @ v / (np.linalg.norm(w) * vel) - c) <= rtol * 100


@pytest.mark.parametrize("pn,rtol", [("f64", 1e-12), ("f32", 2e-5)])
def test_neutral_velocity_draw_is_isotropic(pn, rtol):
    """dmaxwellv: speed vti * vsave_n, polar cosine 1 - 2 Rthe, azimuth
    2 pi Rphi (uniform on the sphere for uniform Rthe, Rphi)."""
    o = Oracle(decks.small(ncx=4, ncy=4), pn)
    for vs, vti, rphi, rthe in [(1.3, 400.0, 0.1, 0.2), (0.7, 250.0, 0.6, 0.9), (2.2, 300.0, 0.35, 0.5)]:
        v = _kin(o, 2, [vs, vti, rphi, rthe])
        s = vs * vti
        assert abs(np.linalg.norm(v) - s) <= rtol * s * 10
        assert abs(v[2] / s - (1 - 2 * rthe)) <= rtol * 10
        phi = math.atan2(v[1], v[0]) % (2 * math.pi)
        assert abs(phi - 2 * math.pi * rphi) <= rtol * 100


@pytest.mark.parametrize("pn,rtol", [("f64", 1e-12), ("f32", 2e-5)])
def test_ionisation_energy_sharing_inverts(pn, rtol):
    """The secondary electron's energy 10 tan(u atan(E/20)) has the CDF
    atan(e/10) / atan(E/20) on [0, E/2]: inverting it returns u."""
    o = Oracle(decks.small(ncx=4, ncy=4), pn)
    for E in (1.0, 20.0, 150.0):
        prev = -1.0
        for u in np.linspace(0.0, 1.0, 11):
            e = _kin(o, 3, [E, u])[0]
            assert 0 <= e <= E / 2 * (1 + rtol * 10) and e >= prev
            prev = e
            assert abs(math.atan(e / 10) / math.atan(E / 20) - u) <= rtol * 100


def _beam(o, G, cell, n, v, species):
    st = {"x": np.full(n, 0.5), "y": np.full(n, 0.5), "vx": np.full(n, v[0]), "vy": np.full(n, v[1]),
          "vz": np.full(n, v[2]), "id": np.arange(n, dtype=np.uint32)}
    cs = np.zeros(G + 1, dtype=np.int32)
    cs[cell + 1:] = n
    st["cell_start"] = cs
    empty = {k: (val[:0] if k != "cell_start" else np.zeros(G + 1, np.int32)) for k, val in st.items()}
    o.set_species(species, st)
    o.set_species(1 - species, empty)


def test_electron_channel_fractions():
    """A mono-energetic electron beam that always collides (P2 = 1): on the
    argon target (P1 = 1) reaction r happens with probability
    sigma_r(E) v / nu_max[Ar], on the metastable target (P1 = 0) reaction 4
    with sigma_4(E) v / nu_max[Ar*] (binomial, 5 sigma per channel)."""
    d = decks.small(ncx=4, ncy=4, capacity_factor=400.0)
    o = Oracle(d, "f64")
    ngy, G, cell, n = 5, 25, 2 * 5 + 2, 60000
    L = o.lib
    L.orc_ar_sigma.restype = C.c_double
    L.orc_ar_sigma.argtypes = [C.c_void_p, C.c_int, C.c_double]
    Es = o.scalar("Escale0")
    for E in (8.0, 30.0, 120.0):
        vel = math.sqrt(E / Es)
        for target, p1 in ((0, 1.0), (1, 0.0)):
            _beam(o, G, cell, n, (0.0, 0.0, vel), 0)
            o.arr("P2")[:] = 1.0
            o.arr("P1")[:] = p1
            o.arr("mccr")[:] = 0
            o.orc_set_tstep(3)
            assert o.orc_mcc() == 0
            ev = o.arr("mccr").reshape(G, 10)[cell]
            sigv = o.arr("sigv")
            th = {1: 11.55, 2: 11.548, 3: 15.76, 4: 4.22}
            rs = (0, 1, 2, 3) if target == 0 else (4,)
            for r in rs:
                sig = L.orc_ar_sigma(o.h, r, E) if E > th.get(r, 0.0) else 0.0
                p = sig * vel / sigv[target]
                mean = n * p
                assert abs(ev[r] - mean) <= 5 * math.sqrt(mean * (1 - p)) + 3, (E, r, ev[r], mean)
            other = (4,) if target == 0 else (0, 1, 2, 3)
            assert all(ev[r] == 0 for r in other)


def test_ion_channel_fractions_and_charge_exchange():
    """A mono-energetic ion beam through a cold background (BackVel1 = 0),
    always colliding (Pion = 1): charge exchange (5) and elastic scattering
    (6) happen with probabilities sigma_r(E) v / nu_max[ion]; after a charge
    exchange the ion has the neutral's velocity (zero here) and after an
    elastic collision its speed is below the incident one."""
    d = decks.small(ncx=4, ncy=4, capacity_factor=400.0)
    o = Oracle(d, "f64")
    ngy, G, cell, n = 5, 25, 2 * 5 + 2, 60000
    L = o.lib
    L.orc_ar_sigma.restype = C.c_double
    L.orc_ar_sigma.argtypes = [C.c_void_p, C.c_int, C.c_double]
    Ei = o.scalar("Escale1")
    for E in (0.5, 5.0, 50.0):
        vel = math.sqrt(E / Ei)
        _beam(o, G, cell, n, (vel, 0.0, 0.0), 1)
        o.arr("Pion")[:] = 1.0
        o.arr("BackVel1")[:] = 0.0
        o.arr("mccr")[:] = 0
        o.orc_set_tstep(2)
        assert o.orc_mcc() == 0
        ev = o.arr("mccr").reshape(G, 10)[cell]
        sigv2 = o.arr("sigv")[2]
        for r in (5, 6):
            p = L.orc_ar_sigma(o.h, r, E) * vel / sigv2
            mean = n * p
            assert abs(ev[r] - mean) <= 5 * math.sqrt(mean * (1 - p)) + 3, (E, r, ev[r], mean)
        s = o.species(1)
        speed = np.sqrt(s["vx"] ** 2 + s["vy"] ** 2 + s["vz"] ** 2)
        assert np.count_nonzero(speed == 0.0) == ev[5]
        assert np.count_nonzero((speed > 0) & (speed < vel * (1 - 1e-12))) == ev[6]
