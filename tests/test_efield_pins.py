"""Closed-form pins of the oracle's E-field stage (oracle/orc_step.c:
orc_voltages, cond_sigma, v_to_efield, orc_efield; orc_accumulate).

The oracle restates src/cuda_Field.cu:2-42 (conductor voltages and their
superposition), 500-558 (surface charge at conductor faces) and 698-846
(node E), and src/cuda_Diagnostic.cu:505-547 (accumulators).  The GPU path
is bit-exact with it, so these pins carry over.  Each check compares with
what the physics fixes, on a grid with dx != dy so that a swapped spacing or
index shows:

- interior and Neumann nodes: E = -grad(phi) for an arbitrary quadratic
  phi, where central differences are exact, and Ex = 0 on the x walls;
- parallel plates: the driven electrode's V(t) = DC + AC sin(2 pi f t +
  phase) (phase in degrees) superposed on the unit Laplace solution gives
  Ey = -V(t) / L everywhere between the plates;
- Gauss's law: a uniform charge density rho between grounded plates has
  phi = rho y (L - y) / (2 eps0) (FP64 direct solve), surface charge -rho L / 2
  on each plate and E = +-rho L / (2 eps0) at the plates;
- accumulators: the sums are the sums of the per-step arrays.

The reference's eps0 is 8.8542e-12.
"""
import math
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "hybrid-final_amd"))
from oracle import Oracle  # noqa: E402
from picmcc import decks  # noqa: E402

EPS0 = 8.8542e-12
DX, DY = 5e-4, 3e-4
NCX, NCY = 7, 6


def _deck(dc=0.0, ac=100.0, phase=0.0, freq=13.56e6):
    d = decks.small(ncx=NCX, ncy=NCY, ppc=2, dx=DX, volts=ac)
    d["Geometry"]["SystemSpec"]["Y_length(m)"] = NCY * DY
    src = d["Geometry"]["Source"][0]
    src["DC(V)"], src["Phase(deg)"], src["Freq(1/s)"] = dc, phase, freq
    return d


def _grid(o):
    ngx, ngy = int(o.scalar("ngx")), int(o.scalar("ngy"))
    g = np.arange(ngx * ngy)
    return ngx, ngy, (g // ngy) * DX, (g % ngy) * DY, g // ngy, g % ngy


def test_grid_spacing():
    o = Oracle(_deck(), "f64")
    assert o.scalar("dx") == DX and abs(o.scalar("dy") - DY) < 1e-18
    assert int(o.scalar("ngx")) == NCX + 1 and int(o.scalar("ngy")) == NCY + 1


def test_interior_and_neumann_field_is_minus_gradient():
    """phi = a x^2 + b x y + c y^2 + d x + e y: central differences are
    exact, so E = -(2 a x + b y + d, b x + 2 c y + e) at interior nodes; on
    the Neumann side walls Ex = 0 and Ey is the same central difference."""
    o = Oracle(_deck(), "f64")
    ngx, ngy, X, Y, ix, iy = _grid(o)
    a, b, c, d, e = 3.0e5, -2.0e5, 7.0e5, 40.0, -90.0
    phi = a * X * X + b * X * Y + c * Y * Y + d * X + e * Y
    o.arr("den")[:] = 0
    o.arr("phi")[:] = phi
    o.orc_set_tstep(0)        # t = 0: every conductor voltage is 0
    o.orc_efield()
    assert np.array_equal(o.arr("TotPot"), phi)
    Ex, Ey = o.arr("Ex"), o.arr("Ey")
    inner = (ix > 0) & (ix < ngx - 1) & (iy > 0) & (iy < ngy - 1)
    np.testing.assert_allclose(Ex[inner], -(2 * a * X + b * Y + d)[inner], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(Ey[inner], -(b * X + 2 * c * Y + e)[inner], rtol=1e-9, atol=1e-9)
    side = ((ix == 0) | (ix == ngx - 1)) & (iy > 0) & (iy < ngy - 1)
    assert np.all(Ex[side] == 0)
    np.testing.assert_allclose(Ey[side], -(b * X + 2 * c * Y + e)[side], rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("dc,ac,phase", [(0.0, 100.0, 0.0), (20.0, 150.0, 30.0), (-35.0, 80.0, 250.0)])
def test_driven_plates_give_uniform_field(dc, ac, phase):
    """Top electrode V(t) = DC + AC sin(2 pi f t + phase), bottom grounded,
    no charge: the potential is linear between the plates and Ey = -V / L at
    every node that is not on an electrode, Ex = 0."""
    freq = 13.56e6
    o = Oracle(_deck(dc, ac, phase, freq), "f64")
    ngx, ngy, X, Y, ix, iy = _grid(o)
    L = NCY * DY
    o.arr("den")[:] = 0
    o.arr("phi")[:] = 0
    dt = o.scalar("dt")
    for k in (0, 37, 111, 250):
        o.orc_set_tstep(k)
        o.orc_efield()
        t = k * dt
        v = dc + ac * math.sin(2 * math.pi * freq * t + math.pi / 180 * phase)
        vc = o.arr("V_cond")
        assert abs(vc[0] - v) <= 1e-12 * (abs(dc) + abs(ac)) and vc[1] == 0
        np.testing.assert_allclose(o.arr("TotPot"), v * Y / L, rtol=1e-12, atol=1e-12 * abs(ac))
        gap = (iy > 0) & (iy < ngy - 1)
        np.testing.assert_allclose(o.arr("Ey")[gap], -v / L, rtol=1e-10, atol=1e-10 * abs(ac) / L)
        np.testing.assert_allclose(o.arr("Ex")[gap], 0.0, atol=1e-10 * abs(ac) / L)


@pytest.mark.parametrize("orient", ["vertical", "horizontal"])
def test_gauss_law_for_uniform_charge(orient):
    """Uniform charge density rho between grounded plates (t = 0): the FP64
    direct solve gives phi = rho u (L - u) / (2 eps0) (u across the gap);
    each plate carries sigma = -rho L / 2 and the field at the plates is
    -+rho L / (2 eps0) (pointing into the plates); inside it is
    -rho (L - 2 u) / (2 eps0)."""
    o = Oracle(_layered_deck(orient, 0, 1.0), "f64")
    ngx, ngy, X, Y, ix, iy = _grid(o)
    if orient == "vertical":
        L, u, ku, kt, nu, nt = NCY * DY, Y, iy, ix, ngy, ngx
    else:
        L, u, ku, kt, nu, nt = NCX * DX, X, ix, iy, ngx, ngy
    q1 = o.scalar("q_density1")     # charge per ion macro-particle
    n0 = 3.0                        # ion macro-particles per node control volume
    rho = n0 * q1 / (DX * DY)
    den = o.arr("den")
    G = ngx * ngy
    den[:] = 0
    den[G:2 * G] = n0
    o.arr("b")[:] = n0 * q1 / EPS0
    o.orc_field()
    want = rho * u * (L - u) / (2 * EPS0)
    np.testing.assert_allclose(o.arr("phi"), want, rtol=1e-9, atol=1e-12 * want.max())
    o.orc_set_tstep(0)
    o.orc_efield()
    En = o.arr("Ey") if orient == "vertical" else o.arr("Ex")
    Et = o.arr("Ex") if orient == "vertical" else o.arr("Ey")
    sig = o.arr("Sigma")
    scale = rho * L / (2 * EPS0)
    low = (ku == 0) & (kt > 0) & (kt < nt - 1)
    high = (ku == nu - 1) & (kt > 0) & (kt < nt - 1)
    np.testing.assert_allclose(sig[low], -rho * L / 2, rtol=1e-9)
    np.testing.assert_allclose(sig[high], -rho * L / 2, rtol=1e-9)
    np.testing.assert_allclose(En[low], -scale, rtol=1e-9)
    np.testing.assert_allclose(En[high], scale, rtol=1e-9)
    assert np.all(Et[low] == 0) and np.all(Et[high] == 0)
    gap = (ku > 0) & (ku < nu - 1)
    np.testing.assert_allclose(En[gap], (-rho * (L - 2 * u) / (2 * EPS0))[gap], rtol=1e-9, atol=1e-9 * scale)
    np.testing.assert_allclose(Et[gap], 0.0, atol=1e-9 * scale)


def test_accumulators_sum_the_steps():
    o = Oracle(_deck(), "f64")
    ngx, ngy, X, Y, ix, iy = _grid(o)
    G = ngx * ngy
    rng = np.random.default_rng(1)
    for name in ("sum_den", "sum_phi", "sum_ex", "sum_ey"):
        o.arr(name)[:] = 0
    want = {k: 0 for k in ("den", "TotPot", "Ex", "Ey")}
    for _ in range(3):
        for k in want:
            a = rng.standard_normal(o.arr(k).size)
            o.arr(k)[:] = a
            want[k] = want[k] + a
        o.orc_accumulate()
    np.testing.assert_allclose(o.arr("sum_den"), want["den"], rtol=1e-15, atol=1e-15)
    np.testing.assert_allclose(o.arr("sum_phi"), want["TotPot"], rtol=1e-15, atol=1e-15)
    np.testing.assert_allclose(o.arr("sum_ex"), want["Ex"], rtol=1e-15, atol=1e-15)
    np.testing.assert_allclose(o.arr("sum_ey"), want["Ey"], rtol=1e-15, atol=1e-15)
    assert o.scalar("n_accum") >= 3
    assert o.arr("sum_den").size == 2 * G


def _layered_deck(orient, nd, eps_r):
    """Plates normal to y ("vertical": grounded bottom, driven top) or to x
    ("horizontal": grounded left, driven right), the other walls Neumann, and
    a dielectric slab of nd cells on the grounded plate."""
    d = _deck()
    g = d["Geometry"]
    if orient == "horizontal":
        g["BoundaryCondition"] = [
            {"X0": 0, "X1": 0, "Y0": 0, "Y1": NCY, "B.C": 1, "Temp": 300.0},
            {"X0": NCX, "X1": NCX, "Y0": 0, "Y1": NCY, "B.C": 1, "Temp": 300.0},
            {"X0": 0, "X1": NCX, "Y0": 0, "Y1": 0, "B.C": 2, "Temp": 300.0},
            {"X0": 0, "X1": NCX, "Y0": NCY, "Y1": NCY, "B.C": 2, "Temp": 300.0}]
        g["ConductorSpec"][0].update({"X0": NCX, "X1": NCX, "Y0": 0, "Y1": NCY})
        g["ConductorSpec"][1].update({"X0": 0, "X1": 0, "Y0": 0, "Y1": NCY})
        slab = {"M_ID": 3, "X0": 0, "X1": nd, "Y0": 0, "Y1": NCY, "Epsilon": eps_r}
    else:
        slab = {"M_ID": 3, "X0": 0, "X1": NCX, "Y0": 0, "Y1": nd, "Epsilon": eps_r}
    g["DielectricSpec"] = [slab] if nd > 0 else []
    return d


@pytest.mark.parametrize("orient", ["vertical", "horizontal"])
def test_layered_dielectric_capacitor(orient):
    """A dielectric slab (eps_r, thickness D) on the grounded plate, uniform
    charge rho in the gas, both plates at 0 V.  The normal displacement is
    continuous at the slab surface, so phi = s1 u in the slab and
    s1 D + eps_r s1 (u - D) - rho (u - D)^2 / (2 eps0) in the gas, with
    s1 = rho (L - D)^2 / (2 eps0 (D + eps_r (L - D))); u is the distance
    from the grounded plate.  The FP64 solve reproduces it node for node, the
    grounded plate carries sigma = -eps0 eps_r s1 and its field -s1 (the
    slab's), the driven plate sigma = -rho (L - D) + eps0 eps_r s1 (the two
    add up to minus the gas charge), the slab's interior and side-wall
    field is -s1, and on the slab surface the field is the gas-side
    difference -(phi(D + h) - phi(D)) / h, as the reference takes it."""
    nd, er = 2, 4.0
    o = Oracle(_layered_deck(orient, nd, er), "f64")
    ngx, ngy, X, Y, ix, iy = _grid(o)
    G = ngx * ngy
    if orient == "vertical":
        h, L, u, ku, kt, nu = DY, NCY * DY, Y, iy, ix, ngy
    else:
        h, L, u, ku, kt, nu = DX, NCX * DX, X, ix, iy, ngx
    D = nd * h
    q1 = o.scalar("q_density1")
    n0 = 2.0
    rho = n0 * q1 / (DX * DY)
    s1 = rho * (L - D) ** 2 / (2 * EPS0 * (D + er * (L - D)))
    gas = ku > nd
    want = np.where(ku <= nd, s1 * u, s1 * D + er * s1 * (u - D) - rho * (u - D) ** 2 / (2 * EPS0))
    den = o.arr("den")
    den[:] = 0
    den[G:2 * G] = np.where(gas, n0, np.where(ku == nd, n0 / 2, 0.0))
    aidx = o.arr("A_idx")
    b = o.arr("b")
    for g_ in range(G):
        if aidx[g_]:
            b[aidx[g_] - 1] = den[G + g_] * q1 / EPS0
    o.orc_field()
    np.testing.assert_allclose(o.arr("phi"), np.where((ku == 0) | (ku == nu - 1), 0.0, want), rtol=1e-9,
                               atol=1e-12 * np.abs(want).max())
    o.orc_set_tstep(0)
    o.orc_efield()
    En = o.arr("Ey") if orient == "vertical" else o.arr("Ex")
    Et = o.arr("Ex") if orient == "vertical" else o.arr("Ey")
    sig = o.arr("Sigma")
    inner_t = (kt > 0) & (kt < (ngx if orient == "vertical" else ngy) - 1)
    ground = (ku == 0) & inner_t
    driven = (ku == nu - 1) & inner_t
    np.testing.assert_allclose(sig[ground], -EPS0 * er * s1, rtol=1e-9)
    np.testing.assert_allclose(En[ground], -s1, rtol=1e-9)
    np.testing.assert_allclose(sig[driven], -rho * (L - D) + EPS0 * er * s1, rtol=1e-9)
    e_L = -(er * s1 - rho * (L - D) / EPS0)
    np.testing.assert_allclose(En[driven], e_L, rtol=1e-9)
    assert np.all(Et[ground] == 0) and np.all(Et[driven] == 0)
    slab = (ku > 0) & (ku < nd)
    np.testing.assert_allclose(En[slab], -s1, rtol=1e-9)
    surface = (ku == nd) & inner_t
    step = s1 * D + er * s1 * h - rho * h * h / (2 * EPS0) - s1 * D
    np.testing.assert_allclose(En[surface], -step / h, rtol=1e-9)
    inside_gas = gas & (ku < nu - 1)
    np.testing.assert_allclose(En[inside_gas], (-(er * s1 - rho * (u - D) / EPS0))[inside_gas], rtol=1e-9,
                               atol=1e-9 * abs(s1))
