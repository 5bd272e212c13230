"""Pins of the oracle's metastable fluid (the checker of the GPU fluid step)
by what the mathematics fixes, not by its own formulas:

- the Thomas recurrence solves tridiagonal systems (numpy's dense solve);
- the collision-integral table reproduces the reference's loop, whose
  missing break makes every in-range value take the last interval's line;
- the plasma lines are the maximal runs of plasma cells (numpy run lengths
  on a deck with a conductor inside the plasma);
- with zero wall flux factors the continuity solve conserves the total
  density, keeps a uniform density uniform, and multiplies each cosine mode
  along x by (1 - lam) / (1 + lam), lam = 2 a (1 - cos(pi k / n)),
  a = D dtc / dx^2 (lam: the cell-centred Neumann Laplacian's eigenvalues;
  implicit in the x sweep, explicit in the y sweep that follows);
- the source update uses N_ave accumulations plus the averaging step, the
  events since the previous average over N_ave steps, and leaves P2 equal to
  the set-up's when the node density has not moved.
"""
import numpy as np
import pytest

from oracle import Oracle
from picmcc import decks


def test_tridiag_solves_the_system():
    o = Oracle(decks.small(ncx=4, ncy=3, ppc=2), "f64")
    import ctypes as C
    L = o.lib
    dp = C.POINTER(C.c_double)
    L.orc_tridiag.argtypes = [dp] * 6 + [C.c_int]
    L.orc_tridiag.restype = C.c_int
    rng = np.random.default_rng(5)
    for n in (1, 2, 7, 64):
        a, c, d = rng.uniform(-1, 0, n), rng.uniform(-1, 0, n), rng.uniform(-5, 5, n)
        b = 2.5 + rng.uniform(0, 1, n)
        x, gam = np.zeros(n), np.zeros(n)
        p = lambda v: v.ctypes.data_as(dp)
        assert L.orc_tridiag(p(a), p(b), p(c), p(d), p(x), p(gam), n) == 1
        M = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
        assert np.allclose(x, np.linalg.solve(M, d), rtol=1e-12, atol=1e-12)


def test_collision_integral_table_reads_as_the_reference_loop():
    o = Oracle(decks.small(ncx=4, ncy=3, ppc=2), "f32")
    import ctypes as C
    f = o.lib.orc_meta_omega
    f.argtypes = [C.c_float]
    f.restype = C.c_float
    assert f(0.1) == np.float32(1.8714) and f(200.0) == np.float32(1.0)
    for v in (0.3, 2.4, 7.0, 60.0):
        t = np.float32((np.float32(v) - np.float32(50.0)) / np.float32(50.0))
        want = np.float32(np.float32(1.0) * t + np.float32(1.0034) * (np.float32(1) - t))
        assert f(v) == want, v


def test_lines_are_the_runs_of_plasma_cells():
    from test_setup_parity import deck_variants
    d = deck_variants()["inner_conductor"]
    o = Oracle(d, "f64")
    ncx, ncy = int(o.scalar("ngx")) - 1, int(o.scalar("ngy")) - 1
    P = o.arr("PlasmaRegion").reshape(ncx, ncy) == 1
    assert not P.all(), "the deck should have non-plasma cells"
    runs_x = []
    for j in range(ncy):
        i = 0
        while i < ncx:
            if P[i, j]:
                i0 = i
                while i + 1 < ncx and P[i + 1, j]:
                    i += 1
                runs_x.append((i0, i, j))
            i += 1
    got = list(zip(o.arr("fl_x1"), o.arr("fl_x2"), o.arr("fl_xy")))
    assert got == runs_x
    runs_y = []
    for i in range(ncx):
        j = 0
        while j < ncy:
            if P[i, j]:
                j0 = j
                while j + 1 < ncy and P[i, j + 1]:
                    j += 1
                runs_y.append((i, j0, j))
            j += 1
    assert list(zip(o.arr("fl_yx"), o.arr("fl_y1"), o.arr("fl_y2"))) == runs_y


def closed_deck(ncx=24, ncy=10):
    """A CCP deck with no wall loss of metastables (quenching 0): every line
    end has a zero flux factor."""
    d = decks.small(ncx=ncx, ncy=ncy, ppc=2)
    d["GasSpecies"]["NeutralSpecies"]["Ar*"]["Quenching"] = 0.0
    return d


def test_zero_flux_solve_conserves_and_keeps_uniform():
    o = Oracle(closed_deck(), "f64")
    assert np.all(o.arr("fl_xg1") == 0) and np.all(o.arr("fl_yg2") == 0)
    den = o.arr("fl_den")
    n0 = den.copy()
    assert np.all(n0 > 1e9)
    for _ in range(5):
        o.orc_fluid_continuity()
    assert np.allclose(den, n0, rtol=1e-12), "a uniform density moved"
    rng = np.random.default_rng(2)
    den[:] = n0 * (1 + 0.5 * rng.uniform(-1, 1, den.size))
    tot = den.sum()
    spread = den.std()
    for _ in range(20):
        o.orc_fluid_continuity()
    assert abs(den.sum() - tot) <= 1e-12 * tot, "total density not conserved"
    assert den.std() < 0.5 * spread, "diffusion did not smooth"


@pytest.mark.parametrize("k", [1, 3])
def test_cosine_mode_decays_at_the_alternating_direction_rate(k):
    ncx, ncy = 24, 10
    o = Oracle(closed_deck(ncx, ncy), "f64")
    # uniform coefficients on every interior face (the deck's temperature
    # field varies, and with it D); the zero-flux line ends stay as set up
    Dc = float(np.median(o.arr("fl_D")))
    dx, dy = o.scalar("dx"), o.scalar("dy")
    ax, bx = o.arr("fl_ax").reshape(ncx + 1, ncy), o.arr("fl_bx").reshape(ncx + 1, ncy)
    ay, by = o.arr("fl_ay").reshape(ncx, ncy + 1), o.arr("fl_by").reshape(ncx, ncy + 1)
    assert np.all(bx[0] == 0) and np.all(ax[ncx] == 0) and np.all(by[:, 0] == 0) and np.all(ay[:, ncy] == 0)
    ax[1:ncx] = bx[1:ncx] = Dc / dx
    ay[:, 1:ncy] = by[:, 1:ncy] = Dc / dy
    a = Dc * o.scalar("dtc") / dx ** 2
    den = o.arr("fl_den")
    base = den[0]
    i = np.arange(ncx)
    mode = np.cos(np.pi * k * (i + 0.5) / ncx)
    amp = 0.2 * base
    den[:] = (base + amp * mode[:, None] * np.ones(ncy)[None, :]).ravel()
    o.orc_fluid_continuity()
    got = den.reshape(ncx, ncy)
    # the x sweep is implicit in x; the y sweep takes the x fluxes of its
    # result explicitly (src/cuda_Fluid.cu:131-163): per call the mode is
    # multiplied by (1 - lam) / (1 + lam), the Cayley factor of 2 dtc
    lam = 2 * a * (1 - np.cos(np.pi * k / ncx))
    want = base + amp * mode * (1 - lam) / (1 + lam)
    assert np.allclose(got, want[:, None], rtol=1e-10, atol=0)


def test_source_window_and_rates():
    d = decks.small(ncx=8, ncy=6, ppc=2)
    d["SimulationMethod"]["AverageIter"] = 2
    o = Oracle(d, "f64")
    G = int(o.scalar("ngx") * o.scalar("ngy"))
    T = int(o.scalar("TnRct"))
    den = o.arr("den")
    mccr = o.arr("mccr").reshape(G, T)
    p2_setup = o.arr("P2").copy()
    fn = o.arr("FluidN")[:G].copy()
    rng = np.random.default_rng(4)
    dens, before = [], mccr.copy()
    for step in range(3):            # two accumulations, then the average
        den[:G] = rng.uniform(0, 50, G)
        dens.append(den[:G].copy())
        mccr += rng.integers(0, 4, size=mccr.shape)
        o.orc_fluid_diag()
        assert int(o.scalar("conti_flag")) == (1 if step == 2 else 0)
    n_ave, dt, ds = 2, o.scalar("dt"), o.scalar("Denscale0")
    rate = ds / n_ave / dt
    cnt = (mccr - before).astype(np.float64)
    ave_e = ds / n_ave * np.sum(dens, axis=0)
    # rate constants of reactions 7 and 9 (src/PhysicsData.c:1715-1717)
    k7 = 2.0e-13 * 0.25 * ave_e * fn
    k9 = 3.0e-21 * o.arr("BackDen1") * fn
    S = rate * cnt[:, 2] - rate * cnt[:, 4] - k7 - 2 * rate * cnt[:, 8] - k9
    ngy = int(o.scalar("ngy"))
    cells = [(g // ngy) * 6 + g % ngy for g in range(G) if g // ngy < 8 and g % ngy < 6]
    nodes = [g for g in range(G) if g // ngy < 8 and g % ngy < 6]
    assert np.allclose(o.arr("fl_src")[cells], S[nodes], rtol=1e-12, atol=1e-3)
    assert int(o.scalar("nave_count")) == 0
    assert np.all(o.arr("fl_win") == 0)
    # the node density did not move yet (no continuity solve has run): P2 stays the set-up's
    assert np.array_equal(o.arr("P2"), p2_setup)


# ------------------------------------------------------------ oxygen gases
# The reaction equations are the deck's Collision_On_Off keys ("N.lhs>rhs",
# hybrid-final_amd/csrc/coll_keys.inc; the reference reads the same keys,
# src/start.c:1036-1230).  Fluid species: O2 gas O2A O2B OP OD; Ar/O2 Ar*
# (the metastable) O2A O2B OP OD.  Naming in the keys: "O" in
# "e+O2B>e+2O" is ground-state O(3P); "Ar*m" is the metastable among the
# tabulated reactions and "Ar*" among the rate-constant ones (68 on).
FLUID = {1: ["O2A", "O2B", "OP", "OD"], 2: ["ARM", "O2A", "O2B", "OP", "OD"]}
PROJ = {1: {"e": 0, "O2^": 1, "O^": 2, "O-": 3}, 2: {"e": 0, "Ar^": 1, "O2^": 2, "O^": 3, "O-": 4}}
NTAB = {1: 58, 2: 68}
# reactions the reference averages without a collision counting them
# (src/cuda_Diagnostic.cu:586-588, 621-632), with the species whose Denscale
# it uses: 69 and 70 (O- projectiles) take O+'s, 81 (Ar* + Ar*) the electron's
LATE = {1: {58: 3}, 2: {69: 3, 70: 3, 79: 1, 80: 1, 81: 0}}
# where the reference's sums differ from the equations: its Ar/O2 O2A loss
# lists reaction 30 (e+O2B>e+O2) where reaction 63 (O^+O2A>O2^+OP) belongs
# (src/cuda_Diagnostic.cu:722-724)
QUIRK = {1: {}, 2: {("O2A", 30): -1, ("O2A", 63): 0}}


def reactions(gas):
    import os
    import re
    path = os.path.join(os.path.dirname(__file__), "..", "hybrid-final_amd", "csrc", "coll_keys.inc")
    text = open(path).read()
    body = re.search(r"pic_coll_keys%d\[\d+\] = \{(.*?)\};" % gas, text).group(1)
    keys = re.findall(r'"([^"]*)"', body)
    out = []
    for r, k in enumerate(keys):
        num, eq = k.split(".", 1)
        assert int(num) == r
        lhs, rhs = eq.strip().split(">")

        def side(t):
            cnt = {}
            for tok in t.split("+"):
                m = re.match(r"(\d*)(.*)$", tok)
                n, name = int(m.group(1) or 1), m.group(2)
                if name == "O":
                    name = "OP"
                if name == "Ar*m" or (name == "Ar*" and r >= NTAB[gas]):
                    name = "ARM"
                if name:
                    cnt[name] = cnt.get(name, 0) + n
            return cnt
        # the projectile: the charged species on the left ("Ar+Ar^>..." too)
        proj = next((t for t in lhs.split("+") if t in PROJ[gas]), None)
        out.append((proj, side(lhs), side(rhs)))
    return out


def net(gas, r, f, eqs):
    if (f, r) in QUIRK[gas]:
        return QUIRK[gas][(f, r)]
    _, l, rt = eqs[r]
    return rt.get(f, 0) - l.get(f, 0)


def oxy_oracle(gas):
    d = decks.oxygen_deck(gas=gas, ncx=6, ncy=4, ppc=2)
    d["SimulationMethod"]["AverageIter"] = 1
    return Oracle(d, "f64")


@pytest.mark.parametrize("gas", [1, 2], ids=["O2", "ArO2"])
def test_oxygen_sources_follow_the_reaction_equations(gas):
    """One reaction's events at a time, fluid densities zero: every fluid
    species' source is (its net change in the equation) x (events) x the
    Denscale of the projectile species / (N_ave dt); reactions with a rate
    constant ignore counted events."""
    o = oxy_oracle(gas)
    eqs = reactions(gas)
    G, T = int(o.scalar("ngx") * o.scalar("ngy")), int(o.scalar("TnRct"))
    assert len(eqs) == T
    ncx, ncy, ngy = int(o.scalar("ngx")) - 1, int(o.scalar("ngy")) - 1, int(o.scalar("ngy"))
    nf = len(FLUID[gas])
    ds, dt = o.arr("Denscale"), o.scalar("dt")
    assert len(set(ds.tolist())) > 2, "the species' Denscale should differ"
    mccr = o.arr("mccr").reshape(G, T)
    src = o.arr("fl_src").reshape(nf, ncx, ncy)
    o.arr("fl_den")[:] = 0
    o.arr("FluidN")[:] = 0
    o.orc_fluid_diag()                    # the window's accumulation step
    counted = set(range(NTAB[gas])) | set(LATE[gas])
    for r in range(T):
        mccr[:, r] += 3
        o.orc_fluid_diag()                # the averaging step
        proj = eqs[r][0]
        sp = LATE[gas].get(r, PROJ[gas].get(proj))
        for fi, f in enumerate(FLUID[gas]):
            want = net(gas, r, f, eqs) * 3 * ds[sp] / 1 / dt if r in counted else 0.0
            got = src[fi]
            assert np.allclose(got, want, rtol=1e-12, atol=0), (r, proj, f, got.ravel()[0], want)
        o.orc_fluid_diag()                # next window's accumulation


@pytest.mark.parametrize("gas", [1, 2], ids=["O2", "ArO2"])
def test_oxygen_rate_constant_sources(gas):
    """No counted events; random node fluid densities: each species' source
    is the sum over the rate-constant reactions of (net change) x RR x the
    densities of the equation's two reactants (fluid species, the O2 or Ar
    background, or 0.25 x the averaged electron density)."""
    o = oxy_oracle(gas)
    eqs = reactions(gas)
    G, T = int(o.scalar("ngx") * o.scalar("ngy")), int(o.scalar("TnRct"))
    ncx, ncy, ngy = int(o.scalar("ngx")) - 1, int(o.scalar("ngy")) - 1, int(o.scalar("ngy"))
    nf = len(FLUID[gas])
    RR, ds = o.arr("RR"), o.arr("Denscale")
    assert np.all(RR[NTAB[gas]:] > 0)
    rng = np.random.default_rng(9)
    den = o.arr("den")
    den[:G] = rng.uniform(0, 40, G)
    o.arr("fl_den")[:] = 0
    FN = o.arr("FluidN").reshape(nf, G)
    o.orc_fluid_diag()
    Fv = rng.uniform(1e14, 1e16, (nf, G))
    FN[:] = Fv
    o.orc_fluid_diag()
    eden = 0.25 * (ds[0] / 1 * (2 * den[:G]))
    bg = {"O2": o.arr("BackDen1") if gas == 1 else o.arr("BackDen2")}
    if gas == 2:
        bg["Ar"] = o.arr("BackDen1")
    src = o.arr("fl_src").reshape(nf, ncx, ncy)
    for fi, f in enumerate(FLUID[gas]):
        want = np.zeros(G)
        for r in range(NTAB[gas], T):
            if r in LATE[gas]:
                continue
            k = net(gas, r, f, eqs)
            if k == 0:
                continue
            fac = RR[r] * np.ones(G)
            for name, n in eqs[r][1].items():
                for _ in range(n):
                    fac = fac * (Fv[FLUID[gas].index(name)] if name in FLUID[gas] else eden if name == "e" else bg[name])
            want += k * fac
        nodes = [i * ngy + j for i in range(ncx) for j in range(ncy)]
        scale = np.max(np.abs(want)) + 1e-300
        assert np.all(np.abs(src[fi].ravel() - want[nodes]) <= 1e-12 * scale), f
