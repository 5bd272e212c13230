"""GPU parity of each stage of the step against the CPU oracle.

Both sides start from the same deck, hence the same loaded particles; each
stage is then run on both and compared.  Integer/ordering results and the
FP32/FP64 arithmetic of push, walls and collisions must be bit-exact; node
density is compared within 1e-5 relative (different summation order); the
potential within the field-solve tolerance.  Dielectric surface hits are summed
on the GPU in 64-bit fixed point (2^-40 of a particle per unit, order-free) and
in particle order by the oracle, so the per-species surface charge and the
Sigma derived from it are compared within 1e-5 (FP32) / 1e-10 (FP64) relative.
"""
import numpy as np
import pytest

from picmcc import Case, FP32_REFINED, FP64_PLAIN, decks
from oracle import Oracle

pytestmark = pytest.mark.gpu

PRECS = [(FP32_REFINED, "f32"), (FP64_PLAIN, "f64")]
FIELDS = ("x", "y", "vx", "vy", "vz", "id", "cell_start")


def same_particles(case, orc, s, what):
    a = case.particles(s)
    b = orc.species(s)
    assert len(a["x"]) == len(b["x"]), f"{what}: species {s} count {len(a['x'])} != {len(b['x'])}"
    for k in FIELDS:
        if not np.array_equal(a[k], b[k]):
            bad = np.nonzero(a[k] != b[k])[0]
            raise AssertionError(f"{what}: species {s} field {k} differs at {len(bad)} places, first {bad[:5]}: "
                                 f"{a[k][bad[:3]]} vs {b[k][bad[:3]]}")


def close(a, b, tol, what):
    """|a - b| <= tol * (|b| + max|b|) elementwise: relative where the charge
    is large, absolute at the scale of the array where species cancel."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.max(np.abs(b)) if b.size else 0.0
    bad = np.abs(a - b) > tol * (np.abs(b) + scale)
    assert not np.any(bad), f"{what}: {np.count_nonzero(bad)} values differ, first at {np.nonzero(bad)[0][:5]}"


def deck_small():
    return decks.small(ncx=16, ncy=12, ppc=40, volts=150.0)


def deck_dense():
    """More than 256 particles per cell: several LDS chunks per cell in both
    passes of the fused step."""
    return decks.small(ncx=8, ncy=6, ppc=420, volts=200.0, pres_torr=0.3)


def _variant(name):
    from test_setup_parity import deck_variants
    d = deck_variants()[name]
    d["Geometry"]["Source"][0]["AC(V)"] = 150.0
    return d


def deck_dielectric():
    """Thick electrodes (conductor cells absorb) and a dielectric block
    (surface charge split onto nodes): the operator is not separable, so the
    FP32 field solve is the Jacobi PCG in FP64 refinement."""
    return _variant("thick_electrodes_dielectric")


def deck_inner_conductor():
    """A Dirichlet right wall (particles deleted there) and a grounded
    conductor inside the plasma."""
    return _variant("inner_conductor")


def deck_tall():
    """Taller than wide (6 x 20 cells): the top-right corner cell's diagonal
    route goes through the structure map's padded corner."""
    return _variant("tall")


def deck_anisotropic():
    """dx != dy (0.7 mm x 0.5 mm cells), a DC bias and a phase on the source."""
    return _variant("anisotropic")


DECKS = [deck_small, deck_dense, deck_dielectric, deck_inner_conductor, deck_tall, deck_anisotropic]
DECK_IDS = ["small", "dense", "dielectric", "inner_conductor", "tall", "anisotropic"]


@pytest.mark.parametrize("deck", DECKS, ids=DECK_IDS)
@pytest.mark.parametrize("prec,pn", PRECS)
def test_stage_by_stage(prec, pn, deck):
    stage_parity(deck(), prec, pn, 3)


def stage_parity(d, prec, pn, steps, field_solver=0):
    """`steps` timesteps of deck d, each stage run on the GPU and in the
    oracle from the same state and compared."""
    case = Case(d, prec, field_solver=field_solver)
    orc = Oracle(d, pn)
    tol = 1e-5 if pn == "f32" else 1e-10
    for s in range(2):
        same_particles(case, orc, s, "load")
    orc.orc_deposit()
    np.testing.assert_allclose(case.array("den"), orc.arr("den"), rtol=1e-5, atol=1e-4)
    for step in range(1, steps + 1):
        case.advance_time()
        orc.orc_set_tstep(step)
        # field solve: GPU PCG vs FP64 direct solve of the same rhs
        case.set_array("b", orc.arr("b"))
        case.field_solve()
        orc.orc_field()
        phi_g, phi_o = case.array("phi"), orc.arr("phi")
        scale = np.max(np.abs(phi_o)) + 1e-30
        assert np.max(np.abs(phi_g - phi_o)) <= 1e-6 * scale + 1e-12, "potential differs beyond solver tolerance"
        # the FP64 solution itself against the FP64 direct solve of the same rhs:
        # the separable refined solve reaches ~1e-13 relative residual, the plain
        # Jacobi PCG stops at 1e-10 relative residual (error <= cond(A) * 1e-10)
        x_g, x_o = case.array("xsol"), orc.arr("xsol")
        xs = np.max(np.abs(x_o)) + 1e-300
        bound = 1e-8 if case.field_solver_kind() == 1 else 1e-6
        assert np.max(np.abs(x_g - x_o)) <= bound * xs, f"FP64 solution error {np.max(np.abs(x_g - x_o)) / xs:.2e}"
        # from here on both use the oracle's potential, density and surface charge
        case.set_array("phi", orc.arr("phi"))
        case.set_array("den", orc.arr("den"))
        case.set_array("Sigma", orc.arr("Sigma"))
        case.efield()
        orc.orc_efield()
        for nm in ("TotPot", "Sigma", "Ex", "Ey"):
            assert np.array_equal(case.array(nm), orc.arr(nm)), f"{nm} differs at step {step}"
        case.move()
        orc.orc_move()
        for s in range(2):
            same_particles(case, orc, s, f"move step {step}")
        case.sort_boundary()
        assert orc.orc_sort() == 0
        for s in range(2):
            same_particles(case, orc, s, f"sort step {step}")
        # particles absorbed by conductors and deleted at walls, per species
        for nm in ("CondCharge", "Lost"):
            assert np.array_equal(case.array(nm).astype(np.int64), orc.arr(nm).astype(np.int64)), f"{nm} differs"
        # surface charge including this step's hits, which the deposit folds in
        sig_g = case.array("sigma_sp").astype(np.float64) + case.array("sigma_hits").astype(np.float64) * 2.0**-40
        close(sig_g, orc.arr("sigma_sp"), tol, f"sigma_sp after sort, step {step}")
        case.mcc()
        assert orc.orc_mcc() == 0
        for s in range(2):
            same_particles(case, orc, s, f"mcc step {step}")
        assert np.array_equal(case.array("mccr").astype(np.int64), orc.arr("mccr")), "collision event counts differ"
        assert np.array_equal(case.array("next_id")[:2].astype(np.int64),
                              [int(orc.scalar("next_id0")), int(orc.scalar("next_id1"))])
        case.deposit()
        orc.orc_deposit()
        np.testing.assert_allclose(case.array("den"), orc.arr("den"), rtol=1e-5, atol=1e-4)
        np.testing.assert_allclose(case.array("b"), orc.arr("b"), rtol=1e-5, atol=1e-3 * np.max(np.abs(orc.arr("b"))))
        assert not np.any(case.array("sigma_hits")), "the deposit leaves surface hits unfolded"
        close(case.array("sigma_sp"), orc.arr("sigma_sp"), tol, f"sigma_sp after deposit, step {step}")
        close(case.array("Sigma"), orc.arr("Sigma"), tol, f"Sigma after deposit, step {step}")
    case.sync()
    case.close()
    orc.close()


@pytest.mark.parametrize("deck", DECKS, ids=DECK_IDS)
@pytest.mark.parametrize("prec,pn", PRECS)
def test_fused_step_equals_staged(prec, pn, deck):
    """pic_step (fused two-pass) and the staged stage functions give the same bits."""
    d = deck()
    a = Case(d, prec)
    b = Case(d, prec)
    for step in range(1, 6):
        a.step(1)
        b.advance_time()
        b.field_solve()
        b.efield()
        b.move()
        b.sort_boundary()
        b.mcc()
        b.deposit()
        b.diagnostic()
        for s in range(2):
            pa, pb = a.particles(s), b.particles(s)
            for k in FIELDS:
                assert np.array_equal(pa[k], pb[k]), f"step {step} species {s} field {k}"
        assert np.array_equal(a.array("mccr"), b.array("mccr"))
    a.sync()
    b.sync()


def test_graph_steps_equal_eager_steps():
    """The fused step replayed from captured graphs (two buffer parities) gives
    the same particles, fields and counters as the same step launched kernel by
    kernel (stage timers on force the eager launches), checked after every
    replay, with host synchronisation and eager work in between."""
    d = deck_small()
    g = Case(d, FP32_REFINED)
    e = Case(d, FP32_REFINED)
    e.stage_timing(True)
    for rep in range(6):
        g.step(1)
        e.step(1)
        g.sync()
        e.sync()
        for s in range(2):
            a, b = g.particles(s), e.particles(s)
            for k in FIELDS:
                assert np.array_equal(a[k], b[k]), f"replay {rep}: species {s} field {k} differs"
        for nm in ("phi", "Ex", "Ey", "den", "sum_den", "mccr", "next_id"):
            assert np.array_equal(g.array(nm), e.array(nm)), f"replay {rep}: {nm} differs"
        assert g.info().tstep == e.info().tstep == rep + 1
    # staged work between graph replays moves the clock and the buffer parity
    g.run_staged(1)
    e.run_staged(1)
    g.step(2)
    e.step(2)
    for s in range(2):
        a, b = g.particles(s), e.particles(s)
        for k in FIELDS:
            assert np.array_equal(a[k], b[k]), f"after staged step: species {s} field {k} differs"


def test_fp64_with_the_separable_field_solve():
    """The FP64 path with the separable FP64 field solve (the solver of the
    FP32 path, chosen with field_solver = 2): stage by stage against the
    oracle, and its graph-replayed fused steps equal the same steps launched
    kernel by kernel."""
    d = deck_small()
    stage_parity(d, FP64_PLAIN, "f64", 3, field_solver=2)
    g = Case(d, FP64_PLAIN, field_solver=2)
    e = Case(d, FP64_PLAIN, field_solver=2)
    assert g.field_solver_kind() == 1
    e.stage_timing(True)
    for rep in range(4):
        g.step(1)
        e.step(1)
        g.sync()
        e.sync()
        for s in range(2):
            a, b = g.particles(s), e.particles(s)
            for k in FIELDS:
                assert np.array_equal(a[k], b[k]), f"replay {rep}: species {s} field {k} differs"
        for nm in ("phi", "Ex", "Ey", "den", "sum_den", "mccr"):
            assert np.array_equal(g.array(nm), e.array(nm)), f"replay {rep}: {nm} differs"
    g.close()
    e.close()
