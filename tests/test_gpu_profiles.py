"""Statistical parity of the FP32-refined path against the FP64 plain path.

The two precisions start from the same loaded particles and the same Philox
stream, but their trajectories separate after a few steps (FP32 vs FP64
rounding feeds back through the field), so they are compared the way the
physics is compared: time-averaged profiles over 10 RF cycles of the 64x64
argon CCP.  Criteria: the x-averaged electron and ion density and potential
profiles and the electron energy distribution (sampled every 40 steps) agree
within 2 % of their peak, the mean electron energy within 2 %, and the total
ionisation count within 5 %.
"""
import numpy as np
import pytest

from picmcc import Case, FP32_REFINED, FP64_PLAIN, decks

pytestmark = pytest.mark.gpu

CYCLES = 10


EEDF_EVERY = 40
EMAX = 40.0


def run(prec):
    d = decks.config1()
    steps = CYCLES * d["SimulationMethod"]["TimeStep_PIC"]
    c = Case(d, prec)
    for _ in range(steps // EEDF_EVERY):
        c.step(EEDF_EVERY)
        c.eedf_sample(EMAX)
    c.sync()
    eedf = c.array("eedf").astype(np.float64)
    inf = c.info()
    ngx, ngy, nsp = inf.ngx, inf.ngy, inf.nsp
    den = c.array("sum_den").reshape(nsp, ngx, ngy) / steps
    phi = c.array("sum_phi").reshape(ngx, ngy) / steps
    mccr = c.array("mccr").reshape(ngx * ngy, inf.TnRct).sum(axis=0)
    c.close()
    return den, phi, mccr, eedf


def test_time_averaged_profiles_fp32_refined_vs_fp64_plain():
    den32, phi32, m32, f32 = run(FP32_REFINED)
    den64, phi64, m64, f64 = run(FP64_PLAIN)
    # x is the Neumann direction: average over it, compare the electrode-to-electrode profile
    for s, name in enumerate(("electron", "ion")):
        a, b = den32[s].mean(axis=0), den64[s].mean(axis=0)
        err = np.max(np.abs(a - b)) / np.max(np.abs(b))
        assert err <= 0.02, f"{name} density profile differs by {err:.3%} of its peak"
    a, b = phi32.mean(axis=0), phi64.mean(axis=0)
    err = np.max(np.abs(a - b)) / np.max(np.abs(b))
    assert err <= 0.02, f"potential profile differs by {err:.3%} of its peak"
    # electron energy distribution sampled every 40 steps (256 bins over 0-40 eV)
    p32, p64 = f32 / f32.sum(), f64 / f64.sum()
    err = np.max(np.abs(p32 - p64)) / np.max(p64)
    assert err <= 0.02, f"EEDF differs by {err:.3%} of its peak"
    centers = (np.arange(len(p64)) + 0.5) * EMAX / len(p64)
    m32e, m64e = float((p32 * centers).sum()), float((p64 * centers).sum())
    assert abs(m32e - m64e) <= 0.02 * m64e, f"mean electron energy {m32e:.4f} vs {m64e:.4f} eV"
    # ionisation events (reaction 3: e + Ar, reaction 4: e + Ar*)
    iz32, iz64 = int(m32[3] + m32[4]), int(m64[3] + m64[4])
    assert iz64 > 0
    assert abs(iz32 - iz64) <= 0.05 * iz64, f"ionisation counts {iz32} vs {iz64}"
