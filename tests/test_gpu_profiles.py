"""Statistical parity of the FP32-refined path against the FP64 plain path.

The two precisions start from the same loaded particles and the same Philox
stream, but their trajectories separate after a few steps (FP32 vs FP64
rounding feeds back through the field), so they are compared the way the
physics is compared: time-averaged profiles over 10 RF cycles of the 64x64
argon CCP.  Criterion: the x-averaged electron and ion density and potential
profiles agree within 2 % of their peak, and the total ionisation count
within 5 %.
"""
import numpy as np
import pytest

from picmcc import Case, FP32_REFINED, FP64_PLAIN, decks

pytestmark = pytest.mark.gpu

CYCLES = 10


def run(prec):
    d = decks.config1()
    steps = CYCLES * d["SimulationMethod"]["TimeStep_PIC"]
    c = Case(d, prec)
    c.step(steps)
    c.sync()
    inf = c.info()
    ngx, ngy, nsp = inf.ngx, inf.ngy, inf.nsp
    den = c.array("sum_den").reshape(nsp, ngx, ngy) / steps
    phi = c.array("sum_phi").reshape(ngx, ngy) / steps
    mccr = c.array("mccr").reshape(ngx * ngy, inf.TnRct).sum(axis=0)
    c.close()
    return den, phi, mccr


def test_time_averaged_profiles_fp32_refined_vs_fp64_plain():
    den32, phi32, m32 = run(FP32_REFINED)
    den64, phi64, m64 = run(FP64_PLAIN)
    # x is the Neumann direction: average over it, compare the electrode-to-electrode profile
    for s, name in enumerate(("electron", "ion")):
        a, b = den32[s].mean(axis=0), den64[s].mean(axis=0)
        err = np.max(np.abs(a - b)) / np.max(np.abs(b))
        assert err <= 0.02, f"{name} density profile differs by {err:.3%} of its peak"
    a, b = phi32.mean(axis=0), phi64.mean(axis=0)
    err = np.max(np.abs(a - b)) / np.max(np.abs(b))
    assert err <= 0.02, f"potential profile differs by {err:.3%} of its peak"
    # ionisation events (reaction 3: e + Ar, reaction 4: e + Ar*)
    iz32, iz64 = int(m32[3] + m32[4]), int(m64[3] + m64[4])
    assert iz64 > 0
    assert abs(iz32 - iz64) <= 0.05 * iz64, f"ionisation counts {iz32} vs {iz64}"
