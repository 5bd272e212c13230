"""Batched Poisson solves against the FP64 direct-solve oracle."""
import numpy as np
import pytest

from picmcc import PoissonBatch, poisson_cases
from poisson_oracle import rel_residual, solve

pytestmark = pytest.mark.gpu


def test_small_batch_matches_direct_solve():
    sizes, bcs, rhs = poisson_cases.make(12, nmin=3, nmax=40, seed=5)
    pb = PoissonBatch(sizes, bcs)
    pb.set_rhs(rhs)
    pb.solve(tol=1e-10, max_cycles=3)
    x, rel, bad = pb.result()
    pb.close()
    assert bad == 0
    off = 0
    for (nx, ny), bc in zip(sizes, bcs):
        n = nx * ny
        xr = solve(nx, ny, bc, rhs[off:off + n])
        err = np.max(np.abs(x[off:off + n] - xr)) / np.max(np.abs(xr))
        assert err <= 1e-8, f"problem {nx}x{ny} {bc}: relative error {err:.2e}"
        off += n
    assert np.all(rel <= 1e-10)


def test_full_size_batch_residuals():
    """64^2-512^2 problems: the residual of each solution, recomputed on the
    host with the oracle's operator, is within the 1e-10 tolerance."""
    sizes, bcs, rhs = poisson_cases.make(24, seed=9)
    pb = PoissonBatch(sizes, bcs)
    pb.set_rhs(rhs)
    pb.solve(tol=1e-10, max_cycles=3)
    x, rel, bad = pb.result()
    pb.close()
    assert bad == 0
    off = 0
    for (nx, ny), bc in zip(sizes, bcs):
        n = nx * ny
        r = rel_residual(nx, ny, bc, rhs[off:off + n], x[off:off + n])
        assert r <= 1e-10, f"problem {nx}x{ny} {bc}: host residual {r:.2e}"
        off += n
