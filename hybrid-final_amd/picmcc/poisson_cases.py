"""Batched Poisson workload: independent 5-point problems with Gaussian
charge blobs.

Grid sizes log-uniform per side over [nmin, nmax]; right-hand side
rho dx^2 / eps0 with rho = 1-8 Gaussian blobs (amplitude N(0, 1e-3) C/m^3,
width 2-20 cells) plus uniform noise; boundary sides drawn from four
patterns, each with at least one Dirichlet side.
"""
import math
import random

import numpy as np

EPS0 = 8.8542e-12
BC_PATTERNS = [(0, 0, 1, 1), (1, 1, 1, 1), (1, 0, 0, 0), (0, 1, 1, 0)]   # (left, right, bottom, top)


def make(nprob, nmin=64, nmax=512, seed=11, dx=2e-4):
    rng = random.Random(seed)
    nrng = np.random.default_rng(seed)
    sizes, bcs, rhs = [], [], []
    for _ in range(nprob):
        nx = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        ny = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        sizes.append((nx, ny))
        bcs.append(BC_PATTERNS[rng.randrange(len(BC_PATTERNS))])
        a = np.arange(nx)[:, None]
        b = np.arange(ny)[None, :]
        rho = np.zeros((nx, ny))
        for _ in range(rng.randint(1, 8)):
            ca, cb = rng.uniform(0, nx), rng.uniform(0, ny)
            w = rng.uniform(2, 20)
            rho += nrng.normal(0.0, 1e-3) * np.exp(-((a - ca) ** 2 + (b - cb) ** 2) / (2 * w * w))
        rho += nrng.uniform(-1e-5, 1e-5, size=(nx, ny))
        rhs.append((rho * dx * dx / EPS0).ravel())
    return sizes, bcs, np.concatenate(rhs)
