"""ORACLE - test infrastructure only.  FP64 direct solve of independent
5-point Poisson problems, the checker of the batched GPU solve.

The operator of one problem (unit spacing, nx x ny unknowns, unknown (a, b)
at a*ny + b) is A = Tx (x) I + I (x) Ty, with Tx, Ty the 1-D second
differences: interior rows [-1 2 -1], a Neumann end row [2 -2] (mirror
node), a Dirichlet end row [2 -1] (0 V beyond).  For uniform permittivity and
dx = dy this is the reference's field operator (src/load.c:121-245); the
tests check that against the restated setup.  The solve is scipy's sparse
LU (SuperLU) in FP64.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spl


def tridiag(n, left, right):
    """left/right: 0 Neumann, 1 Dirichlet."""
    if n == 1:
        return sp.csr_matrix(np.array([[float(left == 1) + float(right == 1)]]))
    T = sp.diags([-np.ones(n - 1), 2.0 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1], format="lil")
    if left == 0:
        T[0, 1] = -2.0
    if right == 0:
        T[n - 1, n - 2] = -2.0
    return T.tocsr()


def operator(nx, ny, bc):
    """bc = (left, right, bottom, top)."""
    Tx = tridiag(nx, bc[0], bc[1])
    Ty = tridiag(ny, bc[2], bc[3])
    return (sp.kron(Tx, sp.identity(ny)) + sp.kron(sp.identity(nx), Ty)).tocsr()


def solve(nx, ny, bc, b):
    return spl.spsolve(operator(nx, ny, bc).tocsc(), np.asarray(b, dtype=np.float64))


def rel_residual(nx, ny, bc, b, x):
    r = np.asarray(b, dtype=np.float64) - operator(nx, ny, bc) @ x
    nb = np.linalg.norm(b)
    return np.linalg.norm(r) / nb if nb > 0 else np.linalg.norm(r)
