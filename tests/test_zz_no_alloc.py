"""The batched Poisson solve that allocates no device memory
(pic_pbatch_solve_noalloc) against the allocating entry points
(pic_pbatch_create / _solve / _result) on the same problems.

Both run the same kernels in the same order, so solutions and residuals
must agree bit for bit.  Two calls in a row, the second on a smaller batch,
show that the shared work space carries nothing over from the last call.
Device buffers come from the HIP runtime the library itself links.
"""
import ctypes as C

import numpy as np
import pytest

from picmcc import PicError, PoissonBatch, lib, pbatch_solve_noalloc, poisson_cases

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip():
    lib()   # libpicmcc.so first: the runtime below is the one it linked
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    return h


def device_array(h, n):
    p = C.c_void_p()
    assert h.hipMalloc(C.byref(p), C.c_size_t(max(1, n) * 8)) == 0
    return p


def solve_both(h, sizes, bcs, rhs, fill):
    pb = PoissonBatch(sizes, bcs)
    pb.set_rhs(rhs)
    pb.solve(1e-10, 3)
    x_ref, rel_ref, bad_ref = pb.result()
    pb.close()
    n = rhs.size
    b_dev, x_dev = device_array(h, n), device_array(h, n)
    try:
        assert h.hipMemcpy(b_dev, rhs.ctypes.data_as(C.c_void_p), n * 8, H2D) == 0
        junk = np.full(n, fill)   # the solve starts from x = 0 whatever x held
        assert h.hipMemcpy(x_dev, junk.ctypes.data_as(C.c_void_p), n * 8, H2D) == 0
        rel, bad = pbatch_solve_noalloc(sizes, bcs, b_dev.value, x_dev.value, 1e-10, 3)
        x = np.empty(n)
        assert h.hipMemcpy(x.ctypes.data_as(C.c_void_p), x_dev, n * 8, D2H) == 0
    finally:
        h.hipFree(b_dev)
        h.hipFree(x_dev)
    assert bad == bad_ref == 0
    assert np.array_equal(rel, rel_ref), "per-problem residuals differ"
    assert np.array_equal(x, x_ref), f"solutions differ at {np.count_nonzero(x != x_ref)} unknowns"
    assert np.all(rel <= 1e-10)


def test_no_alloc_solve_equals_allocating_solve_twice():
    h = hip()
    sizes, bcs, rhs = poisson_cases.make(6, nmin=24, nmax=300, seed=5)
    solve_both(h, sizes, bcs, rhs, 7.0)
    # second call, smaller batch: the work space holds the first call's data
    sizes2, bcs2, rhs2 = poisson_cases.make(2, nmin=16, nmax=90, seed=9)
    assert sum(a * b for a, b in sizes2) < sum(a * b for a, b in sizes)
    solve_both(h, sizes2, bcs2, rhs2, -3.0)


def test_no_alloc_solve_rejects_batches_beyond_its_work_space():
    with pytest.raises(PicError) as e:
        pbatch_solve_noalloc([(1500, 1500)], [(1, 1, 1, 1)], 1, 1)
    assert e.value.code == -3 and "work space" in str(e.value)
