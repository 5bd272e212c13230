"""The C-ABI library loads on a machine without a GPU and exports every
function include/picmcc.h declares."""
import ctypes as C
import os
import re

from picmcc import LIBPATH, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    with open(os.path.join(ROOT, "include", "picmcc.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(\w+)\s*\([^;]*\)\s*;", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_reference_stage_names():
    names = declared_functions()
    for ref in ("main_cuda", "Move_cuda", "SortBounndary_cuda", "MCC_Ar_cuda", "Deposit_cuda", "Deposit_Basic",
                "PCG_SOLVER", "PCG_SOLVER_Laplace", "Efield_cuda_Basic", "start_cuda"):
        assert ref in names


def test_library_exports_every_declared_function():
    L = C.CDLL(LIBPATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_python_mirror_binds():
    L = lib()
    assert L.pic_array_names().decode().split()[0] == "Ex"
    import picmcc
    for m in ("advance_time", "field_solve", "efield", "move", "sort_boundary", "mcc", "deposit", "diagnostic",
              "step", "run_staged", "set_step", "info", "particles", "set_particles", "array", "set_array", "sync",
              "close", "eedf_sample", "stage_timing", "stage_times", "last_field_solve", "field_solver_kind",
              "field_solver_shape", "diag_sample", "diag_history",
              "dump_save", "dump_load", "fluid"):
        assert callable(getattr(picmcc.Case, m, None)), f"Case.{m} is missing"
    for m in ("set_rhs", "solve", "sync", "result", "close"):
        assert callable(getattr(picmcc.PoissonBatch, m, None)), f"PoissonBatch.{m} is missing"
    for m in ("step", "sync", "close"):
        assert callable(getattr(picmcc.Batch, m, None)), f"Batch.{m} is missing"
    assert callable(picmcc.pbatch_solve_noalloc)
