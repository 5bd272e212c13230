"""Python host mirror of the MI355X PIC-MCC library (libpicmcc.so).

A thin ctypes layer over include/picmcc.h.  ``Case`` is one discharge case
resident on the GPU (the reference's stage table, src/cuda_main.cu:112-235,
as methods); ``HostSetup`` is the host-side setup alone (no device needed).
There is no CPU fallback: if the HIP library is missing the import of the
compute path raises.
"""
import ctypes as C
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libpicmcc.so")

FP32_REFINED = 0
FP64_PLAIN = 1

_lib = None


class PicError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"picmcc error {code}: {msg}")
        self.code = code


class Options(C.Structure):
    _fields_ = [("precision", C.c_int), ("device", C.c_int), ("inner_tol", C.c_double),
                ("pcg_maxit", C.c_int), ("load", C.c_int), ("field_solver", C.c_int)]


class Info(C.Structure):
    _fields_ = [("ngx", C.c_int), ("ngy", C.c_int), ("nsp", C.c_int), ("A_size", C.c_int),
                ("CondNUMR", C.c_int), ("TnRct", C.c_int), ("MCCn", C.c_int), ("nvel", C.c_int),
                ("dt", C.c_double), ("dx", C.c_double), ("dy", C.c_double), ("t", C.c_double),
                ("tstep", C.c_longlong), ("precision", C.c_int),
                ("np", C.c_longlong * 5), ("cap", C.c_longlong * 5)]


EXPORTS = {
    "pic_create": (C.c_int, [C.c_char_p, C.POINTER(Options), C.POINTER(C.c_void_p)]),
    "pic_create_from_file": (C.c_int, [C.c_char_p, C.POINTER(Options), C.POINTER(C.c_void_p)]),
    "pic_destroy": (C.c_int, [C.c_void_p]),
    "pic_last_error": (C.c_char_p, []),
    "pic_info_get": (C.c_int, [C.c_void_p, C.POINTER(Info)]),
    "pic_sync": (C.c_int, [C.c_void_p]),
    "pic_query": (C.c_int, [C.c_void_p]),
    "pic_device_memory": (C.c_int, [C.c_int, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "pic_device_sync": (C.c_int, [C.c_int]),
    "pic_eedf_sample": (C.c_int, [C.c_void_p, C.c_double]),
    "pic_advance_time": (C.c_int, [C.c_void_p]),
    "pic_field_solve": (C.c_int, [C.c_void_p]),
    "pic_efield": (C.c_int, [C.c_void_p]),
    "pic_move": (C.c_int, [C.c_void_p]),
    "pic_sort_boundary": (C.c_int, [C.c_void_p]),
    "pic_mcc": (C.c_int, [C.c_void_p]),
    "pic_deposit": (C.c_int, [C.c_void_p]),
    "pic_diagnostic": (C.c_int, [C.c_void_p]),
    "pic_run_staged": (C.c_int, [C.c_void_p, C.c_int]),
    "pic_step": (C.c_int, [C.c_void_p, C.c_int]),
    "pic_get_particles": (C.c_int, [C.c_void_p, C.c_int] + [C.c_void_p] * 7),
    "pic_set_particles": (C.c_int, [C.c_void_p, C.c_int, C.c_longlong] + [C.c_void_p] * 7 + [C.c_uint32]),
    "pic_array_info": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_size_t), C.POINTER(C.c_int),
                                 C.POINTER(C.c_int)]),
    "pic_get_array": (C.c_int, [C.c_void_p, C.c_char_p, C.c_void_p, C.c_size_t]),
    "pic_set_array": (C.c_int, [C.c_void_p, C.c_char_p, C.c_void_p, C.c_size_t]),
    "pic_array_names": (C.c_char_p, []),
    "pic_set_step": (C.c_int, [C.c_void_p, C.c_longlong]),
    "pic_last_field_solve": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                       C.POINTER(C.c_double)]),
    "pic_field_solver_kind": (C.c_int, [C.c_void_p]),
    "pic_field_solver_shape": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "pic_diag_sample": (C.c_int, [C.c_void_p]),
    "pic_dump_save": (C.c_int, [C.c_void_p, C.c_char_p]),
    "pic_fluid_enable": (C.c_int, [C.c_void_p, C.c_int]),
    "pic_batch_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p)]),
    "pic_batch_step": (C.c_int, [C.c_void_p, C.c_int]),
    "pic_batch_sync": (C.c_int, [C.c_void_p]),
    "pic_batch_destroy": (C.c_int, [C.c_void_p]),
    "pic_dump_load": (C.c_int, [C.c_void_p, C.c_char_p]),
    "pic_diag_history": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "pic_stage_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "pic_stage_times": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_longlong)]),
    "pic_pbatch_create": (C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]),
    "pic_pbatch_destroy": (C.c_int, [C.c_void_p]),
    "pic_pbatch_unknowns": (C.c_longlong, [C.c_void_p]),
    "pic_pbatch_set_rhs": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pic_pbatch_solve": (C.c_int, [C.c_void_p, C.c_double, C.c_int]),
    "pic_pbatch_sync": (C.c_int, [C.c_void_p]),
    "pic_pbatch_result": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "pic_pbatch_last_error": (C.c_char_p, []),
    "pic_pbatch_solve_noalloc": (C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                           C.c_void_p, C.c_double, C.c_int, C.c_void_p, C.c_void_p]),
    "pic_setup_open": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "pic_setup_array_info": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_size_t), C.POINTER(C.c_int),
                                       C.POINTER(C.c_int)]),
    "pic_setup_get_array": (C.c_int, [C.c_void_p, C.c_char_p, C.c_void_p, C.c_size_t]),
    "pic_setup_scalar": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_double)]),
    "pic_setup_close": (None, [C.c_void_p]),
    "pic_reaction_names": (C.c_char_p, [C.c_int]),
    "pic_shim_open": (C.c_int, [C.c_char_p, C.c_int]),
    "pic_shim_close": (None, []),
    "pic_shim_ctx": (C.c_void_p, []),
}
SHIMS = ["main_cuda", "start_cuda", "info_Device", "Set_Device_Parameter", "Set_Particle_cuda",
         "Set_NullCollisionTime_cuda", "Set_SortBoundary_cuda", "Set_MatrixPCG_cuda", "PCG_SOLVER_Laplace",
         "PCG_SOLVER", "Efield_cuda_Basic", "Move_cuda", "SortBounndary_cuda", "MCC_Ar_cuda", "MCC_O2_cuda",
         "MCC_ArO2_cuda", "Deposit_cuda", "Deposit_Basic", "Diagnostic"]


def lib():
    """Load libpicmcc.so (raises if it is not built: there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"{LIBPATH} is not built; run `make -C hybrid-final_amd` or __graft_entry__.build()")
        L = C.CDLL(LIBPATH)
        for name, (res, args) in EXPORTS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        for name in SHIMS:
            getattr(L, name).restype = None
            getattr(L, name).argtypes = []
        _lib = L
    return _lib


def _check(code):
    if code != 0:
        raise PicError(code, lib().pic_last_error().decode())
    return code


def device_sync(device=0):
    """Wait for all work on the device; raises PicError on a device error."""
    _check(lib().pic_device_sync(device))


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


_INT_ARRAYS = {"Boundary", "CondID", "Face", "DensRegion", "PlasmaRegion", "StructIdx", "A_idx", "Ai", "Aj",
               "CondCharge", "next_id", "Lost"}


def _dtype(name, elsize, is_float):
    if name == "mccr":
        return np.uint64
    if name == "next_id" or name.endswith(".id"):
        return np.uint32
    if not is_float:
        return np.int32 if elsize == 4 else np.int64
    return np.float32 if elsize == 4 else np.float64


def deck_text(deck):
    return deck if isinstance(deck, str) else json.dumps(deck)


class HostSetup:
    """Host-side case setup only (geometry, operator, tables, initial load)."""

    def __init__(self, deck, precision=FP32_REFINED, load=True):
        L = lib()
        h = C.c_void_p()
        _check(L.pic_setup_open(deck_text(deck).encode(), precision, int(load), C.byref(h)))
        self.h = h
        self.precision = precision

    def array(self, name):
        L = lib()
        n, es, fl = C.c_size_t(), C.c_int(), C.c_int()
        _check(L.pic_setup_array_info(self.h, name.encode(), C.byref(n), C.byref(es), C.byref(fl)))
        out = np.empty(n.value, dtype=_dtype(name, es.value, fl.value))
        _check(L.pic_setup_get_array(self.h, name.encode(), _ptr(out), out.nbytes))
        return out

    def scalar(self, name):
        v = C.c_double()
        _check(lib().pic_setup_scalar(self.h, name.encode(), C.byref(v)))
        return v.value

    def close(self):
        if self.h:
            lib().pic_setup_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Case:
    """One discharge case resident on the GPU."""

    def __init__(self, deck, precision=FP32_REFINED, device=0, inner_tol=0.0, pcg_maxit=0, load=True,
                 field_solver=0):
        L = lib()
        o = Options(precision, device, inner_tol, pcg_maxit, int(load), field_solver)
        h = C.c_void_p()
        _check(L.pic_create(deck_text(deck).encode(), C.byref(o), C.byref(h)))
        self.h = h
        self.precision = precision
        self.real = np.float64 if precision == FP64_PLAIN else np.float32

    # ---- lifecycle
    def close(self):
        if getattr(self, "h", None):
            _check(lib().pic_destroy(self.h))   # refused while the case is in a batch
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        i = Info()
        _check(lib().pic_info_get(self.h, C.byref(i)))
        return i

    def sync(self):
        _check(lib().pic_sync(self.h))

    def eedf_sample(self, emax_ev=40.0):
        """Add the current electron energy histogram to the "eedf" array."""
        _check(lib().pic_eedf_sample(self.h, float(emax_ev)))

    def done(self):
        """True when all enqueued work of this case has finished (non-blocking)."""
        r = lib().pic_query(self.h)
        if r < 0:
            _check(r)
        return r == 1

    # ---- stage table
    def advance_time(self): _check(lib().pic_advance_time(self.h))
    def field_solve(self): _check(lib().pic_field_solve(self.h))
    def efield(self): _check(lib().pic_efield(self.h))
    def move(self): _check(lib().pic_move(self.h))
    def sort_boundary(self): _check(lib().pic_sort_boundary(self.h))
    def mcc(self): _check(lib().pic_mcc(self.h))
    def deposit(self): _check(lib().pic_deposit(self.h))
    def diagnostic(self): _check(lib().pic_diagnostic(self.h))

    def run_staged(self, n=1):
        _check(lib().pic_run_staged(self.h, n))

    def step(self, n=1):
        _check(lib().pic_step(self.h, n))

    def set_step(self, tstep):
        _check(lib().pic_set_step(self.h, tstep))

    STAGES = ("field", "efield", "count_e", "count_i", "bases", "scatter_e", "scatter_i", "deposit", "accumulate")

    def stage_timing(self, enable=True):
        _check(lib().pic_stage_timing(self.h, int(enable)))

    def stage_times(self):
        ms = (C.c_double * len(self.STAGES))()
        n = C.c_longlong()
        _check(lib().pic_stage_times(self.h, ms, len(self.STAGES), C.byref(n)))
        return dict(zip(self.STAGES, list(ms))), n.value

    def fluid(self, on=True):
        """Switches the neutral fluid coupling on or off (Ar*; O2A, O2B, OP,
        OD for O2; Ar* and those four for Ar/O2)."""
        _check(lib().pic_fluid_enable(self.h, int(bool(on))))

    def dump_save(self, path):
        """Writes the case's state in the reference's V000 dump format."""
        _check(lib().pic_dump_save(self.h, os.fsencode(path)))

    def dump_load(self, path):
        """Resumes the case from a V000 dump (same deck: gas, grid and np2c)."""
        _check(lib().pic_dump_load(self.h, os.fsencode(path)))

    def diag_sample(self):
        """Appends a diagnostics record of the last completed step (no host synchronisation)."""
        _check(lib().pic_diag_sample(self.h))

    def diag_history(self):
        """The retained diagnostics records, oldest first, as a dict of arrays:
        t, step, np (records x nsp), ke (J, records x nsp), q_cond (C, records x
        ncond), absorbed (records x nsp x ncond), v_cond (records x ncond)."""
        L = lib()
        n, ln = C.c_int(), C.c_int()
        _check(L.pic_diag_history(self.h, None, 1 << 30, C.byref(n), C.byref(ln)))
        buf = np.empty((max(n.value, 1), ln.value), dtype=np.float64)
        _check(L.pic_diag_history(self.h, _ptr(buf), n.value, C.byref(n), C.byref(ln)))
        buf = buf[:n.value]
        inf = self.info()
        nsp, nc = inf.nsp, inf.CondNUMR
        o = 2
        out = {"t": buf[:, 0], "step": buf[:, 1].astype(np.int64)}
        out["np"] = buf[:, o:o + nsp].astype(np.int64); o += nsp
        out["ke"] = buf[:, o:o + nsp]; o += nsp
        out["q_cond"] = buf[:, o:o + nc]; o += nc
        out["absorbed"] = buf[:, o:o + nsp * nc].reshape(-1, nsp, nc); o += nsp * nc
        out["v_cond"] = buf[:, o:o + nc]
        return out

    def field_solver_shape(self):
        """(blocks, threads per block) of the persistent PCG; (0, 0) with the separable solve."""
        b, t = C.c_int(), C.c_int()
        _check(lib().pic_field_solver_shape(self.h, C.byref(b), C.byref(t)))
        return b.value, t.value

    def field_solver_kind(self):
        return lib().pic_field_solver_kind(self.h)

    def last_field_solve(self):
        it, out, rr = C.c_int(), C.c_int(), C.c_double()
        _check(lib().pic_last_field_solve(self.h, C.byref(it), C.byref(out), C.byref(rr)))
        return it.value, out.value, rr.value

    # ---- state
    def particles(self, isp):
        inf = self.info()
        n = inf.np[isp]
        G = inf.ngx * inf.ngy
        out = {k: np.empty(n, dtype=self.real) for k in ("x", "y", "vx", "vy", "vz")}
        out["id"] = np.empty(n, dtype=np.uint32)
        out["cell_start"] = np.empty(G + 1, dtype=np.int32)
        _check(lib().pic_get_particles(self.h, isp, _ptr(out["x"]), _ptr(out["y"]), _ptr(out["vx"]),
                                       _ptr(out["vy"]), _ptr(out["vz"]), _ptr(out["id"]),
                                       _ptr(out["cell_start"])))
        return out

    def set_particles(self, isp, st, next_id):
        arrs = [np.ascontiguousarray(st[k], dtype=self.real) for k in ("x", "y", "vx", "vy", "vz")]
        ids = np.ascontiguousarray(st["id"], dtype=np.uint32)
        cs = np.ascontiguousarray(st["cell_start"], dtype=np.int32)
        _check(lib().pic_set_particles(self.h, isp, len(arrs[0]), *[_ptr(a) for a in arrs], _ptr(ids), _ptr(cs),
                                       int(next_id)))

    def array(self, name):
        L = lib()
        n, es, fl = C.c_size_t(), C.c_int(), C.c_int()
        _check(L.pic_array_info(self.h, name.encode(), C.byref(n), C.byref(es), C.byref(fl)))
        out = np.empty(n.value, dtype=_dtype(name, es.value, fl.value))
        _check(L.pic_get_array(self.h, name.encode(), _ptr(out), out.nbytes))
        return out

    def set_array(self, name, values):
        L = lib()
        n, es, fl = C.c_size_t(), C.c_int(), C.c_int()
        _check(L.pic_array_info(self.h, name.encode(), C.byref(n), C.byref(es), C.byref(fl)))
        a = np.ascontiguousarray(values, dtype=_dtype(name, es.value, fl.value))
        _check(L.pic_set_array(self.h, name.encode(), _ptr(a), a.nbytes))


class Batch:
    """Several cases stepped together by shared launches (pic_batch_*): one
    graph launch per step for all of them.  The cases keep working as
    before; stepping one outside the batch invalidates the batch."""

    def __init__(self, cases):
        self.cases = list(cases)
        arr = (C.c_void_p * len(self.cases))(*[c.h for c in self.cases])
        h = C.c_void_p()
        _check(lib().pic_batch_create(arr, len(self.cases), C.byref(h)))
        self.h = h.value

    def step(self, n=1):
        _check(lib().pic_batch_step(self.h, int(n)))

    def sync(self):
        _check(lib().pic_batch_sync(self.h))

    def close(self):
        if getattr(self, "h", None):
            lib().pic_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PoissonBatch:
    """Independent 5-point Poisson problems solved together on the GPU
    (include/picmcc.h, pic_pbatch_*).  sizes: [(nx, ny)], bcs: [(left, right,
    bottom, top)] with 0 = Neumann, 1 = Dirichlet."""

    def __init__(self, sizes, bcs, device=0):
        L = lib()
        n = len(sizes)
        self.sizes = [(int(a), int(b)) for a, b in sizes]
        nx = np.array([s[0] for s in self.sizes], dtype=np.int32)
        ny = np.array([s[1] for s in self.sizes], dtype=np.int32)
        bc = np.ascontiguousarray(np.asarray(bcs, dtype=np.int32).reshape(n, 4))
        h = C.c_void_p()
        rc = L.pic_pbatch_create(n, _ptr(nx), _ptr(ny), _ptr(bc), device, C.byref(h))
        if rc != 0:
            raise PicError(rc, L.pic_pbatch_last_error().decode())
        self.h = h
        self.n = n
        self.N = int(L.pic_pbatch_unknowns(self.h))
        self.offsets = np.concatenate([[0], np.cumsum(nx.astype(np.int64) * ny)])

    def _chk(self, rc):
        if rc < 0:
            raise PicError(rc, lib().pic_pbatch_last_error().decode())
        return rc

    def set_rhs(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        assert b.size == self.N
        self._chk(lib().pic_pbatch_set_rhs(self.h, _ptr(b)))

    def solve(self, tol=1e-10, max_cycles=3):
        self._chk(lib().pic_pbatch_solve(self.h, float(tol), int(max_cycles)))

    def sync(self):
        self._chk(lib().pic_pbatch_sync(self.h))

    def result(self):
        x = np.empty(self.N, dtype=np.float64)
        rel = np.empty(self.n, dtype=np.float64)
        bad = self._chk(lib().pic_pbatch_result(self.h, _ptr(x), _ptr(rel)))
        return x, rel, bad

    def close(self):
        if getattr(self, "h", None):
            lib().pic_pbatch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pbatch_solve_noalloc(sizes, bcs, b_dev, x_dev, tol=1e-10, max_cycles=3, device=0, stream=None):
    """The batched solve of PoissonBatch in one call that allocates no device
    memory (include/picmcc.h, pic_pbatch_solve_noalloc): b_dev and x_dev are
    device addresses (ints) of the right-hand sides and the solution, laid
    out as for PoissonBatch.  Returns (per-problem relative residuals, number
    of problems above tol)."""
    n = len(sizes)
    nx = np.array([int(a) for a, _ in sizes], dtype=np.int32)
    ny = np.array([int(b) for _, b in sizes], dtype=np.int32)
    bc = np.ascontiguousarray(np.asarray(bcs, dtype=np.int32).reshape(n, 4))
    rel = np.empty(n, dtype=np.float64)
    L = lib()
    rc = L.pic_pbatch_solve_noalloc(n, _ptr(nx), _ptr(ny), _ptr(bc), int(device), C.c_void_p(int(b_dev)),
                                    C.c_void_p(int(x_dev)), float(tol), int(max_cycles), _ptr(rel),
                                    C.c_void_p(int(stream) if stream else 0))
    if rc < 0:
        raise PicError(rc, L.pic_pbatch_last_error().decode())
    return rel, rc
