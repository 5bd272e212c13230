"""ORACLE - test infrastructure only.

Python front end of the CPU restatement in this directory: maps a case deck
(the reference's JSON format, src/start.c:19-1431) onto the C ``orc_deck``
struct with parson's "missing key reads as 0" rule, loads liborc_f32.so or
liborc_f64.so, and exposes the oracle's arrays as numpy views.  Only tests,
the smoke check and bench.py's CPU-baseline leg use this module.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MAXL, MAXSP, MAXRCT = 32, 5, 100


class OrcDeck(C.Structure):
    _fields_ = [
        ("xlength", C.c_double), ("ylength", C.c_double), ("zlength", C.c_double),
        ("ngx", C.c_int), ("ngy", C.c_int),
        ("nbnd", C.c_int),
        ("bx0", C.c_int * MAXL), ("bx1", C.c_int * MAXL), ("by0", C.c_int * MAXL),
        ("by1", C.c_int * MAXL), ("bbc", C.c_int * MAXL),
        ("btemp", C.c_double * MAXL),
        ("ncond", C.c_int),
        ("cmid", C.c_int * MAXL), ("cx0", C.c_int * MAXL), ("cx1", C.c_int * MAXL),
        ("cy0", C.c_int * MAXL), ("cy1", C.c_int * MAXL),
        ("ctemp", C.c_double * MAXL), ("cR", C.c_double * MAXL), ("cL", C.c_double * MAXL),
        ("cC", C.c_double * MAXL),
        ("nsrc", C.c_int),
        ("smid", C.c_int * MAXL),
        ("sdc", C.c_double * MAXL), ("spow", C.c_double * MAXL), ("sac", C.c_double * MAXL),
        ("sfreq", C.c_double * MAXL), ("sphase", C.c_double * MAXL),
        ("ndiel", C.c_int),
        ("dmid", C.c_int * MAXL), ("dx0", C.c_int * MAXL), ("dx1", C.c_int * MAXL),
        ("dy0", C.c_int * MAXL), ("dy1", C.c_int * MAXL),
        ("deps", C.c_double * MAXL),
        ("gas", C.c_int),
        ("total_pres", C.c_double),
        ("bg_ratio", C.c_double * 2), ("bg_temp", C.c_double * 2),
        ("fl_loadtype", C.c_int),
        ("fl_x0", C.c_double), ("fl_x1", C.c_double), ("fl_y0", C.c_double), ("fl_y1", C.c_double),
        ("fl_dens", C.c_double * 5), ("fl_temp", C.c_double * 5), ("fl_quench", C.c_double * 5),
        ("sp_loadtype", C.c_int),
        ("sp_x0", C.c_double), ("sp_x1", C.c_double), ("sp_y0", C.c_double), ("sp_y1", C.c_double),
        ("sp_dens", C.c_double * MAXSP), ("sp_temp", C.c_double * MAXSP), ("sp_np2c", C.c_double * MAXSP),
        ("sp_ratio", C.c_int * MAXSP), ("sp_maxnp", C.c_int * MAXSP),
        ("dt_pic", C.c_int),
        ("dt_conti", C.c_int),
        ("pcg_tol", C.c_double),
        ("n_ave", C.c_int), ("n_smt", C.c_int), ("np_limit", C.c_int),
        ("coll_flag", C.c_double * MAXRCT),
        ("seed", C.c_int64),
        ("rng_seed", C.c_int64),
    ]


def _num(obj, key):
    """parson json_object_get_number: 0 for a missing key or a non-number."""
    if not isinstance(obj, dict):
        return 0.0
    v = obj.get(key, 0)
    if isinstance(v, bool) or not isinstance(v, (int, float)):
        return 0.0
    return float(v)


def _obj(obj, key):
    if not isinstance(obj, dict):
        return {}
    v = obj.get(key)
    return v if isinstance(v, dict) else {}


def _arr(obj, key):
    if not isinstance(obj, dict):
        return []
    v = obj.get(key)
    return v if isinstance(v, list) else []


def _int(x):
    return int(x)  # C (int) cast truncates toward zero, as Python int() does


AR_KEYS = ["0.e+Ar>e+Ar", "1.e+Ar>e+Ar*", "2.e+Ar>e+Ar*m", "3.e+Ar>2e+Ar^",
           "4.e+Ar*m>2e+Ar^", "5.Ar+Ar^>Ar^+Ar", "6.Ar+Ar^>Ar+Ar^", "7.e+Ar*m>e+Ar",
           "8.Ar*m+Ar>e+Ar^+Ar", "9.Ar+Ar*m>Ar+Ar"]   # src/start.c:1005-1034


def deck_to_struct(deck):
    d = OrcDeck()
    geo = _obj(deck, "Geometry")
    ss = _obj(geo, "SystemSpec")
    d.xlength, d.ylength, d.zlength = _num(ss, "X_length(m)"), _num(ss, "Y_length(m)"), _num(ss, "Z_length(m)")
    d.ngx, d.ngy = _int(_num(ss, "NumGridx")), _int(_num(ss, "NumGridy"))
    bl = _arr(geo, "BoundaryCondition")
    d.nbnd = len(bl)
    for i, b in enumerate(bl[:MAXL]):
        d.bx0[i], d.bx1[i] = _int(_num(b, "X0")), _int(_num(b, "X1"))
        d.by0[i], d.by1[i] = _int(_num(b, "Y0")), _int(_num(b, "Y1"))
        d.bbc[i] = _int(_num(b, "B.C"))
        d.btemp[i] = _num(b, "Temp")
    cl = _arr(geo, "ConductorSpec")
    d.ncond = len(cl)
    for i, c in enumerate(cl[:MAXL]):
        d.cmid[i] = _int(_num(c, "M_ID"))
        d.cx0[i], d.cx1[i] = _int(_num(c, "X0")), _int(_num(c, "X1"))
        d.cy0[i], d.cy1[i] = _int(_num(c, "Y0")), _int(_num(c, "Y1"))
        d.ctemp[i], d.cR[i], d.cL[i], d.cC[i] = (_num(c, "Temp"), _num(c, "R(Ohm)"),
                                               _num(c, "L(H)"), _num(c, "C(F)"))
    sl = _arr(geo, "Source")
    d.nsrc = len(sl)
    for i, s in enumerate(sl[:MAXL]):
        d.smid[i] = _int(_num(s, "M_ID"))
        d.sdc[i], d.spow[i], d.sac[i] = _num(s, "DC(V)"), _num(s, "Power(W)"), _num(s, "AC(V)")
        d.sfreq[i], d.sphase[i] = _num(s, "Freq(1/s)"), _num(s, "Phase(deg)")
    dl = _arr(geo, "DielectricSpec")
    d.ndiel = len(dl)
    for i, x in enumerate(dl[:MAXL]):
        d.dmid[i] = _int(_num(x, "M_ID"))
        d.dx0[i], d.dx1[i] = _int(_num(x, "X0")), _int(_num(x, "X1"))
        d.dy0[i], d.dy1[i] = _int(_num(x, "Y0")), _int(_num(x, "Y1"))
        d.deps[i] = _num(x, "Epsilon")
    gas = _obj(deck, "GasSpecies")
    d.gas = _int(_num(gas, "Type(0:Ar,1:O2,2:Ar/O2)"))
    d.total_pres = _num(gas, "TotalPres(Torr)")
    bg = _obj(gas, "Background")
    for k, nm in enumerate(["Argon", "Oxygen"]):
        d.bg_ratio[k] = _num(_obj(bg, nm), "Ratio")
        d.bg_temp[k] = _num(_obj(bg, nm), "Temp(eV)")
    ns = _obj(gas, "NeutralSpecies")
    d.fl_loadtype = _int(_num(ns, "Loadtype"))
    lp = _obj(ns, "LoadPosition(m)")
    d.fl_x0, d.fl_x1, d.fl_y0, d.fl_y1 = _num(lp, "X0"), _num(lp, "X1"), _num(lp, "Y0"), _num(lp, "Y1")
    flnames = {0: ["Ar*"], 1: ["O2A", "O2B", "OP", "OD"], 2: ["Ar*", "O2A", "O2B", "OP", "OD"]}.get(d.gas, [])
    for k, nm in enumerate(flnames):
        o = _obj(ns, nm)
        d.fl_dens[k], d.fl_temp[k], d.fl_quench[k] = _num(o, "Density"), _num(o, "Temp(eV)"), _num(o, "Quenching")
    cs = _obj(gas, "ChargeSpecies")
    d.sp_loadtype = _int(_num(cs, "Loadtype"))
    lp = _obj(cs, "LoadPosition(m)")
    d.sp_x0, d.sp_x1, d.sp_y0, d.sp_y1 = _num(lp, "X0"), _num(lp, "X1"), _num(lp, "Y0"), _num(lp, "Y1")
    spnames = {0: ["Electron", "Ar+"], 1: ["Electron", "O2+", "O+", "O-"],
               2: ["Electron", "Ar+", "O2+", "O+", "O-"]}.get(d.gas, [])
    for k, nm in enumerate(spnames):
        o = _obj(cs, nm)
        d.sp_dens[k], d.sp_temp[k], d.sp_np2c[k] = _num(o, "Density"), _num(o, "Temp(eV)"), _num(o, "np2c")
        d.sp_ratio[k], d.sp_maxnp[k] = _int(_num(o, "Ratio")), _int(_num(o, "Max_np"))
    sm = _obj(deck, "SimulationMethod")
    d.dt_pic = _int(_num(sm, "TimeStep_PIC"))
    d.dt_conti = _int(_num(sm, "TimeStep_Conti"))
    d.pcg_tol = _num(sm, "PCGMarginOfError")
    d.n_ave = _int(_num(sm, "AverageIter"))
    d.n_smt = _int(_num(sm, "DensitySmoothing"))
    d.np_limit = _int(_num(sm, "ParticleLimit"))
    d.seed = _int(_num(sm, "LoaderSeed")) or 1
    d.rng_seed = _int(_num(sm, "PhiloxSeed"))
    cc = _obj(deck, "Collision_On_Off")
    if d.gas == 0:
        ac = _obj(cc, "ArgonCase")
        for r, k in enumerate(AR_KEYS):
            d.coll_flag[r] = _num(ac, k)
    else:
        # "N.reaction" keys (src/start.c:1036-1230): the number is the index
        ac = _obj(cc, {1: "OxygenCase", 2: "Argon/OxygenCase"}[d.gas])
        for k, v in ac.items():
            d.coll_flag[int(k.split(".")[0])] = float(v)
    return d


class Oracle:
    """One oracle case.  precision is 'f32', 'f64', or 'f64_omp' (the FP64
    oracle built with OpenMP, which adds orc_run_par(nsteps, threads))."""

    def __init__(self, deck, precision="f64"):
        path = os.path.join(HERE, f"liborc_{precision}.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library {path} is not built (run make -C oracle)")
        self.lib = C.CDLL(path)
        L = self.lib
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(OrcDeck), C.POINTER(C.c_int)]
        L.orc_array.restype = C.c_void_p
        L.orc_array.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_long), C.POINTER(C.c_int)]
        L.orc_scalar.restype = C.c_double
        L.orc_scalar.argtypes = [C.c_void_p, C.c_char_p]
        for fn in ["orc_field", "orc_efield", "orc_move", "orc_deposit", "orc_accumulate", "orc_fluid_step",
                   "orc_fluid_continuity", "orc_fluid_diag"]:
            getattr(L, fn).restype = None
            getattr(L, fn).argtypes = [C.c_void_p]
        for fn in ["orc_sort", "orc_mcc", "orc_step"]:
            getattr(L, fn).restype = C.c_int
            getattr(L, fn).argtypes = [C.c_void_p]
        L.orc_run.restype = C.c_int
        L.orc_run.argtypes = [C.c_void_p, C.c_int]
        L.orc_set_np.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_set_tstep.argtypes = [C.c_void_p, C.c_long]
        L.orc_set_fluid.argtypes = [C.c_void_p, C.c_int]
        L.orc_set_fluid.restype = None
        L.orc_destroy.restype = None
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_voltages.argtypes = [C.c_void_p, C.c_double]
        if hasattr(L, "orc_run_par"):
            L.orc_run_par.restype = C.c_int
            L.orc_run_par.argtypes = [C.c_void_p, C.c_int, C.c_int]
        assert L.orc_sizeof_deck() == C.sizeof(OrcDeck), "orc_deck layout mismatch"
        self.real = np.float32 if precision == "f32" else np.float64
        self.deck_struct = deck_to_struct(deck)
        err = C.c_int(0)
        self.h = L.orc_create(C.byref(self.deck_struct), C.byref(err))
        self.err = err.value
        if self.err:
            raise ValueError(f"oracle setup failed with code {self.err}")

    def arr(self, name):
        n, es = C.c_long(0), C.c_int(0)
        p = self.lib.orc_array(self.h, name.encode(), C.byref(n), C.byref(es))
        if not p:
            raise KeyError(name)
        dt = {4: np.int32, 8: np.int64}
        if name in ("Boundary", "CondID", "Face", "DensRegion", "PlasmaRegion", "StructIdx",
                    "A_idx", "Ai", "Aj", "fl_x1", "fl_x2", "fl_xy", "fl_yx", "fl_y1", "fl_y2") or \
                name.endswith((".cell_start", ".tag")):
            dtype = np.int32
        elif name.endswith(".id"):
            dtype = np.uint32
        elif name in ("mccr",):
            dtype = np.int64
        elif es.value == 8 and name in ("xsol", "sum_den", "sum_phi", "sum_ex", "sum_ey", "lu"):
            dtype = np.float64
        else:
            dtype = self.real if es.value == np.dtype(self.real).itemsize else dt[es.value]
        buf = (C.c_char * (n.value * es.value)).from_address(p)
        return np.frombuffer(buf, dtype=dtype, count=n.value)

    def scalar(self, name):
        return self.lib.orc_scalar(self.h, name.encode())

    def close(self):
        """Frees the C case (the arrays returned by arr() die with it)."""
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def species(self, s):
        n = int(self.scalar(f"np{s}"))
        out = {k: self.arr(f"sp{s}.{k}")[:n] for k in ("x", "y", "vx", "vy", "vz", "id")}
        out["cell_start"] = self.arr(f"sp{s}.cell_start")
        return out

    def set_species(self, s, st):
        """Overwrite species s with arrays x,y,vx,vy,vz,id,cell_start."""
        n = len(st["x"])
        cap = int(self.scalar(f"cap{s}"))
        assert n <= cap
        for k in ("x", "y", "vx", "vy", "vz", "id"):
            self.arr(f"sp{s}.{k}")[:n] = st[k]
        self.arr(f"sp{s}.cell_start")[:] = st["cell_start"]
        self.lib.orc_set_np(self.h, s, n)

    def __getattr__(self, name):
        if name.startswith("orc_"):
            fn = getattr(self.lib, name)
            return lambda *a: fn(self.h, *a)
        raise AttributeError(name)


def mcc_o2(precision, k, buckets, next_id, mccr, delcnt=None):
    """Collision step of the oxygen chemistries (orc_mcc_o2) on host copies.

    k: dict of the setup values (G, TnRct, nlogx, nvel, MCCn, idLOGX, dt,
    dtm, idx, idy, np2c, Escale, sigv, seed, step) and arrays (Th_e, xe, cx,
    vsave, BackDen1, BackVel1, FluidN, den; BackDen2 and BackVel2 for Ar/O2),
    gas (1 O2, the default; 2 Ar/O2) and ions (O2 only: 1 = the positive and
    negative ions collide too; default 0).  buckets: one dict per species
    (x, y, vx, vy, vz, id, cell_start, cap): e, O2+, O+, O- for O2; e, Ar+,
    O2+, O+, O- for Ar/O2.  delcnt: per-node count of the O- removed by
    electron detachment, advanced when given.  Returns (new buckets,
    next_id, mccr, rc), or with delcnt (..., rc, delcnt); the inputs are not
    modified.
    """
    path = os.path.join(HERE, f"liborc_{precision}.so")
    if not os.path.exists(path):
        raise RuntimeError(f"oracle library {path} is not built (run make -C oracle)")
    L = C.CDLL(path)
    real = np.float32 if precision == "f32" else np.float64
    R = C.c_float if precision == "f32" else C.c_double
    P = C.POINTER(R)
    gas = int(k.get("gas", 1))
    if gas not in (1, 2):
        raise ValueError("gas must be 1 (O2) or 2 (Ar/O2)")
    nsp = 4 if gas == 1 else 5
    nsig = 20 if gas == 1 else 25

    class In(C.Structure):
        _fields_ = [("G", C.c_int), ("TnRct", C.c_int), ("nlogx", C.c_int), ("nvel", C.c_int), ("MCCn", C.c_int),
                    ("idLOGX", R), ("dt", R), ("dtm", R), ("idx", R), ("idy", R),
                    ("np2c", R * 5), ("Escale", R * 5), ("sigv", R * 25),
                    ("Th_e", P), ("xe", P), ("cx", P), ("vsave", P), ("BackDen1", P), ("BackVel1", P),
                    ("FluidN", P), ("den", P), ("seed", C.c_uint64), ("step", C.c_uint32), ("ions", C.c_int),
                    ("chem", C.c_int), ("BackDen2", P), ("BackVel2", P), ("delcnt", C.POINTER(C.c_int64))]

    class Bucket(C.Structure):
        _fields_ = [("x", P), ("y", P), ("vx", P), ("vy", P), ("vz", P), ("id", C.POINTER(C.c_uint32)),
                    ("cell_start", C.POINTER(C.c_int)), ("n", C.c_int), ("cap", C.c_int)]

    keep = []

    def ptr(a, dt, ct):
        a = np.ascontiguousarray(np.asarray(a, dtype=dt)).copy()
        keep.append(a)
        return a, a.ctypes.data_as(C.POINTER(ct))

    inp = In()
    for f in ("G", "TnRct", "nlogx", "nvel", "MCCn"):
        setattr(inp, f, int(k[f]))
    for f in ("idLOGX", "dt", "dtm", "idx", "idy"):
        setattr(inp, f, float(real(k[f])))
    for f, n in (("np2c", nsp), ("Escale", nsp), ("sigv", nsig)):
        arr = getattr(inp, f)
        for i in range(n):
            arr[i] = float(real(k[f][i]))
    for f in ("Th_e", "xe", "cx", "vsave", "BackDen1", "BackVel1", "FluidN", "den"):
        setattr(inp, f, ptr(k[f], real, R)[1])
    for f in ("BackDen2", "BackVel2"):
        setattr(inp, f, ptr(k[f] if gas == 2 else k[f.replace("2", "1")], real, R)[1])
    inp.seed = int(k["seed"]) & (2 ** 64 - 1)
    inp.step = int(k["step"]) & (2 ** 32 - 1)
    inp.ions = int(k.get("ions", 0))
    inp.chem = gas - 1
    dc = None
    if delcnt is not None:
        dc, inp.delcnt = ptr(delcnt, np.int64, C.c_int64)
    bs = (Bucket * nsp)()
    arrays = []
    for s in range(nsp):
        b = buckets[s]
        cap = int(b["cap"])
        n = len(b["x"])
        d = {}
        for f in ("x", "y", "vx", "vy", "vz"):
            a = np.zeros(cap, dtype=real)
            a[:n] = b[f]
            d[f], p = ptr(a, real, R)
            setattr(bs[s], f, p)
        a = np.zeros(cap, dtype=np.uint32)
        a[:n] = b["id"]
        d["id"], bs[s].id = ptr(a, np.uint32, C.c_uint32)
        d["cell_start"], bs[s].cell_start = ptr(b["cell_start"], np.int32, C.c_int)
        bs[s].n = n
        bs[s].cap = cap
        arrays.append(d)
    nid, nid_p = ptr(next_id, np.uint32, C.c_uint32)
    mc, mc_p = ptr(mccr, np.int64, C.c_int64)
    L.orc_mcc_o2.restype = C.c_int
    L.orc_mcc_o2.argtypes = [C.POINTER(In), C.POINTER(Bucket), C.POINTER(C.c_uint32), C.POINTER(C.c_int64)]
    rc = L.orc_mcc_o2(C.byref(inp), bs, nid_p, mc_p)
    out = []
    for s in range(nsp):
        n = bs[s].n
        d = {f: arrays[s][f][:n].copy() for f in ("x", "y", "vx", "vy", "vz", "id")}
        d["cell_start"] = arrays[s]["cell_start"].copy()
        out.append(d)
    if delcnt is not None:
        return out, nid.copy(), mc.copy(), rc, dc.copy()
    return out, nid.copy(), mc.copy(), rc
