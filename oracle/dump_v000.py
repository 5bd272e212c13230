"""The reference's V000 dump format, restated in numpy as a checker.

TEST INFRASTRUCTURE: only the tests use this module; the library writes and
reads the format itself (pic_dump_save / pic_dump_load).

Layout (all 4-byte little-endian ints and floats except t), following the
writer src/cuda_SaveDump.cu:5-106 with its version header at 151-155, and the
reader src/load.c:5-117:

    KEY2 KEY1 KEY0                              int32 x 3
    t                                           float64
    tstep cstep MainGas                         int32 x 3
    per species: np2c (f32), np (i32), np x [CellID i32, x y vx vy vz f32]
    Conti_Flag                                  int32
    per fluid species: den[ncx][ncy], Source[ncx][ncy] (row by row, the
        two interleaved per row), flux_x[ngx][ncy], flux_y[ncx][ngy]
    hist_count dHIST                            int32 x 2
    t_array[hc], HistPt[nsp][hc], HistFG[nfsp][hc], iter_array[hc]
    per conductor: Current[hc], SP_current[nsp][hc], Volt[hc],
        Volt_cond[hc], Surf_charge[hc]
    hist_ave_count                              int32
    t_ave[hac], per species: ave_np[hac], ave_np_stack[hac]
    per species and node: den, ave_den, sigma  (interleaved)
    per node: Temp BackVel1 BackDen1 BackVel2 BackDen2 Lap_Pot Pois_Pot Ex Ey
    TotPotential, ave_Potential, ave_Source, ave_Sigma, ave_Ex, ave_Ey  [G each]
    ave_MCC_rate[G][TnRct]
"""
import numpy as np

# fluid species per gas type (src/start.c:307-317)
NFSP = {0: 1, 1: 4, 2: 5}


class Dims:
    def __init__(self, nsp, nfsp, ncx, ncy, ncond, tnrct):
        self.nsp, self.nfsp, self.ncx, self.ncy, self.ncond, self.tnrct = nsp, nfsp, ncx, ncy, ncond, tnrct
        self.ngx, self.ngy = ncx + 1, ncy + 1
        self.G = self.ngx * self.ngy


class _Cursor:
    def __init__(self, buf):
        self.b = memoryview(buf)
        self.o = 0

    def take(self, dtype, n):
        dt = np.dtype(dtype)
        end = self.o + dt.itemsize * n
        if end > len(self.b):
            raise ValueError(f"dump ends early: need {end} bytes, have {len(self.b)}")
        a = np.frombuffer(self.b[self.o:end], dtype=dt).copy()
        self.o = end
        return a

    def i32(self):
        return int(self.take("<i4", 1)[0])


def read(data, dims):
    """Parses a V000 dump (bytes or a path) into a dict of numpy arrays."""
    if not isinstance(data, (bytes, bytearray)):
        with open(data, "rb") as f:
            data = f.read()
    d, c = dims, _Cursor(data)
    out = {"key": [c.i32(), c.i32(), c.i32()]}
    out["t"] = float(c.take("<f8", 1)[0])
    out["tstep"], out["cstep"], out["gas"] = c.i32(), c.i32(), c.i32()
    out["np2c"], out["particles"] = [], []
    for _ in range(d.nsp):
        out["np2c"].append(np.float32(c.take("<f4", 1)[0]))
        n = c.i32()
        raw = c.take("<u4", 6 * n).reshape(n, 6)
        out["particles"].append({"cell": raw[:, 0].view(np.int32).copy(),
                                 "x": raw[:, 1].view(np.float32).copy(), "y": raw[:, 2].view(np.float32).copy(),
                                 "vx": raw[:, 3].view(np.float32).copy(), "vy": raw[:, 4].view(np.float32).copy(),
                                 "vz": raw[:, 5].view(np.float32).copy()})
    out["conti_flag"] = c.i32()
    out["fluid"] = []
    for _ in range(d.nfsp):
        den = np.empty((d.ncx, d.ncy), np.float32)
        src = np.empty((d.ncx, d.ncy), np.float32)
        for i in range(d.ncx):
            den[i] = c.take("<f4", d.ncy)
            src[i] = c.take("<f4", d.ncy)
        fx = c.take("<f4", d.ngx * d.ncy).reshape(d.ngx, d.ncy)
        fy = c.take("<f4", d.ncx * d.ngy).reshape(d.ncx, d.ngy)
        out["fluid"].append({"den": den, "source": src, "flux_x": fx, "flux_y": fy})
    hc = c.i32()
    out["hist_count"], out["dhist"] = hc, c.i32()
    h = {"t": c.take("<f4", hc)}
    h["np"] = c.take("<f4", d.nsp * hc).reshape(d.nsp, hc)
    h["fluid"] = c.take("<f4", d.nfsp * hc).reshape(d.nfsp, hc)
    h["iter"] = c.take("<f4", hc)
    h["current"], h["sp_current"], h["volt"], h["volt_cond"], h["surf_charge"] = [], [], [], [], []
    for _ in range(d.ncond):
        h["current"].append(c.take("<f4", hc))
        h["sp_current"].append(c.take("<f4", d.nsp * hc).reshape(d.nsp, hc))
        h["volt"].append(c.take("<f4", hc))
        h["volt_cond"].append(c.take("<f4", hc))
        h["surf_charge"].append(c.take("<f4", hc))
    for k in ("current", "sp_current", "volt", "volt_cond", "surf_charge"):
        h[k] = np.array(h[k], dtype=np.float32)
    out["hist"] = h
    hac = c.i32()
    out["hist_ave_count"] = hac
    out["t_ave"] = c.take("<f4", hac)
    out["ave_np"] = np.empty((d.nsp, hac), np.float32)
    out["ave_np_stack"] = np.empty((d.nsp, hac), np.float32)
    for s in range(d.nsp):
        out["ave_np"][s] = c.take("<f4", hac)
        out["ave_np_stack"][s] = c.take("<f4", hac)
    g3 = c.take("<f4", d.nsp * d.G * 3).reshape(d.nsp, d.G, 3)
    out["den"], out["ave_den"], out["sigma"] = g3[:, :, 0], g3[:, :, 1], g3[:, :, 2]
    g9 = c.take("<f4", d.G * 9).reshape(d.G, 9)
    for k, name in enumerate(("Temp", "BackVel1", "BackDen1", "BackVel2", "BackDen2", "Lap_Pot", "Pois_Pot",
                              "Ex", "Ey")):
        out[name] = g9[:, k].copy()
    for name in ("TotPotential", "ave_Potential", "ave_Source", "ave_Sigma", "ave_Ex", "ave_Ey"):
        out[name] = c.take("<f4", d.G)
    out["ave_MCC_rate"] = c.take("<f4", d.G * d.tnrct).reshape(d.G, d.tnrct)
    out["trailing_bytes"] = len(data) - c.o
    return out


def write(d, dims):
    """The bytes of a V000 dump holding the dict `d` (the layout read() returns)."""
    D = dims
    parts = [np.array(d["key"], "<i4").tobytes(), np.array([d["t"]], "<f8").tobytes(),
             np.array([d["tstep"], d["cstep"], d["gas"]], "<i4").tobytes()]
    for s in range(D.nsp):
        p = d["particles"][s]
        n = len(p["x"])
        parts.append(np.array([d["np2c"][s]], "<f4").tobytes())
        parts.append(np.array([n], "<i4").tobytes())
        raw = np.empty((n, 6), np.uint32)
        raw[:, 0] = np.asarray(p["cell"], np.int32).view(np.uint32)
        for j, k in enumerate(("x", "y", "vx", "vy", "vz")):
            raw[:, j + 1] = np.asarray(p[k], np.float32).view(np.uint32)
        parts.append(raw.astype("<u4").tobytes())
    parts.append(np.array([d["conti_flag"]], "<i4").tobytes())
    for f in d["fluid"]:
        for i in range(D.ncx):
            parts.append(np.asarray(f["den"][i], "<f4").tobytes())
            parts.append(np.asarray(f["source"][i], "<f4").tobytes())
        parts.append(np.asarray(f["flux_x"], "<f4").tobytes())
        parts.append(np.asarray(f["flux_y"], "<f4").tobytes())
    h = d["hist"]
    parts.append(np.array([d["hist_count"], d["dhist"]], "<i4").tobytes())
    for k in ("t", "np", "fluid", "iter"):
        parts.append(np.asarray(h[k], "<f4").tobytes())
    for i in range(D.ncond):
        for k in ("current", "sp_current", "volt", "volt_cond", "surf_charge"):
            parts.append(np.asarray(h[k][i], "<f4").tobytes())
    parts.append(np.array([d["hist_ave_count"]], "<i4").tobytes())
    parts.append(np.asarray(d["t_ave"], "<f4").tobytes())
    for s in range(D.nsp):
        parts.append(np.asarray(d["ave_np"][s], "<f4").tobytes())
        parts.append(np.asarray(d["ave_np_stack"][s], "<f4").tobytes())
    parts.append(np.stack([d["den"], d["ave_den"], d["sigma"]], axis=-1).astype("<f4").tobytes())
    parts.append(np.stack([d[k] for k in ("Temp", "BackVel1", "BackDen1", "BackVel2", "BackDen2", "Lap_Pot",
                                          "Pois_Pot", "Ex", "Ey")], axis=-1).astype("<f4").tobytes())
    for k in ("TotPotential", "ave_Potential", "ave_Source", "ave_Sigma", "ave_Ex", "ave_Ey", "ave_MCC_rate"):
        parts.append(np.asarray(d[k], "<f4").tobytes())
    return b"".join(parts)


def bucket(x, y, ncx, ncy, dtype):
    """Buckets absolute positions (float32, cell units) by cell, keeping file
    order within a cell (Copy_HCPtoGCP, src/cuda_Particle.cu:27-58: the cell is
    the truncated position, the in-cell offset the position minus the cell's
    corner).  A position on the far wall (x == ncx or y == ncy) stays in the
    last cell at the largest offset below 1; the reference has no such cell
    and would index the wall's node.  Returns (order, cell_start, ox, oy):
    order[k] = file index of the k-th particle in bucket order."""
    x = np.asarray(x, np.float32)
    y = np.asarray(y, np.float32)
    ngy = ncy + 1
    xi = x.astype(np.int64)
    yi = y.astype(np.int64)
    below = np.nextafter(dtype(1), dtype(0))
    ox = np.where(xi < ncx, (x - xi.astype(np.float32)).astype(dtype), below)
    oy = np.where(yi < ncy, (y - yi.astype(np.float32)).astype(dtype), below)
    node = np.minimum(xi, ncx - 1) * ngy + np.minimum(yi, ncy - 1)
    order = np.argsort(node, kind="stable")
    G = (ncx + 1) * ngy
    cs = np.zeros(G + 1, np.int64)
    np.add.at(cs, node + 1, 1)
    cs = np.cumsum(cs)
    return order, cs, ox[order], oy[order]
