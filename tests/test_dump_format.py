"""The V000 dump checker (oracle/dump_v000.py) against the reference's layout.

The fixture below is assembled field by field in the order of the reference's
fwrite calls (src/cuda_SaveDump.cu:5-106, version header 151-155), each field
holding distinct values, on a grid whose every dimension differs (ncx=3,
ncy=2, so ngx=4, ngy=3), so that a transposed array, a swapped interleave or a
miscounted section fails the parse.  The file length is also checked against
the size the reference's reader consumes (src/load.c:5-106).
"""
import struct

import numpy as np

import dump_v000 as V

NSP, NFSP, NCX, NCY, NCOND, TNRCT = 2, 1, 3, 2, 2, 4
D = V.Dims(NSP, NFSP, NCX, NCY, NCOND, TNRCT)
G = D.G
HC, HAC = 3, 2


def fixture():
    """(bytes, expected) built from the reference's fwrite sequence."""
    b = bytearray()
    i32 = lambda *v: b.extend(struct.pack(f"<{len(v)}i", *v))
    f32 = lambda *v: b.extend(struct.pack(f"<{len(v)}f", *v))
    i32(0, 0, 0)
    b.extend(struct.pack("<d", 1.25e-7))
    i32(170, 2, 0)                     # tstep, cstep, MainGas
    parts = []
    for s in range(NSP):
        n = 2 + s
        f32(1000.0 + s)
        i32(n)
        rows = []
        for k in range(n):
            row = (k % 4, 0.5 + k + 10 * s, 0.25 + s, 100.0 * k, -3.0 - s, 7.0 + k)
            b.extend(struct.pack("<i5f", *row))
            rows.append(row)
        parts.append(rows)
    i32(1)                             # Conti_Flag
    for i in range(NCX):               # den and Source, row by row
        f32(*[10 * i + j for j in range(NCY)])
        f32(*[-(10 * i + j) for j in range(NCY)])
    f32(*[200 + k for k in range(D.ngx * NCY)])      # flux_x[ngx][ncy]
    f32(*[300 + k for k in range(NCX * D.ngy)])      # flux_y[ncx][ngy]
    i32(HC, 5)
    f32(*[0.1 * r for r in range(HC)])
    for s in range(NSP):
        f32(*[1000 * s + r for r in range(HC)])
    f32(*[42.0] * HC)
    f32(*[7.0 + r for r in range(HC)])
    for i in range(NCOND):
        f32(*[i + 0.5] * HC)
        for s in range(NSP):
            f32(*[i * 10 + s + 0.25] * HC)
        f32(*[i + 100.0] * HC)
        f32(*[i + 200.0] * HC)
        f32(*[i + 300.0] * HC)
    i32(HAC)
    f32(*[9.0, 9.5])
    for s in range(NSP):
        f32(*[s + 1.0] * HAC)
        f32(*[s + 2.0] * HAC)
    for s in range(NSP):
        for g in range(G):
            f32(1000 * s + g, 2000 * s + g + 0.5, -(1000 * s + g))
    for g in range(G):
        f32(*[g + 0.01 * q for q in range(9)])
    for q in range(6):
        f32(*[100 * q + g for g in range(G)])
    for g in range(G):
        f32(*[g * 10 + r for r in range(TNRCT)])
    return bytes(b), parts


def test_reader_follows_the_reference_layout():
    data, parts = fixture()
    d = V.read(data, D)
    assert d["trailing_bytes"] == 0
    assert d["key"] == [0, 0, 0] and d["t"] == 1.25e-7 and (d["tstep"], d["cstep"], d["gas"]) == (170, 2, 0)
    for s in range(NSP):
        assert d["np2c"][s] == np.float32(1000 + s)
        p = d["particles"][s]
        rows = np.array(parts[s], dtype=np.float64)
        assert np.array_equal(p["cell"], rows[:, 0].astype(np.int32))
        for j, k in enumerate(("x", "y", "vx", "vy", "vz")):
            assert np.array_equal(p[k], rows[:, j + 1].astype(np.float32)), k
    assert d["conti_flag"] == 1
    f = d["fluid"][0]
    assert f["den"][2, 1] == 21 and f["source"][2, 1] == -21
    assert f["flux_x"].shape == (D.ngx, NCY) and f["flux_x"][3, 1] == 200 + 3 * NCY + 1
    assert f["flux_y"].shape == (NCX, D.ngy) and f["flux_y"][2, 2] == 300 + 2 * D.ngy + 2
    h = d["hist"]
    assert d["hist_count"] == HC and d["dhist"] == 5
    assert np.allclose(h["t"], [0, 0.1, 0.2]) and h["np"][1, 2] == 1002 and np.all(h["fluid"] == 42)
    assert np.array_equal(h["iter"], [7, 8, 9])
    assert h["current"][1, 0] == 1.5 and h["sp_current"][1, 1, 0] == 11.25
    assert h["volt"][1, 0] == 101 and h["volt_cond"][1, 0] == 201 and h["surf_charge"][1, 0] == 301
    assert d["hist_ave_count"] == HAC and np.array_equal(d["ave_np_stack"][1], [3, 3])
    assert d["den"][1, 5] == 1005 and d["ave_den"][1, 5] == 2005.5 and d["sigma"][1, 5] == -1005
    assert np.isclose(d["Pois_Pot"][7], 7.06) and np.isclose(d["Ey"][7], 7.08) and np.isclose(d["Temp"][7], 7.0)
    assert d["TotPotential"][4] == 4 and d["ave_Ey"][4] == 504
    assert d["ave_MCC_rate"][5, 3] == 53


def test_length_is_what_the_reference_reader_consumes():
    data, parts = fixture()
    n = sum(len(p) for p in parts)
    size = (12 + 8 + 12 + NSP * 8 + 24 * n + 4 + NFSP * 4 * (2 * NCX * NCY + D.ngx * NCY + NCX * D.ngy)
            + 8 + 4 * HC * (2 + NSP + NFSP + NCOND * (4 + NSP)) + 4 + 4 * HAC * (1 + 2 * NSP)
            + 4 * (3 * NSP * G + 9 * G + 6 * G + TNRCT * G))
    assert len(data) == size


def test_writer_round_trips_byte_for_byte():
    data, _ = fixture()
    assert V.write(V.read(data, D), D) == data


def test_bucket_keeps_file_order_and_exact_offsets():
    # ncx=3, ncy=2: node = xi*3 + yi
    x = np.array([2.5, 0.25, 2.75, 3.0, 1.0, 0.999999940395], np.float32)
    y = np.array([1.5, 0.5, 1.25, 0.5, 2.0, 0.0], np.float32)
    order, cs, ox, oy = V.bucket(x, y, 3, 2, np.float32)
    # cells: 2*3+1=7, 0, 7, 2*3+0=6 (x on the far wall), 1*3+1=4 (y on the far wall), 0
    assert list(order) == [1, 5, 4, 3, 0, 2]
    assert cs[0] == 0 and cs[1] == 2 and cs[-1] == 6 and cs[4] == 2 and cs[5] == 3 and cs[7] == 4
    below = np.nextafter(np.float32(1), np.float32(0))
    assert ox[0] == np.float32(0.25) and ox[1] == np.float32(0.999999940395) and oy[1] == 0
    assert ox[2] == 0 and oy[2] == below        # y == ncy
    assert ox[3] == below and oy[3] == np.float32(0.5)   # x == ncx
    assert ox[4] == np.float32(0.5) and ox[5] == np.float32(0.75) and oy[5] == np.float32(0.25)
    # the offset is the exact fractional part: corner + offset gives the position back
    for k, i in enumerate(order):
        if x[i] < 3 and y[i] < 2:
            assert np.float32(np.floor(x[i])) + ox[k] == x[i] and np.float32(np.floor(y[i])) + oy[k] == y[i]
