"""Hand-derived geometry maps of two small decks that are not plain CCPs:
one with a conductor block inside the plasma, one with a dielectric block.

Both decks have 6 x 5 cells with dx != dy (0.5 mm x 0.3 mm), Dirichlet
electrodes on the bottom (M_ID 2, grounded) and top (M_ID 1, driven) node
rows and Neumann side walls.  The block covers nodes x = 2..4, y = 2..3,
that is cells (2, 2) and (3, 2).  The expected values below were worked out
by hand from the rules of the reference's Geometry_setting
(src/start.c:1432-1851) and Face_To_Area (src/start.c:2053-2080), not from
either restatement:

- Boundary: DIRICHLET (1) on the electrode rows (a Dirichlet wall wins over a
  Neumann one at the corners), NEUMANN (2) on the side walls, CONDUCTOR (3)
  or DIELECTRIC (4) on the block's nodes, 0 elsewhere;
- Face names the side of the block a node sits on (UL/UR/LL/LR corners,
  UP/DOWN edges); an electrode row node faces the plasma (UP on the bottom
  row, DOWN on the top row); corner nodes of the domain follow the
  reference's corner rule, which looks at node ID + 1: the bottom-left
  (UR_CORN) and bottom-right (UL_CORN) corners qualify, the top-left does not
  (its ID + 1 is node (1, 0), an electrode node), and for the top-right the
  reference reads one node past the array, which both restatements take as
  CondID 0 (LL_CORN);
- Area: dx for UP/DOWN, dy for LEFT/RIGHT, (dx + dy) / 2 for corners, dy
  (the initial value) for NO_FACE;
- the padded structure map: 0 plasma, the dielectric's M_ID, 100 + the
  conductor's M_ID, -1 next to a Dirichlet wall, -B - 2 on the bottom,
  -B - 6 on the top, -B(j - 1) on the left and -B(j) - 4 on the right of a
  Neumann wall (B the Boundary code of the node named), so the right wall's
  top cell reads the top electrode's Dirichlet code (-5) while the left
  wall's bottom cell reads the bottom one's (-1).
"""
import numpy as np
import pytest

from oracle import Oracle
from picmcc import FP32_REFINED, HostSetup, decks

NCX, NCY = 6, 5
NGX, NGY = NCX + 1, NCY + 1
DX, DY = 5e-4, 3e-4
LEFT, RIGHT, UP, DOWN, UL, UR, LL, LR, NOFACE = range(9)


def _deck(block):
    d = decks.small(ncx=NCX, ncy=NCY, ppc=4)
    g = d["Geometry"]
    g["SystemSpec"]["X_length(m)"] = NCX * DX
    g["SystemSpec"]["Y_length(m)"] = NCY * DY
    if block == "conductor":
        g["ConductorSpec"].append({"M_ID": 3, "X0": 2, "X1": 4, "Y0": 2, "Y1": 3, "Temp": 300.0,
                                   "R(Ohm)": 0.0, "L(H)": 0.0, "C(F)": 1.0})
    else:
        g["DielectricSpec"] = [{"M_ID": 3, "X0": 2, "X1": 4, "Y0": 2, "Y1": 3, "Epsilon": 4.0}]
    return d


def _node(i, j):
    return i * NGY + j


def _expected(block):
    B = np.zeros(NGX * NGY, dtype=np.int32)
    cid = np.zeros(NGX * NGY, dtype=np.int32)
    face = np.full(NGX * NGY, NOFACE, dtype=np.int32)
    for i in range(NGX):
        B[_node(i, 0)] = B[_node(i, NCY)] = 1
        cid[_node(i, 0)], cid[_node(i, NCY)] = 2, 1
    for j in range(1, NCY):
        B[_node(0, j)] = B[_node(NCX, j)] = 2
    for i in range(2, 5):
        for j in range(2, 4):
            B[_node(i, j)] = 3 if block == "conductor" else 4
            if block == "conductor":
                cid[_node(i, j)] = 3
    # the block's own nodes, conductor and dielectric alike
    face[_node(2, 2)], face[_node(3, 2)], face[_node(4, 2)] = LL, DOWN, LR
    face[_node(2, 3)], face[_node(3, 3)], face[_node(4, 3)] = UL, UP, UR
    # electrode rows between the corners face the plasma; side walls: none
    for i in range(1, NCX):
        face[_node(i, 0)], face[_node(i, NCY)] = UP, DOWN
    face[_node(0, 0)], face[_node(NCX, 0)], face[_node(NCX, NCY)] = UR, UL, LL
    area = np.full(NGX * NGY, DY, dtype=np.float64)
    corner = 0.5 * (DX + DY)
    area[np.isin(face, [UP, DOWN])] = DX
    area[np.isin(face, [UL, UR, LL, LR])] = corner
    # padded structure map, (NCX + 2) x (NCY + 2)
    si = np.zeros((NCX + 2, NCY + 2), dtype=np.int32)
    inner = 103 if block == "conductor" else 3
    si[3, 3] = si[4, 3] = inner
    si[1:NCX + 1, 0] = -1
    si[1:NCX + 1, NCY + 1] = -1
    si[0, 1] = -1               # B(0, 0) = 1
    si[0, 2:NCY + 1] = -2       # B(0, j - 1) = 2
    si[NCX + 1, 1:NCY] = -6     # B(6, j) = 2
    si[NCX + 1, NCY] = -5       # B(6, 5) = 1
    si[0, 0] = si[NCX + 1, 0] = si[0, NCY + 1] = si[NCX + 1, NCY + 1] = -1
    plasma = np.ones(NCX * NCY, dtype=np.int32)
    plasma[2 * NCY + 2] = plasma[3 * NCY + 2] = 0
    eps_r = np.ones(NCX * NCY)
    eps_r[[2 * NCY + 2, 3 * NCY + 2]] = 0.0 if block == "conductor" else 4.0
    return dict(Boundary=B, CondID=cid, Face=face, Area=area, StructIdx=si.ravel(), PlasmaRegion=plasma,
                eps_r=eps_r, DensRegion=np.ones(NGX * NGY, dtype=np.int32))


@pytest.mark.parametrize("block", ["conductor", "dielectric"])
@pytest.mark.parametrize("side", ["oracle", "product"])
def test_geometry_maps_match_hand_derivation(block, side):
    d = _deck(block)
    want = _expected(block)
    if side == "oracle":
        o = Oracle(d, "f32")
        get = o.arr
    else:
        h = HostSetup(d, FP32_REFINED, load=False)
        get = h.array
    for name in ("Boundary", "CondID", "Face", "StructIdx", "PlasmaRegion", "DensRegion"):
        got = np.asarray(get(name)).astype(np.int64)
        w = want[name].astype(np.int64)
        assert np.array_equal(got, w), f"{name}: differs at {np.nonzero(got != w)[0].tolist()}"
    np.testing.assert_allclose(np.asarray(get("Area"), dtype=np.float64), want["Area"], rtol=1e-6)
    np.testing.assert_allclose(np.asarray(get("eps_r"), dtype=np.float64), want["eps_r"], rtol=0)
