"""The liquid-box initial condition against an independent construction.

The reference hard-codes its initial free surface per problem (LB.cpp:611-784)
and has no liquid box, so no reference run pins this initialiser. These tests
pin it by what the definition fixes: the liquid nodes are the lattice nodes of
the inclusive box [x0,x1]x[y0,y1]x[z0,z1]; the interface is the box's shell of
nodes that have a gas node among their 18 D3Q19 neighbours (LB.cpp:787-797),
so no shell on a face lying against a wall or wrapping through a periodic
plane onto liquid; and shell nodes left without a fluid neighbour become gas
(LB.cpp:799-814), which removes exactly the box corners where three exposed
faces meet (D3Q19 has no body diagonals). The active-node count is therefore
the product of the inclusive extents minus those corners, written out per case
below, and the expected types are built with numpy slicing, independently of
both implementations; an off-by-one in the box bounds, a missed diagonal
neighbour or a wrong periodic wrap fails them.
"""
import numpy as np
import pytest

from hybird_amd.cfg import base_config
from helpers import fixture_config
from oracle.oracle import OracleSim

# D3Q19 neighbour offsets (lattice.h:36-60), rest direction excluded
_C = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 1, 0), (-1, -1, 0), (-1, 1, 0),
      (1, -1, 0), (0, 1, 1), (0, -1, -1), (0, 1, -1), (0, -1, 1), (1, 0, 1), (-1, 0, -1), (1, 0, -1), (-1, 0, 1)]

# (lattice X, Y, Z, boundary codes, inclusive box, active nodes = extents product - isolated corners)
CASES = {
    # closed tank, box against the floor (z = 1), gas above and around it: the 4 top corners go
    "tank_floor": ((12, 10, 14), (7, 7, 7, 7, 7, 7), (3, 8, 2, 6, 1, 5), 6 * 5 * 5 - 4),
    # box spanning the whole periodic x range: no interface on the x faces, no corner
    "periodic_x_span": ((10, 12, 12), (4, 4, 7, 7, 7, 7), (1, 8, 3, 7, 1, 6), 8 * 5 * 6),
    # box against the periodic x plane from one side: its x=1 face wraps onto gas at x = X-2
    "periodic_x_edge": ((12, 9, 11), (4, 4, 4, 4, 7, 7), (1, 5, 2, 6, 1, 4), 5 * 5 * 4 - 4),
    # three layers in y floating in the tank: one layer of fluid, all 8 corners go
    "slab": ((9, 10, 9), (7, 7, 7, 7, 7, 7), (2, 6, 4, 6, 2, 6), 5 * 3 * 5 - 8),
}


def expected_types(shape, bnd, box):
    X, Y, Z = shape
    t = np.zeros((X, Y, Z), dtype=np.uint8)  # interior starts as gas below
    wall = lambda b: b in (5, 6, 7, 8)
    # boundary layers: x faces, then y and z unless already a wall (LB.cpp:394-431)
    t[0, :, :] = bnd[0]
    t[X - 1, :, :] = bnd[1]
    for sl, b in (((slice(None), 0, slice(None)), bnd[2]), ((slice(None), Y - 1, slice(None)), bnd[3]),
                  ((slice(None), slice(None), 0), bnd[4]), ((slice(None), slice(None), Z - 1), bnd[5])):
        view = t[sl]
        keep = np.vectorize(wall)(view) if view.size else view
        t[sl] = np.where(keep & (view != 0), view, b)
    interior = np.zeros_like(t, dtype=bool)
    interior[1:X - 1, 1:Y - 1, 1:Z - 1] = True
    x0, x1, y0, y1, z0, z1 = box
    liquid = np.zeros_like(interior)
    liquid[x0:x1 + 1, y0:y1 + 1, z0:z1 + 1] = True
    liquid &= interior
    t[interior] = 2
    t[liquid] = 0
    per = [bnd[0] == 4, bnd[2] == 4, bnd[4] == 4]
    size = (X, Y, Z)

    def nb(i, k, d):
        j = i + d
        if per[k] and j == 0:
            return size[k] - 2
        if per[k] and j == size[k] - 1:
            return 1
        return j

    def nbrs(x, y, z):
        return [(nb(x, 0, c[0]), nb(y, 1, c[1]), nb(z, 2, c[2])) for c in _C]

    out = t.copy()
    for (x, y, z) in zip(*np.nonzero(liquid)):
        if any(t[n] == 2 for n in nbrs(x, y, z)):
            out[x, y, z] = 3
    final = out.copy()
    for (x, y, z) in zip(*np.nonzero(out == 3)):
        if not any(out[n] == 0 for n in nbrs(x, y, z)):
            final[x, y, z] = 2
    # linear index x + X (y + Y z): flatten in Fortran order
    return final.ravel(order="F"), int(liquid.sum())


def _config(shape, bnd, box):
    cfg = base_config(problemName="LIQUID_BOX", freeSurfaceSolve=1, lbSizeX=shape[0], lbSizeY=shape[1],
                      lbSizeZ=shape[2], initVisc=0.05, lbFZ=-1.0e-4,
                      **{f"boundary{k}": bnd[k] for k in range(6)})
    return fixture_config(cfg, liquid_box=box)


def _check(types, shape, bnd, box, nact):
    exp, nliq = expected_types(shape, bnd, box)
    x0, x1, y0, y1, z0, z1 = box
    assert nliq == (x1 - x0 + 1) * (y1 - y0 + 1) * (z1 - z0 + 1)
    assert int(((exp == 0) | (exp == 3)).sum()) == nact
    act = (types == 0) | (types == 3)
    assert int(act.sum()) == nact
    assert np.array_equal(types, exp), f"{int(np.sum(types != exp))} node types differ"


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_liquid_box_types(name):
    shape, bnd, box, nact = CASES[name]
    orc = OracleSim(_config(shape, bnd, box), np.zeros((0, 20)))
    _check(orc.get("types"), shape, bnd, box, nact)
    orc.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_liquid_box_types(name):
    from hybird_amd import runtime

    shape, bnd, box, nact = CASES[name]
    print(f"[gpu liquid box] case={name}", flush=True)
    gpu = runtime.Simulation(_config(shape, bnd, box), np.zeros((0, 20)))
    _check(gpu.get("types"), shape, bnd, box, nact)
    gpu.close()
