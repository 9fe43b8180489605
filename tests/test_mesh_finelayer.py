"""Mesh::Quad4::FineLayer (host set-up, gf_mesh_quad4_finelayer*): the reference's golden mesh
(tests/basic/MeshQuad4.cpp:103-256, FineLayer(9, 18)) entry for entry, and the invariants every
FineLayer mesh satisfies (conforming, positive element areas summing to the domain area, every node
used, the middle layer finest), over a sweep of sizes and fine-layer counts."""

import numpy as np
import pytest

import goosefem_amd as gf

SETS = ["nodesBottomEdge", "nodesTopEdge", "nodesLeftEdge", "nodesRightEdge", "nodesBottomOpenEdge",
        "nodesTopOpenEdge", "nodesLeftOpenEdge", "nodesRightOpenEdge"]
CORNERS = ["nodesBottomLeftCorner", "nodesBottomRightCorner", "nodesTopLeftCorner", "nodesTopRightCorner",
           "nodesLeftBottomCorner", "nodesLeftTopCorner", "nodesRightBottomCorner", "nodesRightTopCorner",
           "nodesOrigin"]


def test_finelayer_golden(fixtures):
    g = fixtures["MeshQuad4_FineLayer_9_18"]
    m = gf.Mesh.Quad4.FineLayer(9, 18)
    assert (m.nelem, m.nnode, m.nne, m.ndim) == (g["nelem_"], g["nnode_"], g["nne_"], g["ndim_"])
    assert (m.nelx, m.nely) == (g["shape_x_"], g["shape_y_"])
    assert np.allclose(m.coor, np.array(g["coor_"]))
    assert np.array_equal(m.conn, np.array(g["conn_"], dtype=np.int64))
    assert np.array_equal(m.dofs, np.array(g["dofs_"]))
    for name in SETS:
        assert np.array_equal(getattr(m, name), np.array(g[name + "_"])), name
    for name in CORNERS:
        assert getattr(m, name) == g[name + "_"], name
    assert np.array_equal(m.elementsMiddleLayer(), np.array(g["elementsMiddleLayer_"]))
    assert np.array_equal(m.nodesPeriodic, np.array(g["nodesPeriodic_"]))
    assert np.array_equal(m.dofsPeriodic, np.array(g["dofsPeriodic_"]))


def areas(coor, conn):
    x, y = coor[conn, 0], coor[conn, 1]
    return 0.5 * np.sum(x * np.roll(y, -1, axis=1) - np.roll(x, -1, axis=1) * y, axis=1)


@pytest.mark.parametrize("nelx,nely,nfine", [(1, 1, 1), (3, 3, 1), (9, 18, 1), (27, 27, 1), (27, 60, 3),
                                              (36, 20, 2), (81, 100, 1), (12, 7, 5), (6, 40, 1)])
def test_finelayer_invariants(nelx, nely, nfine):
    h = 0.5
    m = gf.Mesh.Quad4.FineLayer(nelx, nely, h, nfine)
    coor, conn = m.coor, m.conn
    assert conn.shape == (m.nelem, 4) and coor.shape == (m.nnode, 2)
    assert np.array_equal(np.unique(conn), np.arange(m.nnode))  # every node used, none invented
    a = areas(coor, conn)
    assert np.all(a > 0)  # counter-clockwise, non-degenerate
    Lx, Ly = coor[:, 0].max(), coor[:, 1].max()
    assert np.isclose(Lx, nelx * h) and np.isclose(Ly, m.nely * h) and np.isclose(a.sum(), Lx * Ly)
    # conforming: every edge is shared by two elements or lies on the boundary
    edges = {}
    for e in conn:
        for k in range(4):
            key = tuple(sorted((e[k], e[(k + 1) % 4])))
            edges[key] = edges.get(key, 0) + 1
    assert max(edges.values()) <= 2
    boundary = [k for k, v in edges.items() if v == 1]
    on = lambda n: np.isclose(coor[n, 0], 0) or np.isclose(coor[n, 0], Lx) or np.isclose(coor[n, 1], 0) or np.isclose(coor[n, 1], Ly)
    assert all(on(a_) and on(b_) for a_, b_ in boundary)
    # the middle layer is the finest (element size h), spans the width, and sits at mid-height
    mid = m.elementsMiddleLayer()
    assert mid.size == nelx and np.allclose(a[mid], h * h)
    yc = coor[conn[mid], 1].mean()
    assert abs(yc - Ly / 2) <= h
    # edges and periodic pairs are consistent
    assert np.allclose(coor[m.nodesLeftEdge, 0], 0) and np.allclose(coor[m.nodesRightEdge, 0], Lx)
    assert np.allclose(coor[m.nodesBottomEdge, 1], 0) and np.allclose(coor[m.nodesTopEdge, 1], Ly)
    P = m.nodesPeriodic
    d = coor[P[:, 1]] - coor[P[:, 0]]
    assert np.all(np.isclose(np.abs(d[:, 0]), Lx) | np.isclose(d[:, 0], 0))
    assert np.all(np.isclose(np.abs(d[:, 1]), Ly) | np.isclose(d[:, 1], 0))
    assert P.shape[0] == 3 + m.nodesBottomOpenEdge.size + m.nodesLeftOpenEdge.size


def test_tri3_regular_golden(fixtures):
    # tests/basic/MeshTri3.cpp:6-88, and the orientation helpers (MeshTri3.h:240-326)
    g = fixtures["MeshTri3_Regular_5_3"]
    m = gf.Mesh.Tri3.Regular(5, 3)
    # nelem_ is written "15 * 2" in the reference test (MeshTri3.cpp:21)
    assert (m.nelem, m.nnode, m.nne, m.ndim) == (30, g["nnode_"], g["nne_"], g["ndim_"])
    assert np.allclose(m.coor, np.array(g["coor_"]))
    assert np.array_equal(m.conn, np.array(g["conn_"], dtype=np.int64))
    assert np.array_equal(m.dofs, np.array(g["dofs_"]))
    for name in SETS:
        assert np.array_equal(getattr(m, name), np.array(g[name + "_"])), name
    for name in CORNERS:
        assert getattr(m, name) == g[name + "_"], name
    assert np.array_equal(m.nodesPeriodic, np.array(g["nodesPeriodic_"]))
    assert np.array_equal(m.dofsPeriodic, np.array(g["dofsPeriodic_"]))
    # k = (x0 - x1) x (x2 - x1) is negative for the mesh's counter-clockwise triangles: -1
    o = gf.Mesh.Tri3.getOrientation(m.coor, m.conn)
    assert np.all(o == -1)
    flipped = gf.Mesh.Tri3.setOrientation(m.coor, m.conn, +1)
    assert np.all(gf.Mesh.Tri3.getOrientation(m.coor, flipped) == 1)
    assert np.array_equal(flipped[:, 0], m.conn[:, 0]) and np.array_equal(flipped[:, 1], m.conn[:, 2])
    assert np.array_equal(gf.Mesh.Tri3.setOrientation(m.coor, flipped, -1), m.conn)
    assert np.array_equal(gf.Mesh.Tri3.setOrientation(m.coor, m.conn, o, -1), m.conn)
