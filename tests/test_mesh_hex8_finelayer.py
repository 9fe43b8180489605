"""Mesh::Hex8::FineLayer (host set-up, gf_mesh_hex8_finelayer*): the reference's golden mesh
(tests/basic/MeshHex8.cpp:192-377, FineLayer(3, 3, 3): coordinates, connectivity, every node set,
dofs, dofsPeriodic, nodesPeriodic) entry for entry, and, for meshes with transition layers (which
the reference's tests do not cover), the invariants every FineLayer mesh satisfies: positive
Jacobians, element volumes summing to the box volume, conforming faces whose unshared faces tile
the box surface, every node used once, node sets equal to the nodes on their planes in the
reference's order, periodic pairs that differ by box vectors and tie every node of the back, right
and top faces, and a middle layer of nelx * nelz elements centred at Ly / 2."""

import itertools
import json
import os

import numpy as np
import pytest

import goosefem_amd as gf

FIX = os.path.join(os.path.dirname(__file__), "golden", "reference_fixtures.json")

# Hex8 local nodes (ElementHex8.h:20-32): the z = -1 face counter-clockwise, then z = +1
XI = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]],
              float)
FACES = [(0, 1, 2, 3), (4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7)]


def shape_gradients(p):
    """dN/dxi [npt, 8, 3] of the trilinear shape functions at local points p [npt, 3]."""
    f = 1.0 + p[:, None, :] * XI[None, :, :]  # [npt, 8, 3]
    out = np.empty(f.shape)
    for k in range(3):
        o = [j for j in range(3) if j != k]
        out[:, :, k] = XI[None, :, k] * f[:, :, o[0]] * f[:, :, o[1]] / 8.0
    return out


def test_hex8_finelayer_golden():
    fx = json.load(open(FIX))["MeshHex8_FineLayer_3_3_3"]
    m = gf.Mesh.Hex8.FineLayer(3, 3, 3)
    assert len(fx) == 49
    for key, want in fx.items():
        got = getattr(m, key[:-1])
        assert np.array_equal(np.asarray(got), np.asarray(want)), key


CASES = [(9, 17, 27, 1), (27, 40, 9, 1), (9, 31, 27, 3), (6, 20, 9, 1), (27, 60, 27, 1), (9, 18, 9, 2), (1, 5, 1, 1)]


@pytest.mark.parametrize("nelx,nely,nelz,nfine", CASES)
def test_hex8_finelayer_invariants(nelx, nely, nelz, nfine):
    m = gf.Mesh.Hex8.FineLayer(nelx, nely, nelz, 1.0, nfine)
    c, e = m.coor, m.conn
    Lx, Ly, Lz = float(nelx), float(c[:, 1].max()), float(nelz)
    assert (m.nelx, m.nelz, m.nely) == (nelx, nelz, int(Ly))
    assert m.nelem == len(e) and m.nnode == len(c)
    assert len(np.unique(np.round(c, 9), axis=0)) == len(c)
    assert np.array_equal(np.unique(e), np.arange(m.nnode))
    # Jacobians at the 2x2x2 Gauss points (exact volume for trilinear maps) and near the corners
    g = np.array(list(itertools.product([-1.0, 1.0], repeat=3))) / np.sqrt(3.0)
    pts = np.vstack([g, 0.999 * XI])
    J = np.einsum("pak,eaj->epjk", shape_gradients(pts), c[e])
    det = np.linalg.det(J)
    assert det.min() > 0
    assert abs(det[:, :8].sum() - Lx * Ly * Lz) < 1e-9 * Lx * Ly * Lz
    # conforming: no face shared by more than two elements; unshared faces tile the box surface
    faces = e[:, np.array(FACES)].reshape(-1, 4)
    _, first, cnt = np.unique(np.sort(faces, axis=1), axis=0, return_index=True, return_counts=True)
    assert cnt.max() == 2
    area = 0.0
    for f in faces[first[cnt == 1]]:
        P = c[f]
        on = [a for a, L in enumerate((Lx, Ly, Lz)) if np.allclose(P[:, a], 0.0) or np.allclose(P[:, a], L)]
        assert on, "unshared face inside the box"
        q = np.delete(P, on[0], axis=1)
        # the face is a planar quad: its area is half the cross product of its diagonals
        d1, d2 = q[2] - q[0], q[3] - q[1]
        area += 0.5 * abs(d1[0] * d2[1] - d1[1] * d2[0])
    assert abs(area - 2.0 * (Lx * Ly + Ly * Lz + Lz * Lx)) < 1e-9 * area
    tol = 1e-9
    planes = {"Front": (2, 0.0), "Back": (2, Lz), "Left": (0, 0.0), "Right": (0, Lx), "Bottom": (1, 0.0),
              "Top": (1, Ly)}
    for name, (a, v) in planes.items():
        s = getattr(m, "nodes" + name)
        assert sorted(s) == list(np.where(abs(c[:, a] - v) < tol)[0]), name
        if name in ("Bottom", "Top"):
            key = [(c[i, 0], c[i, 2]) for i in s]
            assert key == sorted(key), name  # x outer, z inner
        else:
            assert list(s) == sorted(s), name
    # an edge is the intersection of its two faces; an open edge drops the two corners
    fb = np.intersect1d(m.nodesFront, m.nodesBottom)
    assert np.array_equal(m.nodesFrontBottomEdge, fb)
    assert np.array_equal(m.nodesBottomFrontEdge, fb)
    assert np.array_equal(m.nodesFrontBottomOpenEdge, fb[1:-1])
    assert m.nodesBackTopRightCorner == int(np.where((abs(c - [Lx, Ly, Lz]) < tol).all(axis=1))[0][0])
    t = m.nodesPeriodic
    d = c[t[:, 1]] - c[t[:, 0]]
    for a, L in enumerate((Lx, Ly, Lz)):
        assert np.all((abs(d[:, a]) < tol) | (abs(d[:, a] - L) < tol))
    dep = set(t[:, 1].tolist())
    tied = (abs(c[:, 0] - Lx) < tol) | (abs(c[:, 1] - Ly) < tol) | (abs(c[:, 2] - Lz) < tol)
    assert len(dep) == len(t) and dep == set(np.where(tied)[0].tolist())
    assert set(t[:, 0].tolist()).isdisjoint(dep)
    assert m.dofsPeriodic.max() + 1 == 3 * (m.nnode - len(t))
    ml = m.elementsMiddleLayer()
    assert len(ml) == nelx * nelz
    assert np.allclose(c[e[ml]][:, :, 1].mean(axis=1), Ly / 2)
    # the middle layer is the finest: its elements are unit cubes
    assert np.allclose(det[ml, :8].sum(axis=1), 1.0)


def test_hex8_finelayer_unknown_set():
    m = gf.Mesh.Hex8.FineLayer(3, 3, 3)
    with pytest.raises(Exception):
        m.nodesFrontBackEdge
