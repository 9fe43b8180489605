"""Host mesh helpers against the reference's golden values (tests/basic/Mesh.cpp): edgesize on
Regular(2, 2, h = 10) (:190-197), centers on Regular(2, 2, h = 2) (:207-213) and elemmap2nodemap
for three element maps of Regular(3, 3) (:236-258).  The FineLayer examples of that file need
FineLayer::roll, which is not provided."""

import json
import os

import numpy as np
import pytest

import goosefem_amd as gf

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_fixtures.json")))


def test_edgesize():
    m = gf.Mesh.Quad4.Regular(2, 2, 10.0)
    s = gf.Mesh.edgesize(m.coor, m.conn)
    assert s.shape == (4, 4) and np.allclose(s, 10.0)
    # a distorted element: edges by hand
    coor = np.array([[0.0, 0.0], [3.0, 0.0], [3.0, 4.0], [0.0, 1.0]])
    assert np.allclose(gf.Mesh.edgesize(coor, [[0, 1, 2, 3]]), [[3.0, 4.0, np.hypot(3.0, 3.0), 1.0]])


def test_centers():
    m = gf.Mesh.Quad4.Regular(2, 2, 2.0)
    assert np.array_equal(gf.Mesh.centers(m.coor, m.conn), np.array(FIX["Mesh_centers_Regular_2_2_h2"]))


def test_elemmap2nodemap():
    fx = FIX["Mesh_elemmap2nodemap_Regular_3_3"]
    m = gf.Mesh.Quad4.Regular(3, 3)
    for k in range(3):
        got = gf.Mesh.elemmap2nodemap(fx[f"elmap{k}"], m.coor, m.conn)
        assert np.array_equal(got, fx[f"nodemap{k}"]), k


def test_helpers_reject_other_elements():
    m = gf.Mesh.Hex8.Regular(2, 2, 2)
    for f in (gf.Mesh.edgesize, gf.Mesh.centers):
        with pytest.raises(NotImplementedError):
            f(m.coor, m.conn)
