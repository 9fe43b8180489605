"""Host mesh helpers against the reference's golden values (tests/basic/Mesh.cpp): edgesize on
Regular(2, 2, h = 10) (:190-197), centers on Regular(2, 2, h = 2) (:207-213) and elemmap2nodemap
for three element maps of Regular(3, 3) (:236-258), and FineLayer::roll with elemmap2nodemap on
FineLayer(3, 3) (examples 1 and 2, :260-382: rolls 0..5 permute element and node fields)."""

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


@pytest.mark.parametrize("example", [1, 2])
def test_finelayer_roll_elemmap2nodemap(example):
    fx = FIX[f"Mesh_elemmap2nodemap_FineLayer_3_3_example{example}"]
    m = gf.Mesh.Quad4.FineLayer(3, 3)
    want = {0: "", 1: "_r1", 2: "_r2"}
    ev, nv = np.array(fx["elemval"]), np.array(fx["nodeval"])
    for n in range(6):
        elemmap = m.roll(n)
        nodemap = gf.Mesh.elemmap2nodemap(elemmap, m.coor, m.conn)
        assert np.array_equal(ev[elemmap], fx["elemval" + want[n % 3]]), n
        assert np.array_equal(nv[nodemap], fx["nodeval" + want[n % 3]]), n


def test_finelayer_roll_is_a_layerwise_permutation():
    m = gf.Mesh.Quad4.FineLayer(27, 40)
    c = gf.Mesh.centers(m.coor, m.conn)
    r = m.roll(1)
    assert np.array_equal(np.sort(r), np.arange(m.nelem))
    assert np.allclose(c[r, 1], c[:, 1])  # elements stay in their layer
    assert np.array_equal(m.roll(0), np.arange(m.nelem))
    x0 = m.coor[m.conn[0], 0]
    coarse = int(round(x0.max() - x0.min()))  # the bottom layer's element width
    assert coarse > 1
    assert np.array_equal(m.roll(m.nelx // coarse), np.arange(m.nelem))  # a full turn
