"""Pin the oracle's mesh bookkeeping against the reference's golden arrays (bit-exact integers).

Golden data: tests/golden/reference_fixtures.json, extracted from the reference's own tests
(tests/basic/MeshQuad4.cpp, MeshHex8.cpp, Mesh.cpp, tests/test_Mesh.py).
"""

import numpy as np

import oracle


def test_quad4_regular_golden(fixtures):
    g = fixtures["MeshQuad4_Regular_5_3"]
    coor, conn = oracle.quad4_regular(5, 3)
    assert np.array_equal(coor, np.array(g["coor_"]))
    assert np.array_equal(conn, np.array(g["conn_"], dtype=np.int64))
    assert np.array_equal(oracle.dofs(coor.shape[0], 2), np.array(g["dofs_"]))
    for name in oracle.QUAD4_NODESETS:
        assert np.array_equal(oracle.quad4_nodeset(5, 3, name), np.array(g[name + "_"])), name
    ties = oracle.quad4_nodes_periodic(5, 3)
    assert np.array_equal(ties, np.array(g["nodesPeriodic_"]))
    assert np.array_equal(oracle.dofs_periodic(coor.shape[0], 2, ties), np.array(g["dofsPeriodic_"]))
    assert g["nodesOrigin_"] == 0 and g["nodesBottomRightCorner_"] == 5


def test_quad4_regular_pytest_golden():
    # tests/test_MeshQuad4.py:12-68 builds conn from a node grid; restate that construction
    nodes = np.arange(16).reshape(4, 4)
    conn = np.hstack((nodes[:-1, :-1].reshape(-1, 1), nodes[:-1, 1:].reshape(-1, 1),
                      nodes[1:, 1:].reshape(-1, 1), nodes[1:, :-1].reshape(-1, 1)))
    coor, c = oracle.quad4_regular(3, 3)
    assert np.array_equal(c, conn)
    assert np.array_equal(coor[:, 0], np.tile(np.arange(4.0), 4))


def test_hex8_regular_golden(fixtures):
    g = fixtures["MeshHex8_Regular_3_3_3"]
    coor, conn = oracle.hex8_regular(3, 3, 3)
    assert np.array_equal(coor, np.array(g["coor_"]))
    assert np.array_equal(conn, np.array(g["conn_"], dtype=np.int64))
    assert np.array_equal(oracle.dofs(coor.shape[0], 3), np.array(g["dofs_"]))
    ties = oracle.hex8_nodes_periodic(3, 3, 3)
    assert np.array_equal(ties, np.array(g["nodesPeriodic_"]))
    assert np.array_equal(oracle.dofs_periodic(coor.shape[0], 3, ties), np.array(g["dofsPeriodic_"]))


def test_hex8_periodic_counts():
    # every boundary node except the origin is dependent exactly once
    for n in ((3, 3, 3), (4, 3, 5), (6, 6, 6)):
        ties = oracle.hex8_nodes_periodic(*n)
        nx, ny, nz = n
        assert ties.shape[0] == (nx + 1) * (ny + 1) * (nz + 1) - nx * ny * nz
        assert len(np.unique(ties[:, 1])) == ties.shape[0]
        assert not np.intersect1d(ties[:, 0], ties[:, 1]).size


def test_coordination_elem2node_golden(fixtures):
    coor, conn = oracle.quad4_regular(3, 3)
    assert np.array_equal(oracle.coordination(conn), np.array(fixtures["Mesh_coordination_Regular_3_3"]))
    e2n = oracle.elem2node(conn)
    assert [list(map(int, r)) for r in e2n] == fixtures["Mesh_elem2node_Regular_3_3"]


def test_periodic_bookkeeping_golden(fixtures):
    coor, conn = oracle.quad4_regular(4, 4)
    ties = oracle.quad4_nodes_periodic(4, 4)
    dofs = oracle.dofs_periodic(coor.shape[0], 2, ties)
    real = np.array(fixtures["test_Mesh_coordination_dofs_4_4"]).ravel()
    assert np.array_equal(oracle.coordination(dofs), real)
    n2d = oracle.node2dof(dofs)
    assert [list(map(int, r)) for r in n2d] == fixtures["test_Mesh_node2dof_4_4"]
    nt = oracle.nodaltyings(dofs)
    assert [list(map(int, r)) for r in nt] == fixtures["test_Mesh_nodaltyings_4_4"]
    e2n = np.array(oracle.elem2node(conn, dofs))
    real = np.sort(np.array(fixtures["test_Mesh_elem2node_periodic_4_4"]), axis=1)
    assert np.array_equal(e2n, real)


def test_renumber_docstring_example():
    # Mesh.h:99-110: [[0,1],[5,4]] -> [[0,1],[3,2]]
    assert np.array_equal(oracle.renumber(np.array([[0, 1], [5, 4]])), np.array([[0, 1], [3, 2]]))


def test_periodic_dofs_brute_force():
    # dofsPeriodic = union-find of ties, numbered by first appearance of the sorted representative
    for nx, ny in ((5, 3), (7, 4), (16, 9)):
        coor, conn = oracle.quad4_regular(nx, ny)
        ties = oracle.quad4_nodes_periodic(nx, ny)
        d = np.arange(coor.shape[0] * 2).reshape(-1, 2)
        for i, j in ties:
            d[j] = d[i]
        _, inv = np.unique(d, return_inverse=True)
        assert np.array_equal(oracle.dofs_periodic(coor.shape[0], 2, ties), inv.reshape(d.shape))
        # periodic DOF count: independent nodes = nx * ny
        assert oracle.dofs_periodic(coor.shape[0], 2, ties).max() + 1 == 2 * nx * ny
