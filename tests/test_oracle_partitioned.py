"""Pins the CPU restatement of VectorPartitioned / MatrixDiagonalPartitioned / MatrixPartitioned
(oracle.VectorPartitioned, oracle.MatrixDiagonalPartitioned, oracle.MatrixPartitioned) against the
reference's own tests: the golden arrays of tests/basic/VectorPartitioned.cpp:10-38 and
tests/basic/MatrixDiagonalPartitioned.cpp:12-53, the round trips of VectorPartitioned.cpp:40-95,
MatrixDiagonalPartitioned.cpp:55-94 and MatrixParitioned.cpp:11-108, and the identities that fix the
rest (a permutation round trip, agreement with the unpartitioned Vector, MatrixDiagonal and Matrix
restatements, dense products).  The per-row summation order of MatrixPartitioned's products (first
block's columns, then the second's, one running sum) follows Eigen's evaluation of the reference's
expressions; no reference output pins those last bits, the tests pin the products to rounding."""

import numpy as np

import oracle
from oracle import QUAD4


def tied_dirichlet_9x9(seed=0):
    """Regular(9, 9) with the right open edge's DOFs tied to the left open edge's, and the top and
    bottom edges prescribed (x then y), as in VectorPartitioned.cpp:44-60."""
    coor, conn = oracle.quad4_regular(9, 9)
    dofs = oracle.dofs(coor.shape[0], 2)
    left = oracle.quad4_nodeset(9, 9, "nodesLeftOpenEdge")
    right = oracle.quad4_nodeset(9, 9, "nodesRightOpenEdge")
    top = oracle.quad4_nodeset(9, 9, "nodesTopEdge")
    bottom = oracle.quad4_nodeset(9, 9, "nodesBottomEdge")
    dofs[right] = dofs[left]
    iip = np.concatenate((dofs[top, 0], dofs[top, 1], dofs[bottom, 0], dofs[bottom, 1]))
    u = np.random.default_rng(seed).standard_normal(coor.shape)
    u[right] = u[left]
    return coor, conn, dofs, iip, u


def test_dofs_is_p_golden(fixtures):
    coor, conn = oracle.quad4_regular(3, 3)
    v = oracle.VectorPartitioned(conn, oracle.dofs(coor.shape[0], 2), np.arange(8))
    is_p = np.array(fixtures["VectorPartitioned_is_p_Regular_3_3"])
    assert np.array_equal(v.dofs_is_p(), is_p)
    assert np.array_equal(v.dofs_is_u(), ~is_p)
    assert v.nnp == 8 and v.nnu == 24


def test_lumped_mass_golden(fixtures):
    g = fixtures["MatrixDiagonalPartitioned_lumped_mass_Regular_2_2"]
    coor, conn = oracle.quad4_regular(2, 2)
    dofs = oracle.dofs(coor.shape[0], 2)
    A = oracle.MatrixDiagonalPartitioned(conn, dofs, g["iip"])
    xi, w = oracle.quad_rule(QUAD4, True)
    q = oracle.Quadrature(QUAD4, oracle.Vector(conn, dofs).AsElement(coor), xi, w)
    mass = q.Int_N_scalar_NT_dV(np.ones((q.nelem, q.nip)))
    A.assemble(mass)
    assert np.allclose(A.data(), g["data"], atol=1e-12)
    # the partitioned sum is the unpartitioned one, DOF by DOF, bit for bit
    assert np.array_equal(A.data(), oracle.diag_assemble(conn, dofs, mass))


def test_vector_round_trips():
    coor, conn, dofs, iip, u = tied_dirichlet_9x9()
    v = oracle.VectorPartitioned(conn, dofs, iip)
    assert v.nnp == iip.size and v.nnu == np.unique(dofs).size - iip.size  # tied DOFs leave holes
    u_u, u_p = v.AsDofs_u(u), v.AsDofs_p(u)
    assert np.allclose(v.NodeFromPartitioned(u_u, u_p), u)  # VectorPartitioned.cpp:62-64
    w = np.zeros_like(u)
    assert np.allclose(v.Copy_p(u, v.Copy_u(u, w)), u)  # VectorPartitioned.cpp:90-94
    # dofval: a permutation round trip, and the element/dofval forms agree with the plain Vector
    plain = oracle.Vector(conn, dofs)
    d = plain.AsDofs(u)
    assert np.array_equal(v.AsDofs_u(d), u_u) and np.array_equal(v.AsDofs_p(d), u_p)
    assert np.array_equal(v.DofsFromParitioned(u_u, u_p), d)
    assert np.array_equal(v.ElementFromPartitioned(u_u, u_p), plain.AsElement(d))
    e = plain.AsElement(d)
    assert np.array_equal(v.AsDofs_u(e), u_u) and np.array_equal(v.AsDofs_p(e), u_p)
    # iiu is ascending, iip keeps the caller's order, part numbers the unknowns first
    assert np.all(np.diff(v.iiu) > 0) and np.array_equal(v.iip, iip)
    assert np.array_equal(np.unique(v.part[v.dofs_is_u()]), np.arange(v.nnu))
    assert np.array_equal(np.unique(v.part[v.dofs_is_p()]), v.nnu + np.arange(v.nnp))
    for k, d in enumerate(iip):  # prescribed DOF iip[k] is numbered nnu + k
        assert np.all(v.part[dofs == d] == v.nnu + k)


def test_matrix_dot_solve_reaction():
    # MatrixDiagonalPartitioned.cpp:55-94 (dot, solve), and Reaction = A_pp x_p on the prescribed DOFs
    coor, conn = oracle.quad4_regular(2, 2)
    dofs = oracle.dofs(coor.shape[0], 2)
    iip = np.array([0, 2])
    rng = np.random.default_rng(1)
    a = rng.random(dofs.size)
    x = rng.random(dofs.size)
    A = oracle.MatrixDiagonalPartitioned(conn, dofs, iip)
    A.set(a)
    B = A.Dot(x)
    assert np.array_equal(B, a * x) and np.array_equal(A.data(), a)
    X = A.Solve(B, np.zeros_like(B))
    X[iip] = x[iip]
    assert np.allclose(X, x)
    R = A.Reaction(x, np.zeros_like(x))
    assert np.array_equal(R[iip], a[iip] * x[iip]) and np.all(np.delete(R, iip) == 0.0)
    v = A.v
    assert np.array_equal(A.Dot_u(x[v.iiu], x[v.iip]), a[v.iiu] * x[v.iiu])
    assert np.array_equal(A.Reaction_p(x[v.iiu], x[v.iip]), a[v.iip] * x[v.iip])
    assert np.array_equal(A.Solve_u(B[v.iiu], x[v.iip]), (1.0 / a[v.iiu]) * B[v.iiu])
    # nodevec forms equal the dofval forms through the plain Vector
    plain = oracle.Vector(conn, dofs)
    xn = plain.AsNode(x)
    assert np.array_equal(A.Dot(xn), plain.AsNode(a * x))


def random_symmetric(rng, nelem, k, spd=False):
    a = rng.random((nelem, k, k))
    a = 0.5 * (a + a.transpose(0, 2, 1))
    if spd:
        a = a @ a.transpose(0, 2, 1) + k * np.eye(k)
    return a


def test_matrix_partitioned_blocks_are_the_unpartitioned_entries():
    # every block entry is the unpartitioned Matrix entry (Matrix.h:704-727) of its DOF pair, bit for
    # bit (both sum the same triplets in the same (e, m, i, n, j) order), and nothing is lost
    coor, conn, dofs, iip, _ = tied_dirichlet_9x9()
    rng = np.random.default_rng(2)
    K = conn.shape[1] * dofs.shape[1]
    em = random_symmetric(rng, conn.shape[0], K)
    A = oracle.MatrixPartitioned(conn, dofs, iip)
    A.assemble(em)
    rowptr, col, val = oracle.matrix_assemble(conn, dofs, em)
    full = {}
    for r in range(rowptr.size - 1):
        for k in range(rowptr[r], rowptr[r + 1]):
            full[(r, col[k])] = val[k]
    v = A.v
    rows = (v.iiu, v.iiu, v.iip, v.iip)
    cols = (v.iiu, v.iip, v.iiu, v.iip)
    n = 0
    for b in range(4):
        for (r, c), a in A.blocks[b].items():
            assert a == full[(rows[b][r], cols[b][c])]
        n += len(A.blocks[b])
    assert n == len(full)
    assert np.array_equal(A.Todense(), oracle_dense(rowptr, col, val, v.ndof))


def oracle_dense(rowptr, col, val, n):
    out = np.zeros((n, n))
    for r in range(n):
        out[r, col[rowptr[r]:rowptr[r + 1]]] = val[rowptr[r]:rowptr[r + 1]]
    return out


def test_matrix_partitioned_reference_solve():
    # tests/basic/MatrixParitioned.cpp:11-44: Regular(2, 2), iip {0, 5, 7, 13}, random symmetric
    # element matrices; Solve(A, Dot(x), x) returns x
    coor, conn = oracle.quad4_regular(2, 2)
    dofs = oracle.dofs(coor.shape[0], 2)
    rng = np.random.default_rng(0)
    A = oracle.MatrixPartitioned(conn, dofs, np.array([0, 5, 7, 13]))
    A.assemble(random_symmetric(rng, conn.shape[0], 8))
    x = rng.random(dofs.size)
    b = A.Dot(x)
    assert np.allclose(b, A.Todense() @ x, rtol=1e-14, atol=1e-14)
    assert np.allclose(A.Solve(b, x), x)


def test_matrix_partitioned_dense_matrix():
    # tests/basic/MatrixParitioned.cpp:46-108 fill a 10 x 10 A = a + a^T on dofs arange(10) with
    # iip {0, 2, 4}; the same matrix enters here as one 5-node element's matrix
    rng = np.random.default_rng(4)
    a = rng.random((10, 10))
    Ad = a + a.T
    conn = np.arange(5).reshape(1, 5)
    dofs = np.arange(10).reshape(5, 2)
    K = oracle.MatrixPartitioned(conn, dofs, np.array([0, 2, 4]))
    K.assemble(Ad[None])
    assert np.array_equal(K.Todense(), Ad)
    x = rng.random(10)
    assert np.allclose(K.Dot(x), Ad @ x, rtol=1e-14, atol=1e-14)
    assert np.allclose(K.Solve(K.Dot(x), x), x)
    xn = x.reshape(5, 2)
    assert np.allclose(K.Dot(xn), (Ad @ x).reshape(5, 2), rtol=1e-14, atol=1e-14)
    assert np.allclose(K.Solve(K.Dot(xn), xn), xn)
    # Reaction: the prescribed rows of A x; Dot_u / Dot_p are the row blocks of A x
    v = K.v
    R = K.Reaction(x, np.zeros(10))
    assert np.allclose(R[v.iip], (Ad @ x)[v.iip]) and np.all(R[v.iiu] == 0.0)
    assert np.allclose(K.Dot_u(x[v.iiu], x[v.iip]), (Ad @ x)[v.iiu])
    assert np.allclose(K.Reaction_p(x[v.iiu], x[v.iip]), (Ad @ x)[v.iip])
