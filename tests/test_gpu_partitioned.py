"""GPU parity of VectorPartitioned / MatrixDiagonalPartitioned / MatrixPartitioned (gf_vecpart_*,
gf_diagpart_*, gf_matpart_*) with the oracle's restatement: bit-exact for gathers, scatters, per-DOF
products, the sparse blocks and their products, on the reference's own tied Dirichlet setup
(tests/basic/VectorPartitioned.cpp:40-95) and on a periodic Hex8 cube, plus the reference's golden
arrays and round trips; the iterative solve within a stated tolerance of the oracle's direct one."""

import numpy as np
import pytest

import oracle
from oracle import HEX8, QUAD4, QUAD4_PLANAR

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gf():
    import torch

    import goosefem_amd

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    goosefem_amd.lib()
    return goosefem_amd


def host(t):
    import torch

    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def quad4_tied():
    coor, conn = oracle.quad4_regular(9, 9)
    dofs = oracle.dofs(coor.shape[0], 2)
    left = oracle.quad4_nodeset(9, 9, "nodesLeftOpenEdge")
    right = oracle.quad4_nodeset(9, 9, "nodesRightOpenEdge")
    top = oracle.quad4_nodeset(9, 9, "nodesTopEdge")
    bottom = oracle.quad4_nodeset(9, 9, "nodesBottomEdge")
    dofs[right] = dofs[left]
    iip = np.concatenate((dofs[top, 0], dofs[top, 1], dofs[bottom, 0], dofs[bottom, 1]))
    return coor, conn, dofs, iip


def hex8_periodic():
    coor, conn = oracle.hex8_regular(4, 4, 4)
    dofs = oracle.dofs_periodic(coor.shape[0], 3, oracle.hex8_nodes_periodic(4, 4, 4))
    iip = np.array([5, 0, 17, 3, 40, 41])  # an arbitrary, unsorted prescribed set
    return coor, conn, dofs, iip


@pytest.mark.parametrize("setup", [quad4_tied, hex8_periodic])
def test_vector_partitioned_bit_exact(gf, setup):
    coor, conn, dofs, iip = setup()
    ov = oracle.VectorPartitioned(conn, dofs, iip)
    gv = gf.VectorPartitioned(conn, dofs, iip)
    assert (gv.nnu, gv.nnp) == (ov.nnu, ov.nnp)
    assert np.array_equal(gv.iiu, ov.iiu) and np.array_equal(gv.iip, ov.iip)
    assert np.array_equal(gv.dofs_is_u(), ov.dofs_is_u()) and np.array_equal(gv.dofs_is_p(), ov.dofs_is_p())
    rng = np.random.default_rng(3)
    nd = dofs.shape[1]
    dv = rng.standard_normal(ov.ndof)
    nv = rng.standard_normal((coor.shape[0], nd))  # not consistent on tied nodes: last writer wins
    ev = rng.standard_normal((conn.shape[0], conn.shape[1], nd))
    for arg in (dv, nv, ev):
        assert np.array_equal(host(gv.AsDofs_u(arg)), ov.AsDofs_u(arg))
        assert np.array_equal(host(gv.AsDofs_p(arg)), ov.AsDofs_p(arg))
    u, p = rng.standard_normal(ov.nnu), rng.standard_normal(ov.nnp)
    assert np.array_equal(host(gv.DofsFromParitioned(u, p)), ov.DofsFromParitioned(u, p))
    assert np.array_equal(host(gv.NodeFromPartitioned(u, p)), ov.NodeFromPartitioned(u, p))
    assert np.array_equal(host(gv.ElementFromPartitioned(u, p)), ov.ElementFromPartitioned(u, p))
    dest = rng.standard_normal(nv.shape)
    assert np.array_equal(host(gv.Copy_u(nv, dest)), ov.Copy_u(nv, dest))
    assert np.array_equal(host(gv.Copy_p(nv, dest)), ov.Copy_p(nv, dest))


@pytest.mark.parametrize("setup", [quad4_tied, hex8_periodic])
def test_matrix_diagonal_partitioned_bit_exact(gf, setup):
    coor, conn, dofs, iip = setup()
    oA = oracle.MatrixDiagonalPartitioned(conn, dofs, iip)
    gA = gf.MatrixDiagonalPartitioned(conn, dofs, iip)
    rng = np.random.default_rng(5)
    K = conn.shape[1] * dofs.shape[1]
    elemmat = np.zeros((conn.shape[0], K, K))
    idx = np.arange(K)
    elemmat[:, idx, idx] = rng.uniform(0.5, 2.0, (conn.shape[0], K))
    oA.assemble(elemmat)
    gA.assemble(elemmat)
    assert np.array_equal(host(gA.data), oA.data())
    assert np.array_equal(host(gA.data_uu), oA.Auu) and np.array_equal(host(gA.data_pp), oA.App)
    v = oA.v
    x = rng.standard_normal(v.ndof)
    xn = rng.standard_normal((coor.shape[0], dofs.shape[1]))
    for arg in (x, xn):
        assert np.array_equal(host(gA.Dot(arg)), oA.Dot(arg))
        keep = rng.standard_normal(arg.shape)
        assert np.array_equal(host(gA.Solve(arg, keep)), oA.Solve(arg, keep))
        assert np.array_equal(host(gA.Reaction(arg, keep)), oA.Reaction(arg, keep))
    x_u, x_p = rng.standard_normal(v.nnu), rng.standard_normal(v.nnp)
    assert np.array_equal(host(gA.Dot_u(x_u, x_p)), oA.Dot_u(x_u, x_p))
    assert np.array_equal(host(gA.Dot_p(x_u, x_p)), oA.Dot_p(x_u, x_p))
    assert np.array_equal(host(gA.Reaction_p(x_u, x_p)), oA.Reaction_p(x_u, x_p))
    assert np.array_equal(host(gA.Solve_u(x_u, x_p)), oA.Solve_u(x_u, x_p))
    # set() then data() round trip
    a = rng.uniform(0.5, 2.0, v.ndof)
    oA.set(a)
    gA.set(a)
    assert np.array_equal(host(gA.data), oA.data())


def test_partitioned_reference_golden(gf, fixtures):
    # tests/basic/VectorPartitioned.cpp:10-38 and MatrixDiagonalPartitioned.cpp:12-53 on the GPU
    m = gf.Mesh.Quad4.Regular(3, 3)
    v = gf.VectorPartitioned(m.conn, m.dofs, np.arange(8))
    assert np.array_equal(v.dofs_is_p(), np.array(fixtures["VectorPartitioned_is_p_Regular_3_3"]))
    g = fixtures["MatrixDiagonalPartitioned_lumped_mass_Regular_2_2"]
    m = gf.Mesh.Quad4.Regular(2, 2)
    vec = gf.VectorPartitioned(m.conn, m.dofs, g["iip"])
    A = gf.MatrixDiagonalPartitioned(vec)
    q = gf.Element.Quad4.Quadrature(vec.AsElement(m.coor), gf.Element.Quad4.Nodal.xi(), gf.Element.Quad4.Nodal.w())
    A.assemble(q.Int_N_scalar_NT_dV(np.ones((m.nelem, 4))))
    assert np.allclose(host(A.data), g["data"], atol=1e-12)
    # the reference's solve round trip (MatrixDiagonalPartitioned.cpp:74-93)
    rng = np.random.default_rng(0)
    a, x = rng.random(18), rng.random(18)
    A.set(a)
    B = host(A.Dot(x))
    X = host(A.Solve(B))
    X[g["iip"]] = x[g["iip"]]
    assert np.allclose(B, a * x) and np.allclose(X, x)


def elastic_stiffness(et, conn, coor, seed=0):
    """Element stiffness of a linear elastic body with random shear moduli (FP64, the oracle's)."""
    nd = coor.shape[1]
    x = oracle.Vector(conn, oracle.dofs(coor.shape[0], nd)).AsElement(coor)
    q = oracle.Quadrature(et, x)
    rng = np.random.default_rng(seed)
    G = np.exp(0.4 * rng.standard_normal((q.nelem, q.nip)))
    return q.Int_gradN_dot_tensor4_dot_gradNT_dV(oracle.elastic_tangent(np.full_like(G, 10.0), G))


def random_symmetric(rng, nelem, k, spd=False):
    a = rng.random((nelem, k, k))
    a = 0.5 * (a + a.transpose(0, 2, 1))
    if spd:
        a = a @ a.transpose(0, 2, 1) + k * np.eye(k)
    return a


def no_prescribed():
    coor, conn = oracle.hex8_regular(3, 2, 2)
    return coor, conn, oracle.dofs(coor.shape[0], 3), np.zeros(0, dtype=np.int64)


@pytest.mark.parametrize("setup", [quad4_tied, hex8_periodic, no_prescribed])
def test_matrix_partitioned_bit_exact(gf, setup):
    # assemble (MatrixPartitioned.h:156-210): the four CSR blocks, Todense, Dot / Dot_u / Dot_p,
    # Reaction / Reaction_p on dofval and nodevec, all bit-exact with the oracle's restatement
    coor, conn, dofs, iip = setup()
    rng = np.random.default_rng(7)
    K = conn.shape[1] * dofs.shape[1]
    em = random_symmetric(rng, conn.shape[0], K)
    oA = oracle.MatrixPartitioned(conn, dofs, iip)
    oA.assemble(em)
    gA = gf.MatrixPartitioned(conn, dofs, iip)
    gA.assemble(em)
    for b in range(4):
        rp, cl, val = (host(t) for t in gA.csr(b))
        orp, ocl, oval = oA.csr(b)
        assert np.array_equal(rp, orp) and np.array_equal(cl, ocl) and np.array_equal(val, oval), b
    assert gA.nnz_blocks == tuple(len(oA.blocks[b]) for b in range(4))
    assert np.array_equal(host(gA.Todense()), oA.Todense())
    v = oA.v
    x = rng.standard_normal(v.ndof)
    xn = rng.standard_normal((coor.shape[0], dofs.shape[1]))
    for arg in (x, xn):
        assert np.array_equal(host(gA.Dot(arg)), oA.Dot(arg))
        keep = rng.standard_normal(arg.shape)
        assert np.array_equal(host(gA.Reaction(arg, keep)), oA.Reaction(arg, keep))
    x_u, x_p = rng.standard_normal(v.nnu), rng.standard_normal(v.nnp)
    if v.nnp:
        assert np.array_equal(host(gA.Dot_u(x_u, x_p)), oA.Dot_u(x_u, x_p))
        assert np.array_equal(host(gA.Dot_p(x_u, x_p)), oA.Dot_p(x_u, x_p))
        assert np.array_equal(host(gA.Reaction_p(x_u, x_p)), oA.Reaction_p(x_u, x_p))
    # the products agree with the dense matrix to rounding
    Ad = oA.Todense()
    assert np.allclose(host(gA.Dot(x)), Ad @ x, rtol=1e-13, atol=1e-12)


@pytest.mark.parametrize("setup,et", [(quad4_tied, QUAD4_PLANAR), (hex8_periodic, HEX8)])
def test_matrix_partitioned_solve(gf, setup, et):
    # MatrixPartitionedSolver::Solve / Solve_u (MatrixPartitioned.h:545-631) on an elastic stiffness
    # with Dirichlet DOFs: x_u = A_uu^-1 (b_u - A_up x_p) by the device CG, against the oracle's
    # direct solve.  Tolerance: CG stops at ||r|| <= 1e-12 ||b||; the error bound is cond(A_uu) times
    # that, cond(A_uu) < 500 for both meshes here, so 1e-7 relative is loose.
    coor, conn, dofs, iip = setup()
    em = elastic_stiffness(et, conn, coor)
    oA = oracle.MatrixPartitioned(conn, dofs, iip)
    oA.assemble(em)
    gA = gf.MatrixPartitioned(conn, dofs, iip)
    gA.assemble(em)
    v = oA.v
    rng = np.random.default_rng(11)
    x = rng.standard_normal(v.ndof)
    b = oA.Dot(x)
    solver = gf.MatrixPartitionedSolver(tol=1e-12)
    for bb, xx in ((b, x), (v.NodeFromPartitioned(v.AsDofs_u(b), v.AsDofs_p(b)), v.NodeFromPartitioned(v.AsDofs_u(x), v.AsDofs_p(x)))):
        ref = oA.Solve(bb, xx)
        got = host(solver.Solve(gA, bb, xx))
        assert solver.relres <= 1e-11 and solver.iterations > 0
        assert np.max(np.abs(got - ref)) <= 1e-7 * np.max(np.abs(ref))
        assert np.max(np.abs(got - xx)) <= 1e-7 * np.max(np.abs(xx))  # the round trip of Dot
    b_u, x_p = v.AsDofs_u(b), v.AsDofs_p(x)
    xu = host(solver.Solve_u(gA, b_u, x_p))
    ref_u = oA.Solve_u(b_u, x_p)
    assert np.max(np.abs(xu - ref_u)) <= 1e-7 * np.max(np.abs(ref_u))
    # warm start from the solution: converged before the first iteration, x_u returned unchanged
    xw = host(solver.Solve_u(gA, b_u, x_p, x_u0=ref_u))
    assert solver.iterations == 0 and np.array_equal(xw, ref_u)


def test_partitioned_solve_rejects_bad_input(gf):
    # a non-finite right-hand side fails at once with GF_EINVAL (before any CG iteration); a warm
    # start with non-finite entries is replaced by the zero start, so the solve equals the cold one
    coor, conn, dofs, iip = quad4_tied()
    gA = gf.MatrixPartitioned(conn, dofs, iip)
    gA.assemble(elastic_stiffness(QUAD4_PLANAR, conn, coor))
    x = np.random.default_rng(12).standard_normal(gA.ndof)
    b = host(gA.Dot(x))
    solver = gf.MatrixPartitionedSolver(tol=1e-12)
    bad = b.copy()
    bad[int(np.asarray(gA.iiu)[0])] = np.nan  # an unknown DOF (a prescribed one is not read)
    with pytest.raises(gf.GooseFEMError) as e:
        solver.Solve(gA, bad, x)
    assert e.value.code == gf.GF_EINVAL
    cold = host(solver.Solve(gA, b, x))
    it_cold = solver.iterations
    assert it_cold > 0
    xw = x.copy()
    xw[np.asarray(gA.iiu)] = np.nan  # the start (the unknown entries); the numbering has holes
    warm = host(solver.Solve(gA, b, xw, warm=True))
    assert solver.iterations == it_cold and np.array_equal(warm, cold)


def test_matrix_partitioned_reference_cases(gf):
    # tests/basic/MatrixParitioned.cpp:11-44 (Regular(2, 2), iip {0, 5, 7, 13}) and :46-108 (one
    # 10 x 10 matrix on dofs arange(10), iip {0, 2, 4}), with symmetric positive definite matrices:
    # the device solve is a CG, which needs A_uu SPD (the reference's LDLT takes any symmetric A)
    rng = np.random.default_rng(0)
    m = gf.Mesh.Quad4.Regular(2, 2)
    A = gf.MatrixPartitioned(m.conn, m.dofs, np.array([0, 5, 7, 13]))
    A.assemble(random_symmetric(rng, m.nelem, 8, spd=True))
    solver = gf.MatrixPartitionedSolver()
    x = rng.random(m.nnode * 2)
    assert np.allclose(host(solver.Solve(A, A.Dot(x), x)), x)
    a = rng.random((10, 10))
    Ad = a @ a.T + 10.0 * np.eye(10)
    K = gf.MatrixPartitioned(np.arange(5).reshape(1, 5), np.arange(10).reshape(5, 2), np.array([0, 2, 4]))
    K.assemble(Ad[None])
    assert np.array_equal(host(K.Todense()), Ad)
    x = rng.random(10)
    assert np.allclose(host(K.Dot(x)), Ad @ x)
    assert np.allclose(host(solver.Solve(K, K.Dot(x), x)), x)
    xn = x.reshape(5, 2)
    assert np.allclose(host(K.Dot(xn)), (Ad @ x).reshape(5, 2))
    assert np.allclose(host(solver.Solve(K, K.Dot(xn), xn)), xn)


def test_matrix_partitioned_fp32_assembly(gf):
    # FP32 element matrices: each entry widened to FP64, then the FP64 assembly (bit-exact with the
    # FP64 path on the widened input)
    coor, conn, dofs, iip = quad4_tied()
    rng = np.random.default_rng(3)
    em32 = random_symmetric(rng, conn.shape[0], 8).astype(np.float32)
    a32 = gf.MatrixPartitioned(conn, dofs, iip)
    a32.assemble(em32)
    a64 = gf.MatrixPartitioned(conn, dofs, iip)
    a64.assemble(em32.astype(np.float64))
    for b in range(4):
        assert np.array_equal(host(a32.csr(b)[2]), host(a64.csr(b)[2]))


def test_reference_python_mesh_partitioned_and_center_of_gravity(gf):
    # tests/test_MeshPartitioned.py:10-39 (MatrixPartitioned vs Matrix Todense; the solve round trip,
    # with SPD element matrices for the device CG) and tests/test_Mesh.py:12-21 (center of gravity)
    mesh = gf.Mesh.Quad4.Regular(2, 2)
    rng = np.random.default_rng(9)
    a = random_symmetric(rng, mesh.nelem, 8, spd=True)
    A = gf.MatrixPartitioned(mesh.conn, mesh.dofs, [0, 5, 7, 13])
    B = gf.Matrix(mesh.conn, mesh.dofs)
    A.assemble(a)
    B.assemble(a)
    assert np.array_equal(host(A.Todense()), host(B.Todense()))
    x = rng.random(mesh.nnode * 2)
    assert np.allclose(host(gf.MatrixPartitionedSolver().Solve(A, A.Dot(x), x)), x)
    m3 = gf.Mesh.Quad4.Regular(3, 3)
    assert np.allclose(host(gf.Mesh.center_of_gravity(m3.coor, m3.conn)), [1.5, 1.5])
    assert np.isclose(float(gf.Mesh.nodal_mass(m3.coor, m3.conn).sum()), 2 * 9.0)
