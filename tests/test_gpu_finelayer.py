"""The device kernels on non-regular meshes: Mesh::Quad4::FineLayer (MeshQuad4.h:200-1135) has
transition layers of trapezoidal elements (true Jacobians), nodes shared by three to five elements
and CSR rows of varying length.  Quadrature, vector conversions, CSR assembly and products and the
explicit step are bit-exact with the oracle on it; the partitioned solve matches a direct solve."""

import numpy as np
import pytest

import oracle
from oracle import QUAD4_PLANAR

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


@pytest.mark.parametrize("nelx,nely,nfine", [(27, 40, 1), (9, 18, 1), (36, 21, 3)])
def test_finelayer_element_path_bit_exact(gf, nelx, nely, nfine):
    m = gf.Mesh.Quad4.FineLayer(nelx, nely, 1.0, nfine)
    coor, conn, dofs = m.coor, m.conn, m.dofs
    ov = oracle.Vector(conn, dofs)
    x = ov.AsElement(coor)
    oq = oracle.Quadrature(QUAD4_PLANAR, x)
    gv = gf.Vector(conn, dofs)
    gq = gf.Element.Quad4.QuadraturePlanar(gv.AsElement(coor))
    assert np.array_equal(host(gq.dV), oq.dV())
    assert len(np.unique(np.round(oq.dV(), 12))) > 2  # elements of several sizes and shapes
    rng = np.random.default_rng(0)
    u = rng.standard_normal(coor.shape) * 1e-2
    ue = ov.AsElement(u)
    assert np.array_equal(host(gv.AsElement(u)), ue)
    eps = oq.SymGradN_vector(ue)
    assert np.array_equal(host(gq.SymGradN_vector(ue)), eps)
    K = np.full((oq.nelem, oq.nip), 10.0)
    G = np.exp(0.3 * rng.standard_normal((oq.nelem, oq.nip)))
    sig = oracle.elastic_stress(eps, K, G)
    fe = oq.Int_gradN_dot_tensor2_dV(sig)
    assert np.array_equal(host(gq.Int_gradN_dot_tensor2_dV(sig)), fe)
    assert np.array_equal(host(gv.AssembleNode(fe)), ov.AssembleNode(fe))
    Ke = oq.Int_gradN_dot_tensor4_dot_gradNT_dV(oracle.elastic_tangent(K, G))
    assert np.array_equal(host(gq.Int_gradN_dot_tensor4_dot_gradNT_dV(oracle.elastic_tangent(K, G))), Ke)
    # CSR assembly with rows of different lengths, and the product
    rp, cl, val = oracle.matrix_assemble(conn, dofs, Ke)
    A = gf.Matrix(conn, dofs)
    A.assemble(Ke)
    grp, gcl, gval = (host(t) for t in A.csr())
    assert np.array_equal(grp, rp) and np.array_equal(gcl, cl) and np.array_equal(gval, val)
    assert len(np.unique(np.diff(rp))) >= 3
    xd = rng.standard_normal(dofs.size)
    assert np.array_equal(host(A.Dot(xd)), oracle.csr_dot(rp, cl, val, xd))


def test_finelayer_explicit_bit_exact(gf):
    m = gf.Mesh.Quad4.FineLayer(27, 30)
    coor, conn, dofs = m.coor, m.conn, m.dofs
    ov = oracle.Vector(conn, dofs)
    x = ov.AsElement(coor)
    q = oracle.Quadrature(QUAD4_PLANAR, x)
    xi, w = oracle.quad_rule(QUAD4_PLANAR, True)
    qn = oracle.Quadrature(QUAD4_PLANAR, x, xi, w)
    inv = oracle.diag_inverse(oracle.diag_assemble(conn, dofs, qn.Int_N_scalar_NT_dV(np.ones((q.nelem, 4)))))
    inv[dofs[m.nodesBottomEdge].ravel()] = 0.0
    inv[dofs[m.nodesTopEdge].ravel()] = 0.0
    rng = np.random.default_rng(1)
    Ke = np.full(q.nelem, 10.0)
    Ge = np.exp(0.3 * rng.standard_normal(q.nelem)).astype(np.float32).astype(np.float64)
    u = rng.standard_normal(dofs.size) * 1e-3
    u[dofs[m.nodesBottomEdge].ravel()] = 0.0
    u[dofs[m.nodesTopEdge].ravel()] = 0.0
    v = np.zeros_like(u)
    v[dofs[m.nodesTopEdge, 0]] = 1e-4
    a = np.zeros_like(u)
    X = gf.ExplicitQuad4(gf.Vector(conn, dofs), coor, Ke, Ge, inv, 0.01)
    assert not X.structured
    X.set_state(u, v, a)
    fint = oracle.vv_steps(5, QUAD4_PLANAR, conn, dofs, q.dNx, q.dV(), Ke, Ge, inv, 0.01, u, v, a)
    X.step(5, gf.F64, store_fint=True)
    gu, gvel, ga = (host(t) for t in X.state())
    assert np.array_equal(gu, u) and np.array_equal(gvel, v) and np.array_equal(ga, a)
    assert np.array_equal(host(X.fint()), fint)


def test_finelayer_partitioned_statics(gf):
    # elastic plate on FineLayer(27, 40): bottom fixed, top displaced; the device CG solve against the
    # oracle's direct solve of the same partitioned system
    m = gf.Mesh.Quad4.FineLayer(27, 40)
    coor, conn, dofs = m.coor, m.conn, m.dofs
    iip = np.concatenate((dofs[m.nodesBottomEdge].ravel(), dofs[m.nodesTopEdge].ravel()))
    oq = oracle.Quadrature(QUAD4_PLANAR, oracle.Vector(conn, dofs).AsElement(coor))
    rng = np.random.default_rng(2)
    G = np.exp(0.3 * rng.standard_normal((oq.nelem, oq.nip)))
    Ke = oq.Int_gradN_dot_tensor4_dot_gradNT_dV(oracle.elastic_tangent(np.full_like(G, 10.0), G))
    oA = oracle.MatrixPartitioned(conn, dofs, iip)
    oA.assemble(Ke)
    gA = gf.MatrixPartitioned(conn, dofs, iip)
    gA.assemble(Ke)
    for b in range(4):
        assert np.array_equal(host(gA.csr(b)[2]), oA.csr(b)[2])
    disp = np.zeros_like(coor)
    disp[m.nodesTopEdge, 0] = 0.05
    disp[m.nodesTopEdge, 1] = -0.02
    ref = oA.Solve(np.zeros_like(disp), disp)
    got = host(gf.MatrixPartitionedSolver(tol=1e-12).Solve(gA, np.zeros_like(disp), disp))
    assert np.max(np.abs(got - ref)) <= 1e-8 * np.max(np.abs(ref))


@pytest.mark.parametrize("nelx,nely,nelz", [(9, 17, 27), (27, 24, 9)])
def test_hex8_finelayer_element_path_bit_exact(gf, nelx, nely, nelz):
    """Mesh::Hex8::FineLayer (MeshHex8.h:613-2296): transition layers of wedge-like hexahedra."""
    from oracle import HEX8

    m = gf.Mesh.Hex8.FineLayer(nelx, nely, nelz)
    coor, conn, dofs = m.coor, m.conn, m.dofs
    ov = oracle.Vector(conn, dofs)
    oq = oracle.Quadrature(HEX8, ov.AsElement(coor))
    gv = gf.Vector(conn, dofs)
    gq = gf.Element.Hex8.Quadrature(gv.AsElement(coor))
    assert np.array_equal(host(gq.dV), oq.dV())
    assert len(np.unique(np.round(oq.dV(), 12))) > 2
    assert abs(oq.dV().sum() - np.prod(coor.max(axis=0))) < 1e-9 * np.prod(coor.max(axis=0))
    rng = np.random.default_rng(1)
    ue = ov.AsElement(rng.standard_normal(coor.shape) * 1e-2)
    eps = oq.SymGradN_vector(ue)
    assert np.array_equal(host(gq.SymGradN_vector(ue)), eps)
    K = np.full((oq.nelem, oq.nip), 10.0)
    G = np.exp(0.3 * rng.standard_normal((oq.nelem, oq.nip)))
    fe = oq.Int_gradN_dot_tensor2_dV(oracle.elastic_stress(eps, K, G))
    assert np.array_equal(host(gq.Int_gradN_dot_tensor2_dV(oracle.elastic_stress(eps, K, G))), fe)
    assert np.array_equal(host(gv.AssembleNode(fe)), ov.AssembleNode(fe))
    C = oracle.elastic_tangent(K, G)
    assert np.array_equal(host(gq.Int_gradN_dot_tensor4_dot_gradNT_dV(C)), oq.Int_gradN_dot_tensor4_dot_gradNT_dV(C))
