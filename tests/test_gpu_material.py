"""GPU parity of the material laws (gf_plastic_*, gf_material_elastic_*) with the oracle, and the
statics solves the reference's examples run with them (docs/examples/statics/
fixed-displacement_elastic.py; periodic_elastoplastic.py's Newton loop on a Dirichlet problem):
VectorPartitioned + Quadrature + material refresh + MatrixPartitioned assembly + the partitioned
solve, all on the device."""

import numpy as np
import pytest

import oracle
from oracle import QUAD4_PLANAR

pytestmark = pytest.mark.gpu

K0, G0, SY, H0 = 10.0, 1.0, 0.05, 0.4


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


def sym(rng, shape, scale):
    a = rng.standard_normal(shape + (3, 3)) * scale
    return 0.5 * (a + np.swapaxes(a, -1, -2))


@pytest.mark.parametrize("shape", [(1000,), (37, 4), (1, 1)])
def test_linear_hardening_bit_exact(gf, shape):
    # two increments from a virgin state, mixed elastic and plastic points, random moduli; every
    # output bit-exact with the oracle (same FP64 expression order, no contraction)
    rng = np.random.default_rng(0)
    K = rng.uniform(5.0, 15.0, shape)
    G = rng.uniform(0.5, 1.5, shape)
    sy = rng.uniform(0.02, 0.1, shape)
    H = rng.uniform(0.0, 1.0, shape)
    mat = gf.LinearHardening(K, G, sy, H)
    ref = oracle.LinearHardening(K, G, sy, H)
    for step in range(2):
        E = sym(rng, shape, 0.05)
        E.reshape(-1, 3, 3)[::3] *= 0.01
        mat.increment()
        ref.increment()
        mat.Eps = E
        ref.Eps = E
        mat.refresh()
        ref.refresh()
        assert np.array_equal(host(mat.Sig), ref.Sig)
        assert np.array_equal(host(mat.C), ref.C)
        assert np.array_equal(host(mat.Epsp), ref.Epsp)
        assert np.array_equal(host(mat.epsp), ref.epsp)
    # refresh without the tangent leaves C alone and gives the same stress
    C_before = host(mat.C)
    mat.Eps = 1.01 * E
    ref.Eps = 1.01 * E
    mat.refresh(compute_tangent=False)
    ref.refresh()
    assert np.array_equal(host(mat.Sig), ref.Sig) and np.array_equal(host(mat.C), C_before)


def test_linear_hardening_large_identities(gf):
    # 2^20 points (8192 + ragged blocks): size-independent identities of the return map, on the device
    import torch

    n = (1 << 20) + 77
    rng = np.random.default_rng(1)
    mat = gf.LinearHardening(np.full(n, K0), G0, SY, H0)
    E = torch.as_tensor(sym(rng, (n,), 0.05), device="cuda")
    mat.Eps = E
    mat.refresh()
    S, ep = mat.Sig, mat.epsp
    tr = S.diagonal(dim1=-2, dim2=-1).sum(-1)
    d = S - tr[:, None, None] * torch.eye(3, dtype=S.dtype, device=S.device) / 3
    q = torch.sqrt(1.5 * (d * d).sum((-1, -2)))
    plastic = ep > 0
    assert plastic.float().mean().item() > 0.9
    assert torch.allclose(q[plastic], SY + H0 * ep[plastic], rtol=1e-12, atol=0)
    # the hydrostatic stress is elastic (absolute tolerance: the deviator's rounding enters the trace)
    assert torch.allclose(tr, 3.0 * K0 * E.diagonal(dim1=-2, dim2=-1).sum(-1), rtol=1e-13,
                          atol=1e-14 * S.abs().max().item())
    Ep = mat.Epsp
    assert Ep.diagonal(dim1=-2, dim2=-1).sum(-1).abs().max().item() < 1e-15
    # a sample checked bit for bit against the oracle
    idx = np.arange(0, n, 4099)
    Sr, _, _, epr = oracle.plastic_refresh(host(E)[idx], K0, G0, SY, H0, np.zeros((idx.size, 3, 3)), np.zeros(idx.size))
    assert np.array_equal(host(S)[idx], Sr) and np.array_equal(host(ep)[idx], epr)


def test_elastic_material_class(gf):
    rng = np.random.default_rng(2)
    K = rng.uniform(5.0, 15.0, (20, 4))
    G = rng.uniform(0.5, 1.5, (20, 4))
    mat = gf.Elastic2d(K, G)
    E = sym(rng, (20, 4), 0.01)
    mat.Eps = E
    mat.refresh()
    assert np.array_equal(host(mat.Sig), oracle.elastic_stress(E, K, G))
    assert np.array_equal(host(mat.C), oracle.elastic_tangent(K, G))


def fixed_displacement_setup(gf, n):
    mesh = gf.Mesh.Quad4.Regular(n, n)
    dofs = mesh.dofs
    iip = np.concatenate((dofs[mesh.nodesRightEdge, 0], dofs[mesh.nodesTopEdge, 1],
                          dofs[mesh.nodesLeftEdge, 0], dofs[mesh.nodesBottomEdge, 1]))
    return mesh, dofs, iip


def test_fixed_displacement_elastic_example(gf):
    # the steps of docs/examples/statics/fixed-displacement_elastic.py with this package: one
    # partitioned solve for the prescribed edge displacements, then the residual on the unknown DOFs
    # vanishes (the example's final check), and the result matches the oracle's direct solve
    mesh, dofs, iip = fixed_displacement_setup(gf, 5)
    vector = gf.VectorPartitioned(mesh.conn, dofs, iip)
    K = gf.MatrixPartitioned(vector)
    solver = gf.MatrixPartitionedSolver()
    elem = gf.Element.Quad4.QuadraturePlanar(vector.AsElement(mesh.coor))
    mat = gf.Elastic2d(np.ones((mesh.nelem, elem.nip)), np.ones((mesh.nelem, elem.nip)))
    disp = np.zeros_like(mesh.coor)
    fext = np.zeros_like(mesh.coor)
    ue = vector.AsElement(disp)
    elem.symGradN_vector(ue, mat.Eps)
    mat.refresh()
    fint = vector.AssembleNode(elem.Int_gradN_dot_tensor2_dV(mat.Sig))
    K.assemble(elem.Int_gradN_dot_tensor4_dot_gradNT_dV(mat.C))
    fres = fext - host(fint)
    disp[mesh.nodesRightEdge, 0] = +0.2
    disp[mesh.nodesTopEdge, 1] = -0.2
    solver.solve(K, fres, disp)
    vector.asElement(disp, ue)
    elem.symGradN_vector(ue, mat.Eps)
    mat.refresh()
    fe = elem.Int_gradN_dot_tensor2_dV(mat.Sig)
    fint = host(vector.AssembleNode(fe))
    vector.copy_p(fint, fext)
    assert np.sum(np.abs(fext - fint)) / np.sum(np.abs(fext)) < 1e-10
    # against the oracle's direct solve of the same system
    x = oracle.Vector(mesh.conn, dofs).AsElement(mesh.coor)
    oq = oracle.Quadrature(QUAD4_PLANAR, x)
    oA = oracle.MatrixPartitioned(mesh.conn, dofs, iip)
    oA.assemble(oq.Int_gradN_dot_tensor4_dot_gradNT_dV(oracle.elastic_tangent(np.ones((mesh.nelem, 4)), np.ones((mesh.nelem, 4)))))
    d0 = np.zeros_like(mesh.coor)
    d0[mesh.nodesRightEdge, 0] = +0.2
    d0[mesh.nodesTopEdge, 1] = -0.2
    ref = oA.Solve(np.zeros_like(d0), d0)
    assert np.max(np.abs(disp - ref)) < 1e-9


def newton_elastoplastic(lib, mesh, dofs, iip, nip, nsteps, on_gpu):
    """Incremental Newton solve of a plate with prescribed edge displacements and a J2 hardening
    material, in the loop structure of periodic_elastoplastic.py:85-140 (Dirichlet instead of
    periodic tyings).  lib is this package (GPU) or the oracle (CPU)."""
    shape = (mesh.nelem, nip)
    rng = np.random.default_rng(3)
    sy = SY * rng.uniform(0.5, 1.5, shape)
    if on_gpu:
        vector = lib.VectorPartitioned(mesh.conn, dofs, iip)
        A = lib.MatrixPartitioned(vector)
        solver = lib.MatrixPartitionedSolver(tol=1e-12)
        elem = lib.Element.Quad4.QuadraturePlanar(vector.AsElement(mesh.coor))
        mat = lib.LinearHardening(np.full(shape, K0), G0, sy, H0)
    else:
        vector = oracle.VectorPartitioned(mesh.conn, dofs, iip)
        plain = oracle.Vector(mesh.conn, dofs)
        A = oracle.MatrixPartitioned(mesh.conn, dofs, iip)
        elem = oracle.Quadrature(QUAD4_PLANAR, plain.AsElement(mesh.coor))
        mat = oracle.LinearHardening(np.full(shape, K0), np.full(shape, G0), sy, np.full(shape, H0))
    disp = np.zeros_like(mesh.coor)
    iters = []
    for inc in range(nsteps):
        mat.increment()
        for it in range(30):
            if on_gpu:
                elem.symGradN_vector(vector.AsElement(disp), mat.Eps)
                mat.refresh()
                fint = host(vector.AssembleNode(elem.Int_gradN_dot_tensor2_dV(mat.Sig)))
                A.assemble(elem.Int_gradN_dot_tensor4_dot_gradNT_dV(mat.C))
            else:
                mat.Eps = elem.SymGradN_vector(plain.AsElement(plain.AsDofs(disp)))
                mat.refresh()
                fint = plain.AsNode(plain.AssembleDofs(elem.Int_gradN_dot_tensor2_dV(mat.Sig)))
                A.assemble(elem.Int_gradN_dot_tensor4_dot_gradNT_dV(mat.C))
            fres = -fint
            if it > 0:
                fres[vector.dofs_is_p()] = 0.0
                if np.sum(np.abs(fres)) / np.sum(np.abs(fint)) < 1e-10:
                    iters.append(it)
                    break
            du = np.zeros_like(disp)
            if it == 0:
                du[mesh.nodesRightEdge, 0] = 0.04
                du[mesh.nodesTopEdge, 1] = -0.02
            if on_gpu:
                solver.solve(A, fres, du)
            else:
                du = A.Solve(fres, du)
            disp += du
        else:
            raise AssertionError("Newton did not converge")
    epsp = host(mat.epsp) if on_gpu else mat.epsp
    return disp, epsp, iters


def test_elastoplastic_newton_against_oracle(gf):
    mesh, dofs, iip = fixed_displacement_setup(gf, 8)
    d_gpu, ep_gpu, it_gpu = newton_elastoplastic(gf, mesh, dofs, iip, 4, 6, True)
    d_cpu, ep_cpu, it_cpu = newton_elastoplastic(oracle, mesh, dofs, iip, 4, 6, False)
    assert ep_cpu.max() > 0.0 and np.mean(ep_cpu > 0) > 0.2  # the plate yields
    assert it_gpu == it_cpu  # same Newton path (quadratic convergence of the consistent tangent)
    assert max(it_cpu) <= 6
    assert np.max(np.abs(d_gpu - d_cpu)) <= 1e-9 * np.max(np.abs(d_cpu))
    assert np.max(np.abs(ep_gpu - ep_cpu)) <= 1e-7 * np.max(ep_cpu)


@pytest.mark.parametrize("nx,ny", [(13, 9), (64, 33)])
def test_fused_newton_pass_bit_exact(gf, nx, ny):
    # gf_plastic_newton_q4 equals the unfused chain bit for bit (fe, Ke, Sig, plastic state), over two
    # increments, on a distorted mesh with mixed elastic / plastic points; and the oracle's chain
    coor, conn = oracle.quad4_regular(nx, ny)
    rng = np.random.default_rng(7)
    coor = coor + 0.2 * rng.uniform(-1, 1, coor.shape)
    dofs = oracle.dofs(coor.shape[0], 2)
    vec = gf.Vector(conn, dofs)
    x = vec.AsElement(coor)
    quad = gf.Element.Quad4.QuadraturePlanar(x)
    shape = (conn.shape[0], 4)
    K = rng.uniform(5, 15, shape)
    G = rng.uniform(0.5, 1.5, shape)
    sy = rng.uniform(0.005, 0.02, shape)
    H = rng.uniform(0.0, 1.0, shape)
    fused = gf.LinearHardening(K, G, sy, H)
    nosig = gf.LinearHardening(K, G, sy, H)  # the same pass without the optional stress output
    plain = gf.LinearHardening(K, G, sy, H)
    ref = oracle.LinearHardening(K, G, sy, H)
    oq = oracle.Quadrature(QUAD4_PLANAR, host(x))
    import torch

    for inc in range(2):
        u = rng.standard_normal(coor.shape) * 0.002 * (inc + 1)
        ue = vec.AsElement(u)
        fused.increment()
        nosig.increment()
        plain.increment()
        ref.increment()
        Sig = torch.empty_like(fused.Sig)
        fe, Ke = fused.newton_q4(quad, ue, Sig=Sig)
        fe3, Ke3 = nosig.newton_q4(quad, ue)
        assert np.array_equal(host(fe3), host(fe)) and np.array_equal(host(Ke3), host(Ke))
        assert np.array_equal(host(nosig.Epsp), host(fused.Epsp)) and np.array_equal(host(nosig.epsp), host(fused.epsp))
        quad.symGradN_vector(ue, plain.Eps)
        plain.refresh()
        fe2 = quad.Int_gradN_dot_tensor2_dV(plain.Sig)
        Ke2 = quad.Int_gradN_dot_tensor4_dot_gradNT_dV(plain.C)
        assert np.array_equal(host(fe), host(fe2)) and np.array_equal(host(Ke), host(Ke2))
        assert np.array_equal(host(Sig), host(plain.Sig))
        assert np.array_equal(host(fused.Epsp), host(plain.Epsp)) and np.array_equal(host(fused.epsp), host(plain.epsp))
        ref.Eps = oq.SymGradN_vector(host(ue))
        ref.refresh()
        assert np.array_equal(host(fe), oq.Int_gradN_dot_tensor2_dV(ref.Sig))
        assert np.array_equal(host(Ke), oq.Int_gradN_dot_tensor4_dot_gradNT_dV(ref.C))
    ep = host(fused.epsp)
    assert 0.1 < np.mean(ep > 0) < 0.95
    # the outputs are written as 16-byte streams: an unaligned Sig is a clear error, not a fault
    buf = torch.zeros(Sig.numel() + 1, dtype=torch.float64, device="cuda")
    shifted = buf[1:].view(Sig.shape)
    assert shifted.data_ptr() % 16 != 0
    with pytest.raises(gf.GooseFEMError):
        fused.newton_q4(quad, ue, Sig=shifted)


def test_quadrature_shapes_and_astensor(gf):
    # QuadratureBase shapes, allocation, AsTensor (Element.h:201-420, Allocate.h:158-170) and the
    # write-into forms with numpy outputs, as the examples use them (fixed-displacement_elastic.py)
    mesh = gf.Mesh.Quad4.Regular(3, 2)
    vec = gf.Vector(mesh.conn, mesh.dofs)
    q = gf.Element.Quad4.QuadraturePlanar(vec.AsElement(mesh.coor))
    assert q.shape_elemvec() == (6, 4, 2) and q.shape_elemvec(3) == (6, 4, 3)
    assert q.shape_elemmat() == (6, 8, 8) and q.shape_qscalar() == (6, 4)
    assert q.shape_qvector() == (6, 4, 3) and q.shape_qtensor(2) == (6, 4, 3, 3) and q.shape_qtensor(4, 2) == (6, 4, 2, 2, 2, 2)
    assert tuple(q.allocate_qtensor(2, 0.0).shape) == (6, 4, 3, 3) and float(q.allocate_elemmat(1.5).sum()) == 6 * 64 * 1.5
    dV = host(q.dV)
    T = host(q.AsTensor(2, q.dV))
    assert T.shape == (6, 4, 3, 3) and np.all(T == dV[:, :, None, None])
    M = np.zeros((6, 8, 8))
    q.int_N_scalar_NT_dV(np.ones((6, 4)), M)
    assert np.isclose(M.sum(), 2 * 3 * 2)  # twice the area
    ue = np.zeros((6, 4, 2))
    vec.asElement(mesh.coor, ue)
    assert np.array_equal(ue, host(vec.AsElement(mesh.coor)))
    ax = gf.Element.Quad4.QuadratureAxisymmetric(vec.AsElement(mesh.coor + np.array([0.0, 1.0])))
    assert ax.shape_qtensor(2) == (6, 4, 3, 3) and tuple(ax.AsTensor(2, ax.dV).shape) == (6, 4, 3, 3)
