"""GPU parity of Element::Quad4::QuadratureAxisymmetric (gf_axi_*, ElementQuad4Axisymmetric.h:33-467)
with the oracle's restatement, bit for bit, on distorted meshes of a hollow cylinder with the Gauss
and nodal rules; and a thick-walled cylinder under internal pressure solved on the GPU
(QuadratureAxisymmetric + MatrixPartitioned + the partitioned solve) against Lame's closed form."""

import numpy as np
import pytest

import oracle
from oracle import QUAD4

pytestmark = pytest.mark.gpu

L, R1, R2 = 2.0, 0.5, 1.5


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


def cylinder(nx, ny, distort=0.0, seed=0):
    coor, conn = oracle.quad4_regular(nx, ny)
    zr = np.stack((coor[:, 0] * L / nx, R1 + coor[:, 1] * (R2 - R1) / ny), axis=1)
    if distort:
        inner = (coor[:, 0] > 0) & (coor[:, 0] < nx) & (coor[:, 1] > 0) & (coor[:, 1] < ny)
        rng = np.random.default_rng(seed)
        zr[inner] += distort * rng.uniform(-1, 1, (inner.sum(), 2)) * np.array([L / nx, (R2 - R1) / ny])
    return zr, conn, oracle.dofs(coor.shape[0], 2)


@pytest.mark.parametrize("nodal", [False, True])
@pytest.mark.parametrize("nx,ny,distort", [(7, 5, 0.3), (1, 1, 0.0), (40, 33, 0.2)])
def test_axisymmetric_bit_exact(gf, nodal, nx, ny, distort):
    zr, conn, dofs = cylinder(nx, ny, distort)
    x = oracle.Vector(conn, dofs).AsElement(zr)
    xi, w = oracle.quad_rule(QUAD4, nodal)
    oq = oracle.QuadratureAxisymmetric(x, xi, w)
    gq = gf.Element.Quad4.QuadratureAxisymmetric(x, xi, w)
    assert np.array_equal(host(gq.dV), oq.dV())
    assert np.array_equal(host(gq.B()), oq.B())
    rng = np.random.default_rng(1)
    ue = rng.standard_normal((oq.nelem, 4, 2))
    assert np.array_equal(host(gq.GradN_vector(ue)), oq.GradN_vector(ue))
    assert np.array_equal(host(gq.GradN_vector_T(ue)), oq.GradN_vector_T(ue))
    assert np.array_equal(host(gq.SymGradN_vector(ue)), oq.SymGradN_vector(ue))
    rho = rng.uniform(0.5, 2.0, (oq.nelem, oq.nip))
    assert np.array_equal(host(gq.Int_N_scalar_NT_dV(rho)), oq.Int_N_scalar_NT_dV(rho))
    sig = rng.standard_normal((oq.nelem, oq.nip, 3, 3))
    assert np.array_equal(host(gq.Int_gradN_dot_tensor2_dV(sig)), oq.Int_gradN_dot_tensor2_dV(sig))
    C = rng.standard_normal((oq.nelem, oq.nip, 3, 3, 3, 3))  # a general tensor: every slot pairing counts
    assert np.array_equal(host(gq.Int_gradN_dot_tensor4_dot_gradNT_dV(C)), oq.Int_gradN_dot_tensor4_dot_gradNT_dV(C))
    # update_x recomputes the geometry
    x2 = x * 1.1
    gq.update_x(x2)
    assert np.array_equal(host(gq.dV), oracle.QuadratureAxisymmetric(x2, xi, w).dV())


def test_thick_cylinder_under_pressure(gf):
    # plane strain (u_z = 0 at both ends), internal pressure p: Lame's displacement
    # u_r = (1 + nu) p R1^2 / (E (R2^2 - R1^2)) ((1 - 2 nu) r + R2^2 / r)
    nz, nr = 4, 40
    zr, conn, dofs = cylinder(nz, nr)
    mesh_nodes = np.arange(zr.shape[0])
    ends = mesh_nodes[np.isclose(zr[:, 0], 0.0) | np.isclose(zr[:, 0], L)]
    inner = mesh_nodes[np.isclose(zr[:, 1], R1)]
    iip = np.sort(dofs[ends, 0])
    vector = gf.VectorPartitioned(conn, dofs, iip)
    quad = gf.Element.Quad4.QuadratureAxisymmetric(vector.AsElement(zr))
    K0, G0, p = 10.0, 3.0, 0.01
    mat = gf.Elastic2d(np.full((quad.nelem, quad.nip), K0), np.full((quad.nelem, quad.nip), G0))
    A = gf.MatrixPartitioned(vector)
    A.assemble(quad.Int_gradN_dot_tensor4_dot_gradNT_dV(mat.C))
    # consistent nodal forces of the pressure on the inner face: p 2 pi R1 dz / 2 per node and edge
    fext = np.zeros_like(zr)
    zi = np.sort(zr[inner, 0])
    for a, b in zip(zi[:-1], zi[1:]):
        for zz in (a, b):
            n = inner[np.isclose(zr[inner, 0], zz)][0]
            fext[n, 1] += p * 2 * np.pi * R1 * (b - a) / 2
    u = host(gf.MatrixPartitionedSolver(tol=1e-12).Solve(A, fext, np.zeros_like(zr)))
    E = 9 * K0 * G0 / (3 * K0 + G0)
    nu = (3 * K0 - 2 * G0) / (2 * (3 * K0 + G0))
    r = zr[:, 1]
    exact = (1 + nu) * p * R1**2 / (E * (R2**2 - R1**2)) * ((1 - 2 * nu) * r + R2**2 / r)
    assert np.max(np.abs(u[:, 1] - exact)) < 2e-3 * np.max(np.abs(exact))
    assert np.max(np.abs(u[:, 0])) < 1e-9 * np.max(np.abs(exact))  # plane strain: u_z = 0
    # equilibrium: the internal force balances the pressure on the unknown DOFs
    vector2 = gf.Vector(conn, dofs)
    mat.Eps = quad.SymGradN_vector(vector2.AsElement(u))
    mat.refresh()
    fint = host(vector2.AssembleNode(quad.Int_gradN_dot_tensor2_dV(mat.Sig)))
    res = (fext - fint)[:, 1]
    assert np.max(np.abs(res)) < 1e-9 * np.max(np.abs(fext))
