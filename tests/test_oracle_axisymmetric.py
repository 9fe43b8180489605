"""Pins the CPU restatement of Element::Quad4::QuadratureAxisymmetric
(oracle.QuadratureAxisymmetric, ElementQuad4Axisymmetric.h:33-467).  The reference has no test of
this element, so the restatement is pinned by what the axisymmetric formulation fixes: the exact
volume of a hollow cylinder (also on distorted meshes of it), patch tests of linear displacement
fields (radial expansion gives eps_rr = eps_tt = a), force balance in z for any stress, the
stiffness as the derivative of the internal force (K u = f_int(C : eps(u))), and the lumped mass."""

import numpy as np

import oracle
from oracle import QUAD4

L, R1, R2 = 2.0, 0.5, 1.5


def cylinder(nx, ny, distort=0.0, seed=0):
    coor, conn = oracle.quad4_regular(nx, ny)
    z = coor[:, 0] * L / nx
    r = R1 + coor[:, 1] * (R2 - R1) / ny
    zr = np.stack((z, r), axis=1)
    if distort:
        inner = (coor[:, 0] > 0) & (coor[:, 0] < nx) & (coor[:, 1] > 0) & (coor[:, 1] < ny)
        rng = np.random.default_rng(seed)
        zr[inner] += distort * rng.uniform(-1, 1, (inner.sum(), 2)) * np.array([L / nx, (R2 - R1) / ny])
    dofs = oracle.dofs(coor.shape[0], 2)
    return zr, conn, dofs


def quad(zr, conn, dofs):
    return oracle.QuadratureAxisymmetric(oracle.Vector(conn, dofs).AsElement(zr))


def test_volume_exact():
    for distort in (0.0, 0.3):
        q = quad(*cylinder(6, 5, distort))
        assert np.isclose(q.dV().sum(), np.pi * (R2**2 - R1**2) * L, rtol=1e-13)
        assert np.all(q.dV() > 0)
    # the lumped mass of rho = 1 carries the volume once per displacement component
    q = quad(*cylinder(4, 3))
    M = q.Int_N_scalar_NT_dV(np.ones((q.nelem, q.nip)))
    assert np.isclose(M.sum(), 2 * np.pi * (R2**2 - R1**2) * L, rtol=1e-13)
    assert np.allclose(M, M.transpose(0, 2, 1))


def test_patch_tests():
    zr, conn, dofs = cylinder(5, 4, distort=0.3)
    q = quad(zr, conn, dofs)
    vec = oracle.Vector(conn, dofs)
    a, b, c = 1e-3, -2e-3, 5e-4
    u = np.zeros_like(zr)
    u[:, 1] = a * zr[:, 1]  # radial expansion u_r = a r
    u[:, 0] = b * zr[:, 0] + c * zr[:, 1]  # u_z = b z + c r
    ue = vec.AsElement(u)
    g = q.GradN_vector(ue)
    assert np.allclose(g[..., 0, 0], a, atol=1e-15) and np.allclose(g[..., 1, 1], a, atol=1e-15)
    assert np.allclose(g[..., 2, 2], b, atol=1e-15)
    assert np.allclose(g[..., 0, 2], c, atol=1e-15) and np.allclose(g[..., 2, 0], 0.0, atol=1e-15)
    gT = q.GradN_vector_T(ue)
    assert np.allclose(gT[..., 2, 0], c, atol=1e-15) and np.allclose(gT[..., 0, 2], 0.0, atol=1e-15)
    e = q.SymGradN_vector(ue)
    assert np.allclose(e[..., 0, 2], c / 2, atol=1e-15) and np.array_equal(e[..., 0, 2], e[..., 2, 0])
    assert np.all(e[..., 0, 1] == 0) and np.all(e[..., 1, 2] == 0)
    # B() holds the five non-zero components
    B = q.B()
    assert np.count_nonzero(B[0, 0, 0]) == 5


def test_force_balance_and_stiffness_consistency():
    zr, conn, dofs = cylinder(4, 3, distort=0.25, seed=1)
    q = quad(zr, conn, dofs)
    rng = np.random.default_rng(2)
    sig = rng.standard_normal((q.nelem, q.nip, 3, 3))
    fe = q.Int_gradN_dot_tensor2_dV(sig)
    # a rigid translation along z does no work: the z-forces of each element sum to zero
    assert np.allclose(fe[..., 0].sum(axis=1), 0.0, atol=1e-13)
    K = rng.uniform(5, 15, (q.nelem, q.nip))
    G = rng.uniform(0.5, 1.5, (q.nelem, q.nip))
    C = oracle.elastic_tangent(K, G)
    Ke = q.Int_gradN_dot_tensor4_dot_gradNT_dV(C)
    ue = rng.standard_normal((q.nelem, 4, 2))
    eps = q.SymGradN_vector(ue)
    f = q.Int_gradN_dot_tensor2_dV(np.einsum("eqijkl,eqlk->eqij", C, eps))
    assert np.allclose(np.einsum("eab,eb->ea", Ke, ue.reshape(q.nelem, 8)), f.reshape(q.nelem, 8), rtol=1e-12,
                       atol=1e-12 * np.abs(f).max())
    assert np.allclose(Ke, Ke.transpose(0, 2, 1), atol=1e-12 * np.abs(Ke).max())
    # positive semi-definite with exactly one zero-energy mode per element (rigid z-translation)
    w = np.linalg.eigvalsh(Ke)
    assert np.all(w[:, 0] > -1e-10 * w[:, -1]) and np.all(w[:, 1] > 1e-8 * w[:, -1])
