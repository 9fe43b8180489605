"""Tied-system PCG on the GPU: the matrix-free K' and K'_up of MatrixPartitionedTyings against a
dense restatement built from the oracle's element stiffness, the reordered [u, p, d] dofs and the
tying matrices; the solve against a direct dense solve, with the true residual as the criterion."""

import numpy as np
import pytest

import oracle
from oracle import HEX8, QUAD4_PLANAR

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gf():
    import goosefem_amd

    goosefem_amd.lib()
    return goosefem_amd


def host(t):
    import torch

    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def periodic_system(gf, et, n, seed=1, square=False):
    if et == HEX8:
        mesh = gf.Mesh.Hex8.Regular(n, n, n)
    else:
        mesh = gf.Mesh.Quad4.Regular(n if square else n + 2, n)
    coor, conn = mesh.coor, mesh.conn
    ctrl = gf.Tyings.Control(coor, mesh.dofs)
    iip = np.concatenate((ctrl.controlDofs.ravel(), ctrl.dofs[mesh.nodesOrigin]))
    ty = gf.Tyings.Periodic(ctrl.coor, ctrl.dofs, ctrl.controlDofs, mesh.nodesPeriodic, iip)
    vec = gf.Vector(conn, ty.dofs)
    nd = coor.shape[1]
    ov = oracle.Vector(conn, oracle.dofs(coor.shape[0], nd))
    x = ov.AsElement(coor)
    oq = oracle.Quadrature(et, x)
    quad = (gf.Element.Hex8.Quadrature(x) if et == HEX8 else gf.Element.Quad4.QuadraturePlanar(x))
    rng = np.random.default_rng(seed)
    G = np.exp(0.4 * rng.standard_normal(oq.nelem))
    K = np.full(oq.nelem, 10.0)
    P = gf.TyingsPCG(vec, quad, K, G, ty)
    # dense reference: A in the [u, p, d] numbering, then the reference's condensation
    C = oracle.elastic_tangent(np.repeat(K[:, None], oq.nip, 1), np.repeat(G[:, None], oq.nip, 1))
    Ke = oq.Int_gradN_dot_tensor4_dot_gradNT_dV(C)
    ndof = ty.nni + ty.nnd
    A = np.zeros((ndof, ndof))
    for e in range(oq.nelem):
        d = ty.dofs[conn[e]].ravel()
        A[np.ix_(d, d)] += Ke[e]
    u, p, dd = slice(0, ty.nnu), slice(ty.nnu, ty.nni), slice(ty.nni, ndof)
    Cdu, Cdp = ty.Cdu, ty.Cdp
    Kp = A[u, u] + A[u, dd] @ Cdu + Cdu.T @ A[dd, u] + Cdu.T @ A[dd, dd] @ Cdu
    Kup = A[u, p] + A[u, dd] @ Cdp + Cdu.T @ A[dd, p] + Cdu.T @ A[dd, dd] @ Cdp
    return mesh, ty, P, Kp, Kup


@pytest.mark.parametrize("et,n", [(QUAD4_PLANAR, 6), (HEX8, 3)])
def test_operator_matches_dense_condensation(gf, et, n):
    mesh, ty, P, Kp, Kup = periodic_system(gf, et, n)
    rng = np.random.default_rng(2)
    xu = rng.standard_normal(ty.nnu)
    xp = rng.standard_normal(ty.nnp)
    y = host(P.apply(xu))
    assert np.linalg.norm(y - Kp @ xu) / np.linalg.norm(Kp @ xu) < 1e-13
    y = host(P.apply(None, xp))
    assert np.linalg.norm(y - Kup @ xp) / np.linalg.norm(Kup @ xp) < 1e-13
    assert np.allclose(host(P.inverse_diagonal), 1.0 / np.diag(Kp), rtol=1e-13)
    full = host(P.expand(xu, xp))
    assert np.array_equal(full[: ty.nnu], xu) and np.array_equal(full[ty.nnu:ty.nni], xp)
    assert np.allclose(full[ty.nni:], ty.Cdu @ xu + ty.Cdp @ xp, rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("et,n", [(QUAD4_PLANAR, 10), (HEX8, 4)])
def test_solve_matches_direct_and_true_residual(gf, et, n):
    mesh, ty, P, Kp, Kup = periodic_system(gf, et, n)
    xp = np.zeros(ty.nnp)
    xp[ty.control[0, 1] - ty.nnu] = 1e-3  # simple shear increment through ctrl(0, 1)
    b = -host(P.apply(None, xp))
    assert np.linalg.norm(b + Kup @ xp) / np.linalg.norm(Kup @ xp) < 1e-13
    x, it, res = P.solve(b, tol=1e-10)
    x = host(x)
    xd = np.linalg.solve(Kp, b)
    assert res <= 1e-10 * 10
    assert np.linalg.norm(b - Kp @ x) / np.linalg.norm(b) < 1e-9
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-7
    assert 0 < it < 10 * ty.nnu
    # warm start from the solution: converged at once
    x2, it2, res2 = P.solve(b, x0=xd, tol=1e-10)
    assert it2 == 0 and res2 < 1e-10
    # deterministic: a second solve gives identical bits
    x3, it3, _ = P.solve(b, tol=1e-10)
    assert it3 == it and np.array_equal(host(x3), x)


def test_iteration_cap_reports_nonconvergence(gf):
    mesh, ty, P, Kp, Kup = periodic_system(gf, QUAD4_PLANAR, 10)
    b = np.random.default_rng(0).standard_normal(ty.nnu)
    with pytest.raises(gf.GooseFEMError) as e:
        P.solve(b, tol=1e-12, maxiter=3)
    assert e.value.code == gf.GF_ENOTCONV
    x, it, res = P.solve(b, tol=1e-12, maxiter=3, raise_on_fail=False)
    assert it == 3 and res > 1e-12


def test_structured_periodic_operator_and_solve(gf):
    """Square Quad4 cell (n = 32): the fused structured operator (torus element assembly from one
    shared geometry) against the dense condensation, and its solve against the direct solution."""
    mesh, ty, P, Kp, Kup = periodic_system(gf, QUAD4_PLANAR, 32, square=True)
    assert P.structured
    rng = np.random.default_rng(5)
    xu = rng.standard_normal(ty.nnu)
    y = host(P.apply(xu))
    assert np.linalg.norm(y - Kp @ xu) / np.linalg.norm(Kp @ xu) < 1e-13
    xp = np.zeros(ty.nnp)
    xp[ty.control[0, 1] - ty.nnu] = 1e-3
    b = -Kup @ xp
    # the K'_up x_p term goes through the general path (x_p non-zero)
    assert np.linalg.norm(host(P.apply(None, xp)) - Kup @ xp) / np.linalg.norm(Kup @ xp) < 1e-13
    x, it, res = P.solve(b, tol=1e-10)
    x = host(x)
    xd = np.linalg.solve(Kp, b)
    assert res <= 1e-9 and np.linalg.norm(b - Kp @ x) / np.linalg.norm(b) < 1e-9
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-6
    x2, it2, _ = P.solve(b, tol=1e-10)
    assert it2 == it and np.array_equal(host(x2), x)
    # the fused two-launch iteration honours the cap and warm starts
    _, it3, res3 = P.solve(b, tol=1e-12, maxiter=3, raise_on_fail=False)
    assert it3 == 3 and res3 > 1e-12
    _, it4, res4 = P.solve(b, x0=xd, tol=1e-10)
    assert it4 == 0 and res4 < 1e-10
    # the non-square cell of the other tests stays on the general path
    _, _, P2, _, _ = periodic_system(gf, QUAD4_PLANAR, 6)
    assert not P2.structured


def test_structured_periodic_hex8_operator_and_solve(gf):
    """Hex8 cube (n = 8): the structured operator (K_e = K KK + G KG from one shared geometry,
    node-centric sums over the 8 elements of each node) against the dense condensation, and its solve
    against the direct solution."""
    mesh, ty, P, Kp, Kup = periodic_system(gf, HEX8, 8)
    assert P.structured
    rng = np.random.default_rng(6)
    xu = rng.standard_normal(ty.nnu)
    y = host(P.apply(xu))
    assert np.linalg.norm(y - Kp @ xu) / np.linalg.norm(Kp @ xu) < 1e-13
    xp = np.zeros(ty.nnp)
    xp[ty.control[0, 1] - ty.nnu] = 5e-4
    xp[ty.control[1, 0] - ty.nnu] = 5e-4
    b = -Kup @ xp
    x, it, res = P.solve(b, tol=1e-10)
    x = host(x)
    xd = np.linalg.solve(Kp, b)
    assert res <= 1e-9 and np.linalg.norm(b - Kp @ x) / np.linalg.norm(b) < 1e-9
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-6
    x2, it2, _ = P.solve(b, tol=1e-10)
    assert it2 == it and np.array_equal(host(x2), x)
    _, it3, res3 = P.solve(b, tol=1e-12, maxiter=3, raise_on_fail=False)
    assert it3 == 3 and res3 > 1e-12
    _, it4, res4 = P.solve(b, x0=xd, tol=1e-10)
    assert it4 == 0 and res4 < 1e-10


@pytest.mark.parametrize("et,n,square", [(QUAD4_PLANAR, 32, True), (QUAD4_PLANAR, 10, False), (HEX8, 8, False)])
def test_bad_start_falls_back(gf, et, n, square):
    """A learned (or any) start whose residual is not finite or larger than ||b|| is replaced by the
    previous converged solution, if that one does better than zero, else by zero: NaN, 1e6-scaled
    and random starts reach the true residual 1e-10 in the plain solve's iteration count (the zero
    fallback repeats the plain solve bit for bit); a non-finite right-hand side fails at once."""
    mesh, ty, P, Kp, Kup = periodic_system(gf, et, n, square=square)
    xp = np.zeros(ty.nnp)
    xp[ty.control[0, 1] - ty.nnu] = 1e-3
    b = -Kup @ xp
    xd = np.linalg.solve(Kp, b)
    rng = np.random.default_rng(11)
    starts = {
        "nan": np.full(ty.nnu, np.nan),
        "scaled": 1e6 * xd,
        "random": np.abs(xd).max() * rng.standard_normal(ty.nnu),
    }
    x_plain = it_plain = None
    for name, x0 in starts.items():
        # a fresh solver has no previous solution: the fallback is zero, i.e. the plain solve
        _, _, Q, _, _ = periodic_system(gf, et, n, square=square)
        if x_plain is None:
            xz, it_plain, _ = Q.solve(b, tol=1e-10)
            x_plain = host(xz)
            assert Q.start_used == gf.GF_START_GIVEN
            _, _, Q, _, _ = periodic_system(gf, et, n, square=square)
        x, it, res = Q.solve(b, x0=x0, tol=1e-10)
        x = host(x)
        assert Q.start_used == gf.GF_START_ZERO, name
        assert res <= 1e-10 and np.linalg.norm(b - Kp @ x) / np.linalg.norm(b) < 1e-9, name
        assert it == it_plain and np.array_equal(x, x_plain), name
    # with a previous solution: a doubled load from a NaN start restarts from that solution, whose
    # residual is half of ||b||, and needs no more iterations than the plain solve
    P.solve(b, tol=1e-10)
    x, it, res = P.solve(2.0 * b, x0=np.full(ty.nnu, np.nan), tol=1e-10)
    assert P.start_used == gf.GF_START_PREVIOUS
    assert res <= 1e-10 and it <= it_plain
    assert np.linalg.norm(host(x) - 2.0 * xd) / np.linalg.norm(2.0 * xd) < 1e-6
    # a good start is kept
    P.solve(b, x0=xd + 1e-6 * np.abs(xd).max() * rng.standard_normal(ty.nnu), tol=1e-10)
    assert P.start_used == gf.GF_START_GIVEN
    # a non-finite right-hand side is rejected before any iteration
    bad = b.copy()
    bad[3] = np.inf
    with pytest.raises(gf.GooseFEMError) as e:
        P.solve(bad, tol=1e-10)
    assert e.value.code == gf.GF_EINVAL
