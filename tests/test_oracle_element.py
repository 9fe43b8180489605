"""Pin the oracle's quadrature / vector / matrix / tyings / CG against closed forms, the reference's
known-answer tests and brute force (numpy einsum, dense assembly, scipy direct solve).

Each check is chosen so that a plausible slip (a transposed operand, a dropped term, a wrong
index) fails it; the brute-force references are written independently of oracle.cpp.
"""

import numpy as np
import pytest

import oracle
from oracle import HEX8, QUAD4, QUAD4_PLANAR


def _mesh(et, n=3):
    if et == HEX8:
        return oracle.hex8_regular(n, n, n)
    return oracle.quad4_regular(n, n)


def _quad(et, n=3, nodal=False, coor=None):
    c, conn = _mesh(et, n)
    if coor is not None:
        c = coor
    vec = oracle.Vector(conn, oracle.dofs(c.shape[0], c.shape[1]))
    xi, w = oracle.quad_rule(et, nodal)
    return c, conn, vec, oracle.Quadrature(et, vec.AsElement(c), xi, w)


def _distorted(et, n, seed=3):
    c, conn = _mesh(et, n)
    rng = np.random.default_rng(seed)
    return c + 0.15 * rng.uniform(-1, 1, c.shape)


# ---- geometry ---------------------------------------------------------------------------------

@pytest.mark.parametrize("et,dv", [(QUAD4, 0.25), (QUAD4_PLANAR, 0.25), (HEX8, 0.125)])
def test_dV_closed_form(et, dv):
    # tests/basic/ElementQuad4.cpp:17-26, ElementQuad4Planar.cpp:17-26, ElementHex8.cpp:19-28
    _, _, _, q = _quad(et)
    assert np.allclose(q.dV(), dv, rtol=1e-15, atol=0)
    # regular h=1 mesh: J = 1/2 I, so dNdx = 2 dN/dxi (up to rounding of the irrational xi)
    assert np.allclose(q.GradN(), np.broadcast_to(2.0 * q.dNxi, q.GradN().shape), rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("et", [QUAD4, QUAD4_PLANAR, HEX8])
def test_gradN_matches_finite_difference_isoparametric_map(et):
    # dN/dx on distorted elements vs central differences of the isoparametric map inverse
    c = _distorted(et, 2)
    _, conn = _mesh(et, 2)
    vec = oracle.Vector(conn, oracle.dofs(c.shape[0], c.shape[1]))
    x = vec.AsElement(c)
    q = oracle.Quadrature(et, x)
    nd = oracle.NDIM[et]
    h = 1e-6
    for e in range(x.shape[0]):
        for p in range(q.nip):
            # J_ij = dx_j / dxi_i via FD of x(xi) = sum_m N_m(xi) x_m
            J = np.zeros((nd, nd))
            for i in range(nd):
                xp = q.xi[p].copy()
                xm = q.xi[p].copy()
                xp[i] += h
                xm[i] -= h
                Np = _shape(et, xp)
                Nm = _shape(et, xm)
                J[i] = (Np @ x[e] - Nm @ x[e]) / (2 * h)
            dNxi = np.zeros((q.nne, nd))
            for i in range(nd):
                xp = q.xi[p].copy()
                xm = q.xi[p].copy()
                xp[i] += h
                xm[i] -= h
                dNxi[:, i] = (_shape(et, xp) - _shape(et, xm)) / (2 * h)
            ref = dNxi @ np.linalg.inv(J).T
            assert np.allclose(q.GradN()[e, p], ref, atol=1e-8)
            assert np.isclose(q.dV()[e, p], np.linalg.det(J) * q.w[p], atol=1e-9)


def _shape(et, xi):
    if oracle.NDIM[et] == 2:
        s = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]])
        return 0.25 * np.prod(1 + s * xi, axis=1)
    s = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                  [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]])
    return 0.125 * np.prod(1 + s * xi, axis=1)


def test_hex8_q1_quirk_equal_on_regular_differs_on_distorted():
    # Element.h:1220 indexes dNxi(m,i) for dNxi(m,j); identical when J is diagonal.  On Regular
    # meshes J's off-diagonal entries are round-off (~1e-16), so the two agree to a few ulp.
    _, _, vec, q = _quad(HEX8, 3)
    qf = oracle.Quadrature(HEX8, q.x, faithful_q1=True)
    assert np.allclose(q.GradN(), qf.GradN(), rtol=4e-15, atol=0)
    c = _distorted(HEX8, 2)
    _, conn = _mesh(HEX8, 2)
    x = oracle.Vector(conn, oracle.dofs(c.shape[0], 3)).AsElement(c)
    assert not np.allclose(oracle.Quadrature(HEX8, x).GradN(),
                           oracle.Quadrature(HEX8, x, faithful_q1=True).GradN())


# ---- patch tests (tests/basic/ElementQuad4*.cpp:41-153, ElementHex8.cpp:43-155) ---------------

@pytest.mark.parametrize("et", [QUAD4, QUAD4_PLANAR, HEX8])
def test_patch_affine_displacement(et):
    c, conn, vec, q = _quad(et)
    nd, td = oracle.NDIM[et], oracle.TDIM[et]
    F = np.zeros((td, td))
    F[0, 1] = 0.1
    disp = c @ F[:nd, :nd].T
    ue = vec.AsElement(disp)
    eps = np.zeros((td, td))
    eps[0, 1] = eps[1, 0] = 0.05
    assert np.allclose(q.GradN_vector(ue), F.T)
    assert np.allclose(q.GradN_vector_T(ue), F)
    assert np.allclose(q.SymGradN_vector(ue), eps)
    # uniform strain on a periodic mesh: zero net internal force (ElementQuad4.cpp:191-218)
    n = 3
    ties = (oracle.hex8_nodes_periodic(n, n, n) if et == HEX8 else oracle.quad4_nodes_periodic(n, n))
    dp = oracle.dofs_periodic(c.shape[0], nd, ties)
    vp = oracle.Vector(conn, dp)
    fe = q.Int_gradN_dot_tensor2_dV(q.SymGradN_vector(ue))
    assert np.allclose(vp.AssembleDofs(fe), 0.0, atol=1e-14)


@pytest.mark.parametrize("et", [QUAD4, QUAD4_PLANAR, HEX8])
def test_interp_and_int_N_vector(et):
    # InterpQuad of a constant field is the field; Int_N_vector_dV + AssembleNode reproduces a
    # constant field on a periodic mesh (tests/basic/ElementQuad4.cpp:155-170)
    c, conn, vec, q = _quad(et)
    nd = oracle.NDIM[et]
    assert np.allclose(q.InterpQuad_vector(vec.AsElement(np.ones_like(c))), 1.0)
    n = 3
    ties = (oracle.hex8_nodes_periodic(n, n, n) if et == HEX8 else oracle.quad4_nodes_periodic(n, n))
    vp = oracle.Vector(conn, oracle.dofs_periodic(c.shape[0], nd, ties))
    qv = np.zeros((q.nelem, q.nip, nd))
    qv[..., 0] = 2.0
    qv[..., 1] = 5.0
    node = vp.AssembleNode(q.Int_N_vector_dV(qv))
    assert np.allclose(node[:, 0], 2.0) and np.allclose(node[:, 1], 5.0)


# ---- brute force (einsum) on random fields and distorted elements ----------------------------

@pytest.mark.parametrize("et", [QUAD4, QUAD4_PLANAR, HEX8])
def test_kernels_vs_einsum(et):
    c = _distorted(et, 3, seed=11)
    _, conn = _mesh(et, 3)
    vec = oracle.Vector(conn, oracle.dofs(c.shape[0], c.shape[1]))
    q = oracle.Quadrature(et, vec.AsElement(c))
    rng = np.random.default_rng(5)
    nd, td = oracle.NDIM[et], oracle.TDIM[et]
    ue = rng.standard_normal((q.nelem, q.nne, nd))
    g, vol = q.GradN(), q.dV()
    grad = np.zeros((q.nelem, q.nip, td, td))
    grad[..., :nd, :nd] = np.einsum("eqmi,emj->eqij", g, ue)
    assert np.allclose(q.GradN_vector(ue), grad, rtol=1e-13, atol=1e-13)
    assert np.allclose(q.GradN_vector_T(ue), np.swapaxes(grad, 2, 3), rtol=1e-13, atol=1e-13)
    assert np.allclose(q.SymGradN_vector(ue), 0.5 * (grad + np.swapaxes(grad, 2, 3)), rtol=1e-13, atol=1e-13)
    sig = rng.standard_normal((q.nelem, q.nip, td, td))
    fe = np.einsum("eqmi,eqij,eq->emj", g, sig[..., :nd, :nd], vol)
    assert np.allclose(q.Int_gradN_dot_tensor2_dV(sig), fe, rtol=1e-12, atol=1e-12)
    C = rng.standard_normal((q.nelem, q.nip) + (td,) * 4)
    Cn = C[..., :nd, :nd, :nd, :nd]
    Ke = np.einsum("eqmi,eqijkl,eqnl,eq->emjnk", g, Cn, g, vol).reshape(q.nelem, q.nne * nd, q.nne * nd)
    assert np.allclose(q.Int_gradN_dot_tensor4_dot_gradNT_dV(C), Ke, rtol=1e-12, atol=1e-12)
    rho = rng.uniform(0.5, 2, (q.nelem, q.nip))
    M = np.einsum("qm,eq,qn,eq->emn", q.N, rho, q.N, vol)
    Mref = np.zeros((q.nelem, q.nne, nd, q.nne, nd))
    for i in range(nd):
        Mref[:, :, i, :, i] = M
    assert np.allclose(q.Int_N_scalar_NT_dV(rho), Mref.reshape(q.nelem, q.nne * nd, -1), rtol=1e-13)


@pytest.mark.parametrize("et", [QUAD4_PLANAR, HEX8])
def test_stiffness_consistency_symmetry_nullspace(et):
    # Ke u_e == Int_gradN_dot_tensor2_dV(C : SymGradN(u_e)) for the isotropic tangent; Ke = Ke^T;
    # rigid-body modes (translations + rotations) are in the null space.
    c = _distorted(et, 2, seed=7)
    _, conn = _mesh(et, 2)
    vec = oracle.Vector(conn, oracle.dofs(c.shape[0], c.shape[1]))
    x = vec.AsElement(c)
    q = oracle.Quadrature(et, x)
    nd = oracle.NDIM[et]
    rng = np.random.default_rng(2)
    K = np.full((q.nelem, q.nip), 10.0)
    G = rng.uniform(0.5, 1.5, (q.nelem, q.nip))
    C = oracle.elastic_tangent(K, G)
    Ke = q.Int_gradN_dot_tensor4_dot_gradNT_dV(C)
    assert np.allclose(Ke, np.swapaxes(Ke, 1, 2), atol=1e-12)
    ue = rng.standard_normal((q.nelem, q.nne, nd))
    fe = q.Int_gradN_dot_tensor2_dV(oracle.elastic_stress(q.SymGradN_vector(ue), K, G))
    assert np.allclose(np.einsum("eab,eb->ea", Ke, ue.reshape(q.nelem, -1)), fe.reshape(q.nelem, -1),
                       atol=1e-12)
    modes = [np.eye(nd)[i] + 0 * x for i in range(nd)]
    rot = np.zeros_like(x)
    rot[..., 0], rot[..., 1] = -x[..., 1], x[..., 0]
    modes.append(rot)
    for mvec in modes:
        r = np.einsum("eab,eb->ea", Ke, mvec.reshape(q.nelem, -1))
        assert np.abs(r).max() < 1e-11


def test_material_closed_form():
    # parity unpinned by the reference (external GMat); closed-form identities only
    rng = np.random.default_rng(4)
    e = rng.standard_normal((50, 3, 3))
    e = 0.5 * (e + np.swapaxes(e, 1, 2))
    K, G = 10.0, rng.uniform(0.5, 2, 50)
    s = oracle.elastic_stress(e, K, G)
    tr = np.trace(e, axis1=1, axis2=2)
    assert np.allclose(np.trace(s, axis1=1, axis2=2), 3 * K * tr)  # volumetric part
    dev_s = s - np.trace(s, axis1=1, axis2=2)[:, None, None] / 3 * np.eye(3)
    dev_e = e - tr[:, None, None] / 3 * np.eye(3)
    assert np.allclose(dev_s, 2 * G[:, None, None] * dev_e)
    C = oracle.elastic_tangent(np.full(50, K), G)
    assert np.allclose(np.einsum("pijkl,plk->pij", C, e), s)


# ---- lumped mass known answer (tests/basic/MatrixDiagonal.cpp:12-53) ---------------------------

def test_lumped_mass_golden(fixtures):
    for et in (QUAD4, QUAD4_PLANAR):
        c, conn = oracle.quad4_regular(2, 2)
        dofs = oracle.dofs(c.shape[0], 2)
        vec = oracle.Vector(conn, dofs)
        xi, w = oracle.quad_rule(et, nodal=True)
        q = oracle.Quadrature(et, vec.AsElement(c), xi, w)
        A = oracle.diag_assemble(conn, dofs, q.Int_N_scalar_NT_dV(np.ones((q.nelem, q.nip))))
        assert np.allclose(A, fixtures["MatrixDiagonal_lumped_mass_Regular_2_2"], atol=1e-12)


@pytest.mark.parametrize("et", [QUAD4, QUAD4_PLANAR, HEX8])
def test_lumped_mass_periodic_all_ones(et):
    # tests/basic/ElementQuad4.cpp:172-189, ElementHex8.cpp:157-174
    c, conn, vec, q = _quad(et, nodal=True)
    n = 3
    ties = (oracle.hex8_nodes_periodic(n, n, n) if et == HEX8 else oracle.quad4_nodes_periodic(n, n))
    dp = oracle.dofs_periodic(c.shape[0], oracle.NDIM[et], ties)
    A = oracle.diag_assemble(conn, dp, q.Int_N_scalar_NT_dV(np.ones((q.nelem, q.nip))))
    assert np.allclose(A, 1.0)


# ---- Vector (tests/basic/Vector.cpp:11-167) ----------------------------------------------------

def test_vector_periodic_golden(fixtures):
    g = fixtures["Vector_periodic_Regular_2_2"]
    c, conn = oracle.quad4_regular(2, 2)
    dp = oracle.dofs_periodic(c.shape[0], 2, oracle.quad4_nodes_periodic(2, 2))
    vec = oracle.Vector(conn, dp)
    v = np.array(g["asDofs - nodevec"]["input"])
    for V in (vec.AsDofs(v), vec.AsDofs(vec.AsElement(vec.AsDofs(v)))):
        assert V.size == 8
        for k, (n, i) in g["asDofs - nodevec"]["expected"]:
            assert V[k] == v[n, i]
    f = np.array(g["asDofs - assembleDofs"]["input"])
    for F in (vec.AssembleDofs(f), vec.AssembleDofs(vec.AsElement(f))):
        assert F.size == 8
        for k, val in g["asDofs - assembleDofs"]["expected"]:
            assert abs(F[k] - val) < 1e-12


def test_vector_brute_force_periodic():
    c, conn = oracle.quad4_regular(4, 3)
    dp = oracle.dofs_periodic(c.shape[0], 2, oracle.quad4_nodes_periodic(4, 3))
    vec = oracle.Vector(conn, dp)
    rng = np.random.default_rng(0)
    ev = rng.standard_normal((conn.shape[0], 4, 2))
    ref = np.zeros(vec.ndof)
    for e in range(conn.shape[0]):
        for m in range(4):
            for i in range(2):
                ref[dp[conn[e, m], i]] += ev[e, m, i]
    assert np.array_equal(vec.AssembleDofs(ev), ref)
    assert np.array_equal(vec.AssembleNode(ev), ref[dp])  # Q3: tied nodes receive the DOF sum
    dv = rng.standard_normal(vec.ndof)
    assert np.array_equal(vec.AsElement(dv), dv[dp[conn]])
    # AsDofs(elemvec) overwrite: last (e, m, i) in loop order wins
    last = np.zeros(vec.ndof)
    for e in range(conn.shape[0]):
        for m in range(4):
            for i in range(2):
                last[dp[conn[e, m], i]] = ev[e, m, i]
    assert np.array_equal(vec.AsDofs(ev), last)


# ---- sparse Matrix (tests/basic/Matrix.cpp:10-94) -----------------------------------------------

def test_matrix_assemble_dense_brute_force_and_roundtrip():
    c, conn = oracle.quad4_regular(3, 2)
    dofs = oracle.dofs(c.shape[0], 2)
    rng = np.random.default_rng(9)
    a = rng.random((conn.shape[0], 8, 8))
    a = 0.5 * (a + np.swapaxes(a, 1, 2)) + 8 * np.eye(8)
    rowptr, col, val = oracle.matrix_assemble(conn, dofs, a)
    ndof = dofs.max() + 1
    dense = np.zeros((ndof, ndof))
    for e in range(conn.shape[0]):
        d = dofs[conn[e]].ravel()
        dense[np.ix_(d, d)] += a[e]
    mine = np.zeros_like(dense)
    for r in range(ndof):
        mine[r, col[rowptr[r]:rowptr[r + 1]]] = val[rowptr[r]:rowptr[r + 1]]
    assert np.allclose(mine, dense, rtol=1e-14)
    x = rng.random(ndof)
    assert np.allclose(oracle.csr_dot(rowptr, col, val, x), dense @ x, rtol=1e-13)
    # Dot -> Solve round trip (Matrix.cpp:10-36) via the plain CG
    b = oracle.csr_dot(rowptr, col, val, x)
    xs, it, res = oracle.pcg_csr(rowptr, col, val, np.diag(dense).copy(), b, tol=1e-13)
    assert res < 1e-12 and np.allclose(xs, x, rtol=1e-9)


# ---- Tyings + K' (TyingsPeriodic.h, MatrixPartitionedTyings.h:655-674) ------------------------

def periodic_problem(nx, ny, gamma=1e-3, seed=1):
    """Restatement of docs/examples/statics/periodic_elastoplastic.py:12-47 (elastic)."""
    coor, conn = oracle.quad4_regular(nx, ny)
    dofs = oracle.dofs(coor.shape[0], 2)
    co, do, cdofs, cnodes = oracle.tyings_control(coor, dofs)
    iip = np.concatenate((cdofs.ravel(), do[0]))
    ty = oracle.tyings_periodic(co, do, cdofs, oracle.quad4_nodes_periodic(nx, ny), iip)
    return coor, conn, co, ty


def test_tyings_structure_and_closed_form():
    nx, ny = 5, 4
    coor, conn, co, ty = periodic_problem(nx, ny)
    nties = oracle.quad4_nodes_periodic(nx, ny).shape[0]
    assert ty["nnd"] == 2 * nties and ty["nnp"] == 6
    assert ty["nni"] == 2 * (nx * ny + 2)
    assert ty["nnu"] + ty["nnp"] == ty["nni"]
    # Reorder is a permutation of all DOFs
    assert np.array_equal(np.sort(ty["dofs"].ravel()), np.arange(ty["nni"] + ty["nnd"]))
    # u_d = u_i + (x_d - x_i) . ctrl : apply Cdi to an affine field u = F x
    F = np.array([[0.01, 0.03], [-0.02, 0.005]])
    u = np.zeros((co.shape[0], 2))
    u[: coor.shape[0]] = coor @ F.T
    u[coor.shape[0]:] = F  # control node j holds row j of F (ctrl(j,k) = F_jk)
    ui = np.zeros(ty["nni"])
    ud_true = np.zeros(ty["nnd"])
    for n in range(co.shape[0]):
        for i in range(2):
            d = ty["dofs"][n, i]
            if d < ty["nni"]:
                ui[d] = u[n, i]
            else:
                ud_true[d - ty["nni"]] = u[n, i]
    assert np.allclose(ty["Cdi"] @ ui, ud_true, atol=1e-14)
    assert np.allclose(ty["Cdi"], np.hstack((ty["Cdu"], ty["Cdp"])))


def test_tyings_structure_and_closed_form_hex8():
    """The 3-D restatement: sizes, a permutation of all DOFs, and u_d = u_i + (x_d - x_i) . ctrl
    for an affine field (TyingsPeriodic.h:244-256)."""
    nx, ny, nz = 4, 3, 5
    coor, conn = oracle.hex8_regular(nx, ny, nz)
    dofs = oracle.dofs(coor.shape[0], 3)
    co, do, cdofs, cnodes = oracle.tyings_control(coor, dofs)
    iip = np.concatenate((cdofs.ravel(), do[0]))
    ties = oracle.hex8_nodes_periodic(nx, ny, nz)
    ty = oracle.tyings_periodic(co, do, cdofs, ties, iip)
    assert len(ties) == (nx + 1) * (ny + 1) * (nz + 1) - nx * ny * nz
    assert ty["nnd"] == 3 * len(ties) and ty["nnp"] == 12
    assert ty["nni"] == 3 * (nx * ny * nz + 3)
    assert np.array_equal(np.sort(ty["dofs"].ravel()), np.arange(ty["nni"] + ty["nnd"]))
    F = np.array([[0.01, 0.03, -0.007], [-0.02, 0.005, 0.011], [0.004, -0.013, 0.021]])
    u = np.zeros((co.shape[0], 3))
    u[: coor.shape[0]] = coor @ F.T
    u[coor.shape[0]:] = F
    d = ty["dofs"]
    indep = d < ty["nni"]
    ui = np.zeros(ty["nni"])
    ud = np.zeros(ty["nnd"])
    ui[d[indep]] = u[indep]
    ud[d[~indep] - ty["nni"]] = u[~indep]
    assert np.abs(ty["Cdi"] @ ui - ud).max() < 1e-14
    assert np.allclose(ty["Cdi"], np.hstack((ty["Cdu"], ty["Cdp"])))


def test_Kprime_cg_vs_scipy_direct():
    scipy = pytest.importorskip("scipy.sparse.linalg")
    nx, ny = 6, 5
    coor, conn, co, ty = periodic_problem(nx, ny)
    dofs = ty["dofs"]
    vec = oracle.Vector(conn, dofs[: coor.shape[0]])
    q = oracle.Quadrature(QUAD4_PLANAR, vec.AsElement(coor))
    rng = np.random.default_rng(3)
    G = np.exp(0.4 * rng.standard_normal(q.nelem))[:, None] * np.ones((1, q.nip))
    C = oracle.elastic_tangent(np.full_like(G, 10.0), G)
    Ke = q.Int_gradN_dot_tensor4_dot_gradNT_dV(C)
    # dense A over all DOFs in the [u, p, d] numbering, then the reference's condensation
    ndof = ty["nni"] + ty["nnd"]
    A = np.zeros((ndof, ndof))
    for e in range(q.nelem):
        d = dofs[conn[e]].ravel()
        A[np.ix_(d, d)] += Ke[e]
    nnu, nni = ty["nnu"], ty["nni"]
    u, p, d = slice(0, nnu), slice(nnu, nni), slice(nni, ndof)
    Cdu, Cdp = ty["Cdu"], ty["Cdp"]
    Kp = A[u, u] + A[u, d] @ Cdu + Cdu.T @ A[d, u] + Cdu.T @ A[d, d] @ Cdu
    Kup = A[u, p] + A[u, d] @ Cdp + Cdu.T @ A[d, p] + Cdu.T @ A[d, d] @ Cdp
    xp = np.zeros(ty["nnp"])
    xp[ty["control"][0, 1] - nnu] = 1e-3  # simple shear increment, ctrl(0, 1)
    b = -Kup @ xp
    xd = np.linalg.solve(Kp, b)
    import scipy.sparse as sp
    Ks = sp.csr_matrix(Kp)
    xs = scipy.spsolve(Ks.tocsc(), b)
    assert np.allclose(xd, xs, rtol=1e-10, atol=1e-16)
    x, it, res = oracle.pcg_csr(Ks.indptr.astype(np.int64), Ks.indices.astype(np.int64), Ks.data,
                               np.diag(Kp).copy(), b, tol=1e-12)
    assert res < 1e-12 and np.allclose(x, xs, rtol=1e-8, atol=1e-14)
    # K' equals assembly with dependent DOFs merged into their independent DOFs (controls fixed)
    merged = np.where(dofs >= nni, -1, dofs)
    ties = oracle.quad4_nodes_periodic(nx, ny)
    for i_, d_ in ties:
        merged[d_] = dofs[i_]
    Km = np.zeros((nni, nni))
    for e in range(q.nelem):
        dd = merged[conn[e]].ravel()
        Km[np.ix_(dd, dd)] += Ke[e]
    assert np.allclose(Km[u, u], Kp, atol=1e-12)


# ---- velocity-Verlet (docs/examples/dynamics/Elastic-VelocityVerlet/main.cpp:110-157) ----------

def test_vv_step_matches_python_restatement_and_conserves_energy():
    nx = ny = 6
    c, conn = oracle.quad4_regular(nx, ny)
    dofs = oracle.dofs(c.shape[0], 2)
    vec = oracle.Vector(conn, dofs)
    q = oracle.Quadrature(QUAD4_PLANAR, vec.AsElement(c))
    xi, w = oracle.quad_rule(QUAD4_PLANAR, nodal=True)
    qn = oracle.Quadrature(QUAD4_PLANAR, vec.AsElement(c), xi, w)
    M = oracle.diag_assemble(conn, dofs, qn.Int_N_scalar_NT_dV(np.ones((q.nelem, 4))))
    inv = oracle.diag_inverse(M)
    Ke = np.full(q.nelem, 10.0)
    Ge = np.exp(0.3 * oracle.randn(2, q.nelem))
    u = oracle.randn(1, dofs.size, 0.0, 1e-3)
    v = np.zeros_like(u)
    a = np.zeros_like(u)
    u0, dt = u.copy(), 0.01
    # python restatement of one step
    Gq = np.repeat(Ge[:, None], 4, 1)
    un = u0 + dt * v + 0.5 * dt * dt * a
    eps = q.SymGradN_vector(vec.AsElement(un))
    fint = vec.AssembleDofs(q.Int_gradN_dot_tensor2_dV(oracle.elastic_stress(eps, 10.0, Gq)))
    an = inv * (0.0 - fint)
    vn = v + 0.5 * dt * (a + an)
    f = oracle.vv_steps(1, QUAD4_PLANAR, conn, dofs, q.dNx, q.dV(), Ke, Ge, inv, dt, u, v, a)
    assert np.array_equal(u, un) and np.array_equal(f, fint)
    assert np.array_equal(a, an) and np.array_equal(v, vn)

    # energy: E_kin + E_pot drifts O(dt^2) over many steps
    def energy(u, v):
        eps = q.SymGradN_vector(vec.AsElement(u))
        sig = oracle.elastic_stress(eps, 10.0, Gq)
        return 0.5 * np.sum(M * v * v) + 0.5 * np.sum(np.einsum("eqij,eqij->eq", eps, sig) * q.dV())

    E0 = energy(u, v)
    oracle.vv_steps(200, QUAD4_PLANAR, conn, dofs, q.dNx, q.dV(), Ke, Ge, inv, dt, u, v, a)
    assert abs(energy(u, v) - E0) / E0 < 1e-3


def test_randn_moments_and_determinism():
    z = oracle.randn(7, 200001)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01
    assert np.array_equal(z[:1000], oracle.randn(7, 1000))


def test_vv_steps_parallel_bit_identical():
    """The all-cores CPU baseline (OpenMP element chunks + ordered per-DOF gather) repeats the serial
    restatement bit for bit, on a renumbered mesh as well."""
    nx, ny = 37, 21
    coor, conn = oracle.quad4_regular(nx, ny)
    rng = np.random.default_rng(4)
    conn = conn[rng.permutation(conn.shape[0])]
    dofs = oracle.dofs(coor.shape[0], 2)[rng.permutation(coor.shape[0])]
    q = oracle.Quadrature(QUAD4_PLANAR, oracle.Vector(conn, dofs).AsElement(coor))
    E = conn.shape[0]
    K = np.full(E, 10.0)
    G = np.exp(0.3 * rng.standard_normal(E))
    inv = 1.0 / (1.0 + rng.random(2 * coor.shape[0]))
    u0 = 1e-3 * rng.standard_normal(inv.size)
    out = []
    for fn in (oracle.vv_steps, oracle.vv_steps_parallel):
        u, v, a = u0.copy(), np.zeros_like(u0), np.zeros_like(u0)
        r = fn(5, QUAD4_PLANAR, conn, dofs, q.dNx, q.dV(), K, G, inv, 0.01, u, v, a)
        out.append((u, v, a, r[0] if isinstance(r, tuple) else r))
    for x, y in zip(*out):
        assert np.array_equal(x, y)
