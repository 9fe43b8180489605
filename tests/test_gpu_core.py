"""GPU parity of the HIP path against the FP64 oracle (bit-exact for FP64 and integer work,
stated tolerances for FP32 kernels).  Runs on an MI355X; every call goes through the C ABI."""

import numpy as np
import pytest

import oracle
from oracle import HEX8, QUAD4, QUAD4_PLANAR

pytestmark = pytest.mark.gpu

# FP32 element kernels vs FP64: norm-wise relative error bound (stated tolerance)
TOL_F32 = 1e-6


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


def relerr(a, b):
    return np.linalg.norm(np.asarray(a, dtype=np.float64) - b) / max(np.linalg.norm(b), 1e-300)


def mesh(et, n, distort=0.0, seed=3):
    if et == HEX8:
        c, conn = oracle.hex8_regular(n, n, n)
    else:
        c, conn = oracle.quad4_regular(n + 1, n)
    if distort:
        c = c + distort * np.random.default_rng(seed).uniform(-1, 1, c.shape)
    return c, conn


QUAD_CLS = {QUAD4: "Quadrature", QUAD4_PLANAR: "QuadraturePlanar", HEX8: "Quadrature"}


def gquad(gf, et, x, nodal=False):
    ns = gf.Element.Hex8 if et == HEX8 else gf.Element.Quad4
    cls = getattr(ns, QUAD_CLS[et])
    if nodal:
        return cls(x, ns.Nodal.xi(), ns.Nodal.w())
    return cls(x)


@pytest.mark.parametrize("et", [QUAD4, QUAD4_PLANAR, HEX8])
@pytest.mark.parametrize("distort", [0.0, 0.2])
def test_quadrature_fp64_bit_exact(gf, et, distort):
    c, conn = mesh(et, 5, distort)
    nd = c.shape[1]
    vec = oracle.Vector(conn, oracle.dofs(c.shape[0], nd))
    x = vec.AsElement(c)
    q = oracle.Quadrature(et, x)
    gq = gquad(gf, et, x)
    assert np.array_equal(host(gq.GradN), q.GradN())
    assert np.array_equal(host(gq.dV), q.dV())
    rng = np.random.default_rng(1)
    ue = rng.standard_normal((q.nelem, q.nne, nd))
    for name in ("GradN_vector", "GradN_vector_T", "SymGradN_vector", "InterpQuad_vector"):
        assert np.array_equal(host(getattr(gq, name)(ue)), getattr(q, name)(ue)), name
    td = q.tdim
    sig = rng.standard_normal((q.nelem, q.nip, td, td))
    assert np.array_equal(host(gq.Int_gradN_dot_tensor2_dV(sig)), q.Int_gradN_dot_tensor2_dV(sig))
    qv = rng.standard_normal((q.nelem, q.nip, nd))
    assert np.array_equal(host(gq.Int_N_vector_dV(qv)), q.Int_N_vector_dV(qv))
    rho = rng.uniform(0.5, 2.0, (q.nelem, q.nip))
    assert np.array_equal(host(gq.Int_N_scalar_NT_dV(rho)), q.Int_N_scalar_NT_dV(rho))
    C = rng.standard_normal((q.nelem, q.nip) + (td,) * 4)
    assert np.array_equal(host(gq.Int_gradN_dot_tensor4_dot_gradNT_dV(C)), q.Int_gradN_dot_tensor4_dot_gradNT_dV(C))


@pytest.mark.parametrize("et", [QUAD4, QUAD4_PLANAR])
@pytest.mark.parametrize("nodal", [False, True])
def test_quad4_stiffness_fp64_blocks(gf, et, nodal):
    # the Quad4 FP64 K_e kernel (one thread per element row, 64 elements per workgroup, rows
    # streamed through LDS) over many full workgroups and a partial one, bit-exact with the oracle;
    # an unaligned output takes the generic kernel and gives the same bits
    import torch

    c, conn = mesh(et, 45, 0.2, seed=8)  # 46 x 45 = 2070 elements: 32 full workgroups + 22
    x = oracle.Vector(conn, oracle.dofs(c.shape[0], 2)).AsElement(c)
    q = oracle.Quadrature(et, x, *oracle.quad_rule(et, nodal))
    gq = gquad(gf, et, x, nodal)
    rng = np.random.default_rng(12)
    C = rng.standard_normal((q.nelem, q.nip) + (q.tdim,) * 4)
    Ke = q.Int_gradN_dot_tensor4_dot_gradNT_dV(C)
    assert np.array_equal(host(gq.Int_gradN_dot_tensor4_dot_gradNT_dV(C)), Ke)
    buf = torch.zeros(Ke.size + 1, dtype=torch.float64, device="cuda")
    shifted = buf[1:].view(Ke.shape)
    assert shifted.data_ptr() % 16 != 0
    gq.Int_gradN_dot_tensor4_dot_gradNT_dV(torch.from_numpy(C).cuda(), out=shifted)
    assert np.array_equal(host(shifted), Ke)


@pytest.mark.parametrize("rule", ["centroid", "gauss3"])
def test_hex8_gradients_custom_rules(gf, rule):
    """Hex8 gradients of rules with 1 point (the stored-geometry kernel) and 27 points (geometry
    recomputed from the coordinates) bit-exact with the oracle, FP64; FP32 equal to FP32 arithmetic on
    the FP64 dN/dx rounded once."""
    import torch

    c, conn = mesh(HEX8, 4, 0.2)
    vec = oracle.Vector(conn, oracle.dofs(c.shape[0], 3))
    x = vec.AsElement(c)
    if rule == "centroid":
        xi, w = np.zeros((1, 3)), np.array([8.0])
    else:
        a = np.sqrt(3.0 / 5.0)
        p1, w1 = np.array([-a, 0.0, a]), np.array([5.0, 8.0, 5.0]) / 9.0
        xi = np.array([[p, q, r] for p in p1 for q in p1 for r in p1])
        w = np.array([wp * wq * wr for wp in w1 for wq in w1 for wr in w1])
    q = oracle.Quadrature(HEX8, x, xi, w)
    gq = gf.Element.Hex8.Quadrature(x, xi, w)
    ue = np.random.default_rng(8).standard_normal((q.nelem, 8, 3))
    for name in ("GradN_vector", "GradN_vector_T", "SymGradN_vector"):
        assert np.array_equal(host(getattr(gq, name)(ue)), getattr(q, name)(ue)), name
    u32 = ue.astype(np.float32)
    g32 = q.GradN().astype(np.float32)
    got = host(gq.GradN_vector(torch.from_numpy(u32).cuda()))
    want = np.zeros((q.nelem, q.nip, 3, 3), dtype=np.float32)
    for m in range(8):  # the kernel's accumulation order, in FP32
        want += g32[:, :, m, :, None] * u32[:, None, m, None, :]
    assert np.array_equal(got, want)


@pytest.mark.parametrize("et", [QUAD4_PLANAR, HEX8])
def test_quadrature_fp32_tolerance(gf, et):
    import torch

    c, conn = mesh(et, 6, 0.1)
    nd = c.shape[1]
    vec = oracle.Vector(conn, oracle.dofs(c.shape[0], nd))
    q = oracle.Quadrature(et, vec.AsElement(c))
    gq = gquad(gf, et, vec.AsElement(c))
    rng = np.random.default_rng(2)
    ue = rng.standard_normal((q.nelem, q.nne, nd))
    ue32 = torch.from_numpy(ue.astype(np.float32)).cuda()
    eps32 = gq.SymGradN_vector(ue32)
    assert eps32.dtype == torch.float32
    assert relerr(host(eps32), q.SymGradN_vector(ue)) < TOL_F32
    sig = q.SymGradN_vector(ue) * 3.0
    f32 = gq.Int_gradN_dot_tensor2_dV(torch.from_numpy(sig.astype(np.float32)).cuda())
    assert relerr(host(f32), q.Int_gradN_dot_tensor2_dV(sig)) < TOL_F32


def test_material_bit_exact(gf):
    rng = np.random.default_rng(4)
    for td in (2, 3):
        e = rng.standard_normal((40, 4, td, td))
        K = np.full((40, 4), 10.0)
        G = rng.uniform(0.5, 1.5, (40, 4))
        assert np.array_equal(host(gf.elastic_stress(e, gf.to_device(K), gf.to_device(G))),
                              oracle.elastic_stress(e, K, G))
    K = np.full((7, 8), 10.0)
    G = rng.uniform(0.5, 1.5, (7, 8))
    assert np.array_equal(host(gf.elastic_tangent(K, G)), oracle.elastic_tangent(K, G))


@pytest.mark.parametrize("td", [2, 3])
def test_elastic_tangent_streams(gf, td):
    """The tangent kernel's vector path (full 256-point workgroups, 16-byte stores) and its tail,
    FP64 bit-exact with the oracle and FP32 equal to the oracle's FP64 rounded once."""
    import torch

    rng = np.random.default_rng(7)
    K = rng.uniform(5.0, 15.0, (301, 8))  # 2408 points: 9 full workgroups and a tail of 104
    G = rng.uniform(0.5, 1.5, (301, 8))
    ref = oracle.elastic_tangent(K, G, td)
    Kd, Gd = gf.to_device(K), gf.to_device(G)
    assert np.array_equal(host(gf.elastic_tangent(Kd, Gd, td=td)), ref)
    c32 = gf.elastic_tangent(Kd, Gd, td=td, dtype=torch.float32)
    assert c32.dtype == torch.float32
    assert np.array_equal(host(c32), ref.astype(np.float32))


@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("et", [QUAD4, HEX8])
def test_vector_bit_exact(gf, et, periodic):
    c, conn = mesh(et, 4)
    nd = c.shape[1]
    if periodic:
        ties = oracle.hex8_nodes_periodic(4, 4, 4) if et == HEX8 else oracle.quad4_nodes_periodic(5, 4)
        dofs = oracle.dofs_periodic(c.shape[0], nd, ties)
    else:
        dofs = oracle.dofs(c.shape[0], nd)
    ov = oracle.Vector(conn, dofs)
    gv = gf.Vector(conn, dofs)
    assert gv.ndof == ov.ndof
    rng = np.random.default_rng(7)
    ev = rng.standard_normal(ov.conn.shape + (nd,))
    nv = rng.standard_normal((c.shape[0], nd))
    dv = rng.standard_normal(ov.ndof)
    assert np.array_equal(host(gv.AsElement(nv)), ov.AsElement(nv))
    assert np.array_equal(host(gv.AsElement(dv)), ov.AsElement(dv))
    assert np.array_equal(host(gv.AsDofs(nv)), ov.AsDofs(nv))
    assert np.array_equal(host(gv.AsDofs(ev)), ov.AsDofs(ev))
    assert np.array_equal(host(gv.AsNode(dv)), ov.AsNode(dv))
    assert np.array_equal(host(gv.AsNode(ev)), ov.AsNode(ev))
    assert np.array_equal(host(gv.AssembleDofs(nv)), ov.AssembleDofs(nv))
    assert np.array_equal(host(gv.AssembleDofs(ev)), ov.AssembleDofs(ev))
    assert np.array_equal(host(gv.AssembleNode(ev)), ov.AssembleNode(ev))


def test_vector_reference_golden(gf, fixtures):
    # tests/basic/Vector.cpp:11-167 on the GPU path
    g = fixtures["Vector_periodic_Regular_2_2"]
    m = gf.Mesh.Quad4.Regular(2, 2)
    vec = gf.Vector(m.conn, m.dofsPeriodic)
    v = np.array(g["asDofs - nodevec"]["input"])
    for V in (host(vec.AsDofs(v)), host(vec.AsDofs(vec.AsElement(vec.AsDofs(v))))):
        for k, (n, i) in g["asDofs - nodevec"]["expected"]:
            assert V[k] == v[n, i]
    f = np.array(g["asDofs - assembleDofs"]["input"])
    for F in (host(vec.AssembleDofs(f)), host(vec.AssembleDofs(vec.AsElement(f)))):
        assert np.allclose(F, 0.0, atol=1e-12)


def test_lumped_mass_golden_and_diag(gf, fixtures):
    m = gf.Mesh.Quad4.Regular(2, 2)
    vec = gf.Vector(m.conn, m.dofs)
    x = vec.AsElement(m.coor)
    q = gquad(gf, QUAD4, x, nodal=True)
    M = gf.MatrixDiagonal(vec, None)
    M.assemble(q.Int_N_scalar_NT_dV(np.ones((m.nelem, 4))))
    assert np.allclose(host(M.data), fixtures["MatrixDiagonal_lumped_mass_Regular_2_2"], atol=1e-12)
    # bit-exact vs oracle on a larger periodic Hex8 mesh; Dot / Solve round trip
    c, conn = oracle.hex8_regular(3, 3, 3)
    dp = oracle.dofs_periodic(c.shape[0], 3, oracle.hex8_nodes_periodic(3, 3, 3))
    ov = oracle.Vector(conn, oracle.dofs(c.shape[0], 3))
    xi, w = oracle.quad_rule(HEX8, True)
    oq = oracle.Quadrature(HEX8, ov.AsElement(c), xi, w)
    mm = oq.Int_N_scalar_NT_dV(np.ones((oq.nelem, 8)))
    A = oracle.diag_assemble(conn, dp, mm)
    gM = gf.MatrixDiagonal(conn, dp)
    gM.assemble(mm)
    assert np.array_equal(host(gM.data), A)
    assert np.allclose(A, 1.0)
    rng = np.random.default_rng(0)
    xv = rng.random(A.size)
    assert np.array_equal(host(gM.Dot(xv)), A * xv)
    assert np.array_equal(host(gM.Solve(A * xv)), oracle.diag_inverse(A) * (A * xv))


@pytest.mark.parametrize("et,periodic", [(QUAD4, False), (QUAD4, True), (HEX8, False), (HEX8, True)])
def test_matrix_csr_bit_exact(gf, et, periodic):
    c, conn = mesh(et, 3)
    nd = c.shape[1]
    if periodic:
        ties = oracle.hex8_nodes_periodic(3, 3, 3) if et == HEX8 else oracle.quad4_nodes_periodic(4, 3)
        dofs = oracle.dofs_periodic(c.shape[0], nd, ties)
    else:
        dofs = oracle.dofs(c.shape[0], nd)
    rng = np.random.default_rng(11)
    K = conn.shape[1] * nd
    a = rng.standard_normal((conn.shape[0], K, K))
    rowptr, col, val = oracle.matrix_assemble(conn, dofs, a)
    A = gf.Matrix(conn, dofs)
    A.assemble(a)
    grp, gcol, gval = (host(t) for t in A.csr())
    assert np.array_equal(grp, rowptr) and np.array_equal(gcol, col)
    assert np.array_equal(gval, val)
    x = rng.standard_normal(rowptr.size - 1)
    assert np.array_equal(host(A.Dot(x)), oracle.csr_dot(rowptr, col, val, x))
    dense = np.zeros((rowptr.size - 1,) * 2)
    for r in range(rowptr.size - 1):
        dense[r, col[rowptr[r]:rowptr[r + 1]]] = val[rowptr[r]:rowptr[r + 1]]
    assert np.array_equal(host(A.Todense()), dense)


@pytest.mark.parametrize("nfan", [20, 40, 70])
def test_matrix_csr_long_rows(gf, nfan):
    # a fan of Quad4 elements around hub node 0: the hub's rows hold 2 (3 nfan + 1) columns, which
    # exercises the 128- and 256-entry row capacities and (nfan 70: 422 columns) the row-serial path
    conn = np.array([[0, 1 + 3 * i, 2 + 3 * i, 3 + 3 * i] for i in range(nfan)], dtype=np.int64)
    dofs = oracle.dofs(3 * nfan + 1, 2)
    rng = np.random.default_rng(nfan)
    a = rng.standard_normal((nfan, 8, 8))
    rowptr, col, val = oracle.matrix_assemble(conn, dofs, a)
    assert rowptr[1] - rowptr[0] == 2 * (3 * nfan + 1)
    A = gf.Matrix(conn, dofs)
    A.assemble(a)
    grp, gcol, gval = (host(t) for t in A.csr())
    assert np.array_equal(grp, rowptr) and np.array_equal(gcol, col) and np.array_equal(gval, val)
    x = rng.standard_normal(rowptr.size - 1)
    assert np.array_equal(host(A.Dot(x)), oracle.csr_dot(rowptr, col, val, x))


@pytest.mark.parametrize("et", [QUAD4, HEX8])
@pytest.mark.parametrize("periodic", [False, True])
def test_single_element_mesh(gf, et, periodic):
    # the smallest mesh the generators make; periodic ties every node to node 0, so all element
    # contributions collapse onto ndim DOFs (the longest ordered accumulation per entry)
    if et == HEX8:
        c, conn = oracle.hex8_regular(1, 1, 1)
        ties = oracle.hex8_nodes_periodic(1, 1, 1)
    else:
        c, conn = oracle.quad4_regular(1, 1)
        ties = oracle.quad4_nodes_periodic(1, 1)
    nd = c.shape[1]
    dofs = oracle.dofs_periodic(c.shape[0], nd, ties) if periodic else oracle.dofs(c.shape[0], nd)
    ov = oracle.Vector(conn, dofs)
    gv = gf.Vector(conn, dofs)
    assert gv.ndof == ov.ndof == (nd if periodic else c.size)
    rng = np.random.default_rng(21)
    ev = rng.standard_normal(ov.conn.shape + (nd,))
    dv = rng.standard_normal(ov.ndof)
    assert np.array_equal(host(gv.AssembleDofs(ev)), ov.AssembleDofs(ev))
    assert np.array_equal(host(gv.AsElement(dv)), ov.AsElement(dv))
    x = oracle.Vector(conn, oracle.dofs(c.shape[0], nd)).AsElement(c)
    q = oracle.Quadrature(et, x)
    gq = gquad(gf, et, x)
    assert np.array_equal(host(gq.dV), q.dV())
    C = rng.standard_normal((1, q.nip) + (q.tdim,) * 4)
    Ke = q.Int_gradN_dot_tensor4_dot_gradNT_dV(C)
    assert np.array_equal(host(gq.Int_gradN_dot_tensor4_dot_gradNT_dV(C)), Ke)
    rowptr, col, val = oracle.matrix_assemble(conn, dofs, Ke)
    A = gf.Matrix(conn, dofs)
    A.assemble(Ke)
    grp, gcol, gval = (host(t) for t in A.csr())
    assert np.array_equal(grp, rowptr) and np.array_equal(gcol, col) and np.array_equal(gval, val)
    Kd = Ke * np.eye(Ke.shape[1])  # MatrixDiagonal takes diagonal element matrices
    diag = oracle.diag_assemble(conn, dofs, Kd)
    gM = gf.MatrixDiagonal(conn, dofs)
    gM.assemble(Kd)
    assert np.array_equal(host(gM.data), diag)


def _explicit_setup(nx, ny, seed=1, shuffle=False, fixed=True, kvar=False, distort=False, tied=False, permdofs=False):
    coor, conn = oracle.quad4_regular(nx, ny)
    if distort:  # structured connectivity, perturbed coordinates: the kernel must read them
        coor = coor + 0.15 * oracle.randn(seed + 300, coor.size).reshape(coor.shape)
    if shuffle:  # same mesh, elements renumbered: defeats the structured-grid detection
        perm = np.random.default_rng(5).permutation(conn.shape[0])
        conn = conn[perm]
    dofs = oracle.dofs(coor.shape[0], 2)
    if tied:  # periodic DOFs: a DOF gathers the contributions of several nodes
        dofs = oracle.dofs_periodic(coor.shape[0], 2, oracle.quad4_nodes_periodic(nx, ny))
    if permdofs:  # every DOF still belongs to one node component, numbered in another order
        dofs = np.random.default_rng(7).permutation(dofs.size)[dofs]
    vec = oracle.Vector(conn, dofs)
    x = vec.AsElement(coor)
    q = oracle.Quadrature(QUAD4_PLANAR, x)
    xi, w = oracle.quad_rule(QUAD4_PLANAR, True)
    qn = oracle.Quadrature(QUAD4_PLANAR, x, xi, w)
    M = oracle.diag_assemble(conn, dofs, qn.Int_N_scalar_NT_dV(np.ones((q.nelem, 4))))
    inv = oracle.diag_inverse(M)
    if fixed:
        bot = oracle.quad4_nodeset(nx, ny, "nodesBottomEdge")
        top = oracle.quad4_nodeset(nx, ny, "nodesTopEdge")
        inv[dofs[bot].ravel()] = 0.0
        inv[dofs[top].ravel()] = 0.0
    Ke = np.full(q.nelem, 10.0)
    if kvar:
        Ke = (10.0 + oracle.randn(seed + 200, q.nelem)).astype(np.float32).astype(np.float64)
    Ge = np.exp(0.3 * oracle.randn(seed + 100, q.nelem)).astype(np.float32).astype(np.float64)
    u = oracle.randn(seed, int(dofs.max()) + 1, 0.0, 1e-3)
    v = np.zeros_like(u)
    if fixed:
        v[dofs[top, 0]] = 1e-4
        u[dofs[bot].ravel()] = 0.0
        u[dofs[top].ravel()] = 0.0
    a = np.zeros_like(u)
    return coor, conn, dofs, q, inv, Ke, Ge, u, v, a


@pytest.mark.parametrize("nx,ny,shuffle,kvar,distort,tied", [(64, 64, False, False, False, False),
                                                            (130, 17, False, True, False, False),
                                                            (3, 5, False, False, False, False),
                                                            (24, 20, True, True, False, False),
                                                            (40, 24, False, False, True, False),
                                                            (70, 45, True, False, True, False),
                                                            (24, 20, True, True, False, True),
                                                            (33, 29, True, True, True, "perm"),
                                                            (128, 20, False, True, False, False),
                                                            (40, 32, False, False, True, False)])
def test_explicit_fp64_bit_exact(gf, nx, ny, shuffle, kvar, distort, tied):
    # tied: False, True (periodic DOFs) or "perm" (a permuted numbering, each DOF still on one node)
    coor, conn, dofs, q, inv, Ke, Ge, u, v, a = _explicit_setup(nx, ny, shuffle=shuffle, kvar=kvar, distort=distort,
                                                                tied=tied is True, permdofs=tied == "perm")
    vec = gf.Vector(conn, dofs)
    X = gf.ExplicitQuad4(vec, coor, Ke, Ge, inv, 0.01)
    assert X.structured == (not shuffle)
    assert X.path == ("structured" if not shuffle else "gather" if tied is True else "blocks")
    X.set_state(u, v, a)
    for nsteps in (1, 9):
        fint = oracle.vv_steps(nsteps, QUAD4_PLANAR, conn, dofs, q.dNx, q.dV(), Ke, Ge, inv, 0.01, u, v, a)
        X.step(nsteps, gf.F64, store_fint=True)
        gu, gv_, ga = (host(t) for t in X.state())
        assert np.array_equal(gu, u) and np.array_equal(gv_, v) and np.array_equal(ga, a)
        assert np.array_equal(host(X.fint()), fint)


@pytest.mark.parametrize("shuffle", [False, True])
def test_explicit_fp64_bit_exact_any_mass(gf, shuffle):
    # M^-1 with thousands of distinct values (a random density): the step reads the full array
    # instead of its palette of at most 256 values; both bit-exact with the oracle
    coor, conn, dofs, q, inv, Ke, Ge, u, v, a = _explicit_setup(40, 32, shuffle=shuffle)
    inv = inv * (1.0 + 0.2 * oracle.randn(77, inv.size))
    assert np.unique(inv).size > 256
    X = gf.ExplicitQuad4(gf.Vector(conn, dofs), coor, Ke, Ge, inv, 0.01)
    X.set_state(u, v, a)
    fint = oracle.vv_steps(7, QUAD4_PLANAR, conn, dofs, q.dNx, q.dV(), Ke, Ge, inv, 0.01, u, v, a)
    X.step(7, gf.F64, store_fint=True)
    gu, gv_, ga = (host(t) for t in X.state())
    assert np.array_equal(gu, u) and np.array_equal(gv_, v) and np.array_equal(ga, a)
    assert np.array_equal(host(X.fint()), fint)


@pytest.mark.parametrize("case", ["structured", "blocks", "blocks-permuted-dofs", "gather"])
def test_explicit_fp64_bit_exact_fext(gf, case):
    # an external force (MatrixDiagonalPartitioned::solve of f_ext - f_int, main.cpp:152-155) on each
    # of the step's kernels, bit-exact with the oracle
    coor, conn, dofs, q, inv, Ke, Ge, u, v, a = _explicit_setup(
        30, 26, shuffle=case != "structured", kvar=True, tied=case == "gather", permdofs=case == "blocks-permuted-dofs")
    fext = 1e-2 * oracle.randn(55, inv.size)
    X = gf.ExplicitQuad4(gf.Vector(conn, dofs), coor, Ke, Ge, inv, 0.01, fext=fext)
    assert X.path == case.split("-")[0]
    X.set_state(u, v, a)
    fint = oracle.vv_steps(5, QUAD4_PLANAR, conn, dofs, q.dNx, q.dV(), Ke, Ge, inv, 0.01, u, v, a, fext=fext)
    X.step(5, gf.F64, store_fint=True)
    gu, gv_, ga = (host(t) for t in X.state())
    assert np.array_equal(gu, u) and np.array_equal(gv_, v) and np.array_equal(ga, a)
    assert np.array_equal(host(X.fint()), fint)


@pytest.mark.parametrize("prec", ["F32", "F64", "F64_FAST"])
@pytest.mark.parametrize("nx,ny,kvar,distort,sx,fext", [(130, 17, True, False, 0.37, False),
                                                       (64, 48, False, True, 1.0, True),
                                                       (3, 5, False, False, 1.0, True),
                                                       (67, 9, True, False, 1.0, False)])
def test_explicit_multistep_equals_single_steps(gf, prec, nx, ny, kvar, distort, sx, fext):
    # a multi-step call (state buffers swapped on the device between its steps) gives the
    # trajectory of single-step calls bit for bit, in every precision (FP32 included), on partial
    # tiles, with fixed DOFs, an external force, per-element K and a caller's initial acceleration
    coor, conn, dofs, q, inv, Ke, Ge, u, v, a = _explicit_setup(nx, ny, kvar=kvar, distort=distort)
    coor = coor * np.array([sx, 1.0])
    fe = 1e-2 * oracle.randn(56, inv.size) if fext else None
    p = getattr(gf, prec)
    X1 = gf.ExplicitQuad4(gf.Vector(conn, dofs), coor, Ke, Ge, inv, 0.01, fext=fe)
    X2 = gf.ExplicitQuad4(gf.Vector(conn, dofs), coor, Ke, Ge, inv, 0.01, fext=fe)
    assert X1.structured
    a = 1e-3 * oracle.randn(57, inv.size) * (inv != 0)  # a caller's acceleration, read by the first step
    X1.set_state(u, v, a)
    X2.set_state(u, v, a)
    X1.step(6, p, store_fint=True)
    for _ in range(6):
        X2.step(1, p, store_fint=True)
    for s1, s2 in zip(X1.state(), X2.state()):
        assert np.array_equal(host(s1), host(s2))
    assert np.array_equal(host(X1.fint()), host(X2.fint()))
    # and a second call continues from the stored state
    X1.step(3, p)
    for _ in range(3):
        X2.step(1, p)
    for s1, s2 in zip(X1.state(), X2.state()):
        assert np.array_equal(host(s1), host(s2))


@pytest.mark.parametrize("shuffle,kvar,distort", [(False, False, False), (False, True, False), (True, False, False),
                                                 (False, False, True)])
def test_explicit_fp32_tolerance(gf, shuffle, kvar, distort):
    coor, conn, dofs, q, inv, Ke, Ge, u, v, a = _explicit_setup(64, 48, shuffle=shuffle, kvar=kvar, distort=distort)
    vec = gf.Vector(conn, dofs)
    X = gf.ExplicitQuad4(vec, coor, Ke, Ge, inv, 0.01)
    X.set_state(u, v, a)
    fint = oracle.vv_steps(1, QUAD4_PLANAR, conn, dofs, q.dNx, q.dV(), Ke, Ge, inv, 0.01, u, v, a)
    X.step(1, gf.F32, store_fint=True)
    assert relerr(host(X.fint()), fint) < TOL_F32
    gu, gv_, ga = (host(t) for t in X.state())
    assert np.array_equal(gu, u)  # displacement update is FP64 in both
    assert relerr(ga, a) < TOL_F32


@pytest.mark.parametrize("nx,ny,sx,sy", [(130, 17, 2.0, 0.5), (64, 48, 0.25, 3.0)])
def test_explicit_fp32_rectangles(gf, nx, ny, sx, sy):
    # Regular coordinates (i sx, j sy) with sx != sy: the FP32 step uses the closed-form rectangle
    # body on partly filled tiles; f_int and a against the FP64 oracle (u is FP64 but its second
    # step carries the first step's FP32 acceleration, weighted by dt^2 / 2)
    coor, conn, dofs, q, inv, Ke, Ge, u, v, a = _explicit_setup(nx, ny, kvar=True)
    coor = coor * np.array([sx, sy])
    vec_o = oracle.Vector(conn, dofs)
    q = oracle.Quadrature(QUAD4_PLANAR, vec_o.AsElement(coor))
    xi, w = oracle.quad_rule(QUAD4_PLANAR, True)
    qn = oracle.Quadrature(QUAD4_PLANAR, vec_o.AsElement(coor), xi, w)
    inv_all = oracle.diag_inverse(oracle.diag_assemble(conn, dofs, qn.Int_N_scalar_NT_dV(np.ones((q.nelem, 4)))))
    inv = np.where(inv == 0.0, 0.0, inv_all)
    X = gf.ExplicitQuad4(gf.Vector(conn, dofs), coor, Ke, Ge, inv, 0.01)
    assert X.structured
    X.set_state(u, v, a)
    fint = oracle.vv_steps(2, QUAD4_PLANAR, conn, dofs, q.dNx, q.dV(), Ke, Ge, inv, 0.01, u, v, a)
    X.step(2, gf.F32, store_fint=True)
    assert relerr(host(X.fint()), fint) < TOL_F32
    gu, gv_, ga = (host(t) for t in X.state())
    assert relerr(ga, a) < TOL_F32 and relerr(gu, u) < 1e-8


@pytest.mark.parametrize("nx,ny,shuffle,kvar,distort,sx", [(64, 64, False, False, False, 1.0),
                                                          (130, 17, False, True, False, 0.37),
                                                          (24, 20, True, True, False, 1.0),
                                                          (40, 24, False, False, True, 1.0)])
def test_explicit_fp64_fast(gf, nx, ny, shuffle, kvar, distort, sx):
    """GF_F64_FAST (FP64 with FMA, one division per Gauss point, the closed-form rectangle body on
    Regular coordinates) against the reference-order oracle: f_int and a within 1e-14 relative after
    one step and after ten."""
    coor, conn, dofs, q, inv, Ke, Ge, u, v, a = _explicit_setup(nx, ny, shuffle=shuffle, kvar=kvar, distort=distort)
    if sx != 1.0:
        coor = coor * np.array([sx, 1.0])
        vec_o = oracle.Vector(conn, dofs)
        q = oracle.Quadrature(QUAD4_PLANAR, vec_o.AsElement(coor))
        xi, w = oracle.quad_rule(QUAD4_PLANAR, True)
        qn = oracle.Quadrature(QUAD4_PLANAR, vec_o.AsElement(coor), xi, w)
        inv_all = oracle.diag_inverse(oracle.diag_assemble(conn, dofs, qn.Int_N_scalar_NT_dV(np.ones((q.nelem, 4)))))
        inv = np.where(inv == 0.0, 0.0, inv_all)
    X = gf.ExplicitQuad4(gf.Vector(conn, dofs), coor, Ke, Ge, inv, 0.01)
    assert X.structured == (not shuffle)
    X.set_state(u, v, a)
    for nsteps in (1, 9):
        fint = oracle.vv_steps(nsteps, QUAD4_PLANAR, conn, dofs, q.dNx, q.dV(), Ke, Ge, inv, 0.01, u, v, a)
        X.step(nsteps, gf.F64_FAST, store_fint=True)
        gu, gv_, ga = (host(t) for t in X.state())
        assert relerr(host(X.fint()), fint) < 1e-14
        assert relerr(ga, a) < 1e-14 and relerr(gv_, v) < 1e-14 and relerr(gu, u) < 1e-14


def _fint_extended(u, G, K, nx, ny, rel=1e-25):
    """f_int of the Regular nx x ny (h = 1) plane-strain mesh at displacement u, in extended precision
    (long double) with the exact geometry (J = I / 2, dN/dx = 2 dN/dxi, dV = 1 / 4), 2x2 Gauss.  Only
    element rows touching a node with |u| > rel max|u| are evaluated: the nodes elsewhere carry
    forces below rel of the norm.  Returns (f, node mask of the evaluated rows)."""
    L = np.longdouble
    U = u.reshape(ny + 1, nx + 1, 2)
    moving = np.where(np.abs(U).max(axis=(1, 2)) > rel * np.abs(U).max())[0]
    ey0, ey1 = max(int(moving.min()) - 1, 0), min(int(moving.max()) + 1, ny)  # element rows [ey0, ey1)
    ex = np.arange(nx)
    ey = np.arange(ey0, ey1)
    n0 = (ey[:, None] * (nx + 1) + ex[None, :]).ravel()
    nodes = np.stack((n0, n0 + 1, n0 + nx + 2, n0 + nx + 1), axis=1)  # Mesh::Quad4::Regular order
    e = (ey[:, None] * nx + ex[None, :]).ravel()
    ue = u.reshape(-1, 2)[nodes].astype(L)
    KL, GL = K[e].astype(L), G[e].astype(L)
    a = L(1) / np.sqrt(L(3))
    sm = np.array([-1, 1, 1, -1], dtype=L)
    tm = np.array([-1, -1, 1, 1], dtype=L)
    f = np.zeros(ue.shape, dtype=L)
    for xi, eta in ((-a, -a), (a, -a), (a, a), (-a, a)):
        dx = 2 * sm * (1 + tm * eta) / 4
        dy = 2 * tm * (1 + sm * xi) / 4
        gxx, gyy = ue[:, :, 0] @ dx, ue[:, :, 1] @ dy
        gxy, gyx = ue[:, :, 1] @ dx, ue[:, :, 0] @ dy
        tr = gxx + gyy
        sxx = KL * tr + 2 * GL * (gxx - tr / 3)
        syy = KL * tr + 2 * GL * (gyy - tr / 3)
        sxy = GL * (gxy + gyx)
        f[:, :, 0] += (np.outer(sxx, dx) + np.outer(sxy, dy)) / 4
        f[:, :, 1] += (np.outer(sxy, dx) + np.outer(syy, dy)) / 4
    out = np.zeros(u.size, dtype=L)
    np.add.at(out, (2 * nodes[:, :, None] + np.arange(2)).ravel(), f.ravel())
    mask = np.zeros(u.size // 2, dtype=bool)
    mask[ey0 * (nx + 1):(ey1 + 1) * (nx + 1)] = True
    return out.astype(np.float64), np.repeat(mask, 2)


def test_explicit_full_size_fp32_vs_fp64(gf):
    """The Quad4 2048 x 2048 explicit workload (log-normal G, sigma 0.3; K 10; bottom edge fixed, top
    edge driven at v_x 1e-4; dt 0.01), 1000 velocity-Verlet steps: the FP32 element body against the
    FP64 reference-order path (bit-exact with the oracle on the smaller meshes above), u and f_int
    within 1e-6 relative; the FP64 fast body within 1e-12 of it.  At the same displacement the fast
    body is within 1e-14 of an extended-precision evaluation with exact geometry, while the
    reference order is not: it forms the Jacobian from absolute coordinates (up to 2048), which
    costs it about |x| eps of relative error in dN/dx."""
    import torch

    n = 2048
    mesh = gf.Mesh.Quad4.Regular(n, n)
    coor, conn, dofs = mesh.coor, mesh.conn, mesh.dofs
    vec = gf.Vector(conn, dofs)
    nodal = gf.Element.Quad4.QuadraturePlanar(vec.AsElement(coor), gf.Element.Quad4.Nodal.xi(),
                                              gf.Element.Quad4.Nodal.w())
    M = gf.MatrixDiagonal(vec, None)
    M.assemble(nodal.Int_N_scalar_NT_dV(torch.ones((mesh.nelem, 4), dtype=torch.float64, device="cuda")))
    inv = M.Solve(torch.ones(vec.ndof, dtype=torch.float64, device="cuda"))
    del nodal
    top = mesh.nodesTopEdge
    fixed = np.concatenate((dofs[mesh.nodesBottomEdge].ravel(), dofs[top].ravel()))
    inv[torch.from_numpy(fixed).cuda()] = 0.0
    G = np.exp(0.3 * gf.randn(2, mesh.nelem)).astype(np.float32)
    K = np.full(mesh.nelem, 10.0, dtype=np.float32)
    u0 = np.zeros(vec.ndof)
    v0 = np.zeros(vec.ndof)
    v0[dofs[top, 0]] = 1e-4
    X = gf.ExplicitQuad4(vec, coor, K, G, inv, 0.01)
    out = {}
    for prec in (gf.F64, gf.F32, gf.F64_FAST):
        X.set_state(u0, v0, np.zeros(vec.ndof))
        X.step(999, prec)
        s999 = [host(t).copy() for t in X.state()]
        X.step(1, prec, store_fint=True)
        out[prec] = (host(X.state()[0]), host(X.fint()), s999)
    u64, f64, s999 = out[gf.F64]
    assert relerr(out[gf.F32][0], u64) < 1e-6
    assert relerr(out[gf.F32][1], f64) < 1e-6
    assert relerr(out[gf.F64_FAST][0], u64) < 1e-12
    assert relerr(out[gf.F64_FAST][1], f64) < 1e-12
    # one step of the fast body from the reference path's state: the same displacement
    X.set_state(*s999)
    X.step(1, gf.F64_FAST, store_fint=True)
    assert np.array_equal(host(X.state()[0]), u64)
    ffast = host(X.fint())
    fext, mask = _fint_extended(u64, G.astype(np.float64), K.astype(np.float64), n, n)
    nrm = np.linalg.norm(fext)
    assert np.linalg.norm(f64[~mask]) < 1e-20 * nrm and np.linalg.norm(ffast[~mask]) < 1e-20 * nrm
    e_fast = np.linalg.norm(ffast[mask] - fext[mask]) / nrm
    e_ref = np.linalg.norm(f64[mask] - fext[mask]) / nrm
    e_f32 = relerr(out[gf.F32][1], f64)
    print(f"\n2048^2 x 1000: FP32 u {relerr(out[gf.F32][0], u64):.2e} f_int {e_f32:.2e}; FP64 fast u "
          f"{relerr(out[gf.F64_FAST][0], u64):.2e} f_int {relerr(out[gf.F64_FAST][1], f64):.2e}; at u_1000 vs "
          f"extended precision: fast {e_fast:.2e}, reference order {e_ref:.2e}")
    assert e_fast < 1e-14
    assert relerr(ffast, f64) < 1e-12


def test_explicit_full_size_finelayer_fp32_vs_fp64(gf):
    """The general-mesh path at full size: FineLayer(2187, 2400, nfine 1800) (3.95 M elements, element
    sizes 1 to 81), log-normal G (sigma 0.3), K 10, bottom edge fixed, top edge driven at v_x 1e-4,
    a random initial displacement (1e-3, every free DOF, so that every element is loaded), dt 0.01,
    1000 steps, one fused kernel per step.  The FP64 reference-order path is bit-exact with the
    oracle on the FineLayer meshes of test_gpu_finelayer; here the FP32 body stays within 1e-6 of it
    (u and f_int) and the FP64 fast body within 1e-11."""
    import torch

    mesh = gf.Mesh.Quad4.FineLayer(2187, 2400, 1.0, 1800)
    coor, conn, dofs = mesh.coor, mesh.conn, mesh.dofs
    vec = gf.Vector(conn, dofs)
    nodal = gf.Element.Quad4.QuadraturePlanar(vec.AsElement(coor), gf.Element.Quad4.Nodal.xi(),
                                              gf.Element.Quad4.Nodal.w())
    M = gf.MatrixDiagonal(vec, None)
    M.assemble(nodal.Int_N_scalar_NT_dV(torch.ones((mesh.nelem, 4), dtype=torch.float64, device="cuda")))
    inv = M.Solve(torch.ones(vec.ndof, dtype=torch.float64, device="cuda"))
    del nodal
    top = mesh.nodesTopEdge
    fixed = np.concatenate((dofs[mesh.nodesBottomEdge].ravel(), dofs[top].ravel()))
    inv[torch.from_numpy(fixed).cuda()] = 0.0
    G = np.exp(0.3 * gf.randn(3, mesh.nelem)).astype(np.float32)
    K = np.full(mesh.nelem, 10.0, dtype=np.float32)
    v0 = np.zeros(vec.ndof)
    v0[dofs[top, 0]] = 1e-4
    u0 = 1e-3 * gf.randn(4, vec.ndof)
    u0[fixed] = 0.0
    X = gf.ExplicitQuad4(vec, coor, K, G, inv, 0.01)
    assert X.path == "blocks"
    out = {}
    for prec in (gf.F64, gf.F32, gf.F64_FAST):
        X.set_state(u0, v0, np.zeros(vec.ndof))
        X.step(1000, prec, store_fint=True)
        out[prec] = (host(X.state()[0]), host(X.fint()))
    u64, f64 = out[gf.F64]
    e = {p: (relerr(out[p][0], u64), relerr(out[p][1], f64)) for p in (gf.F32, gf.F64_FAST)}
    print(f"\nFineLayer x 1000: FP32 u {e[gf.F32][0]:.2e} f_int {e[gf.F32][1]:.2e}; FP64 fast u "
          f"{e[gf.F64_FAST][0]:.2e} f_int {e[gf.F64_FAST][1]:.2e}")
    assert np.abs(u64).max() > 0.0
    assert e[gf.F32][0] < 1e-6 and e[gf.F32][1] < 1e-6
    assert e[gf.F64_FAST][0] < 1e-11 and e[gf.F64_FAST][1] < 1e-11


@pytest.mark.parametrize("shuffle", [False, True])
def test_explicit_energy(gf, shuffle, tmp_path):
    """Per-step energies of the explicit step (Elastic-VelocityVerlet/main.cpp:155-160) against
    independent evaluations: Ekin = sum m v^2 / 2 (numpy) and Epot = sum_q sigma : eps dV / 2 from the
    oracle's quadrature of the same displacement; then the example's HDF5 output of the series."""
    from goosefem_amd import h5

    coor, conn, dofs, q, inv, Ke, Ge, u, v, a = _explicit_setup(24, 20, shuffle=shuffle, kvar=True)
    mass = np.where(inv > 0, 1.0 / np.where(inv > 0, inv, 1.0), 0.75)  # any positive mass of every DOF
    X = gf.ExplicitQuad4(gf.Vector(conn, dofs), coor, Ke, Ge, inv, 0.01)
    X.set_state(u, v, a)
    with pytest.raises(gf.GooseFEMError):
        X.energy(mass)  # no f_int stored yet
    hist = gf.to_device(np.zeros((6, 2)))
    for inc in range(6):
        X.step(1, gf.F64, store_fint=True)
        X.energy(mass, out=hist[inc])
    gu, gv, _ = (host(t) for t in X.state())
    ek, ep = host(hist)[-1]
    assert abs(ek - 0.5 * np.sum(mass * gv * gv)) <= 1e-13 * abs(ek)
    ov = oracle.Vector(conn, dofs)
    eps = q.SymGradN_vector(ov.AsElement(gu.reshape(-1, 2)))
    sig = oracle.elastic_stress(eps, np.repeat(Ke[:, None], 4, 1), np.repeat(Ge[:, None], 4, 1))
    w = 0.5 * np.einsum("eqij,eqji->eq", sig, eps)
    assert abs(ep - np.sum(w * q.dV())) <= 1e-12 * abs(ep)
    path = str(tmp_path / "example.hdf5")
    with h5.File(path, h5.File.Overwrite) as f:
        H = host(hist)
        for inc in range(6):
            h5.dump(f, "/global/Ekin", H[inc, 0], inc)
            h5.dump(f, "/global/Epot", H[inc, 1], inc)
        h5.dump(f, "/mesh/conn", conn.astype(np.uint64))
        h5.dump(f, "/mesh/coor", coor)
        h5.dump(f, "/mesh/disp", gu.reshape(-1, 2))
    with h5.File(path, "r") as f:
        assert np.array_equal(h5.load(f, "/global/Epot"), H[:, 1])
        assert np.array_equal(h5.load(f, "/mesh/disp"), gu.reshape(-1, 2))


def test_explicit_rejects_bad_input(gf):
    import ctypes

    m = gf.Mesh.Hex8.Regular(2, 2, 2)
    vec = gf.Vector(m.conn, m.dofs)
    with pytest.raises(gf.GooseFEMError):
        gf.ExplicitQuad4(vec, np.zeros((27, 3)), np.ones(8), np.ones(8), np.ones(81), 0.01)
    f = gf.lib().gf_quad_int_gradN_dot_tensor2_dV
    f.restype = ctypes.c_int
    assert f(None, None, None, ctypes.c_int(0), None) == gf.GF_EINVAL


def test_zero_element_inputs_fail_cleanly(gf):
    # a mesh with no elements is rejected with an error (GF_EINVAL) by the calls that move element
    # arrays to or from the device (an empty array has no data pointer): no launch with an empty
    # grid, no fault
    conn = np.zeros((0, 4), dtype=np.int64)
    dofs = np.arange(8).reshape(4, 2)
    calls = [
        lambda: gf.Vector(conn, dofs).AsElement(np.ones(8)),
        lambda: gf.Vector(conn, dofs).AssembleDofs(np.zeros((0, 4, 2))),
        lambda: gf.Element.Quad4.QuadraturePlanar(np.zeros((0, 4, 2))),
        lambda: gf.Element.Quad4.Quadrature(np.zeros((0, 4, 2))),
        lambda: gf.Element.Hex8.Quadrature(np.zeros((0, 8, 3))),
        lambda: (lambda A: (A.assemble(np.zeros((0, 8, 8))), A.Todense()))(gf.Matrix(conn, dofs)),
        lambda: (lambda A: (A.assemble(np.zeros((0, 8, 8))), A.data))(gf.MatrixDiagonal(conn, dofs)),
        lambda: gf.elemmat_dot(np.zeros((0, 8, 8)), np.zeros((0, 8))),
        lambda: gf.LinearHardening(np.zeros((0, 4)), 1.0, 0.01, 0.1),
    ]
    for call in calls:
        with pytest.raises(gf.GooseFEMError):
            call()
    # the library is usable afterwards
    c, cn = oracle.quad4_regular(3, 2)
    v = gf.Vector(cn, oracle.dofs(c.shape[0], 2))
    assert np.array_equal(host(v.AsElement(c)), oracle.Vector(cn, oracle.dofs(c.shape[0], 2)).AsElement(c))
