"""C-ABI checks that need no GPU: the library loads, exports every symbol declared in include/*.h,
and its host-side set-up logic (mesh bookkeeping, periodic tyings) is bit-exact with the
reference's golden arrays and with the oracle."""

import ctypes
import glob
import os
import re

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols(header="goosefem_amd.h"):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(gf_\w+)\s*\(", src, flags=re.M)))


@pytest.fixture(scope="module")
def gf():
    import goosefem_amd

    goosefem_amd.lib()
    return goosefem_amd


def test_every_header_belongs_to_a_library():
    assert sorted(os.path.basename(h) for h in glob.glob(os.path.join(ROOT, "include", "*.h"))) == [
        "goosefem_amd.h", "goosefem_amd_h5.h"]


def test_library_exports_every_declared_symbol(gf):
    syms = declared_symbols()
    assert len(syms) > 60
    lib = gf.lib()
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_h5_library_exports_every_declared_symbol():
    from goosefem_amd import h5

    syms = declared_symbols("goosefem_amd_h5.h")
    assert len(syms) == 8
    lib = h5.lib()
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_single_hip_runtime(gf):
    assert len(gf._hip_runtimes_mapped()) == 1


def test_quad4_mesh_golden(gf, fixtures):
    g = fixtures["MeshQuad4_Regular_5_3"]
    m = gf.Mesh.Quad4.Regular(5, 3)
    assert np.array_equal(m.coor, np.array(g["coor_"]))
    assert np.array_equal(m.conn, np.array(g["conn_"]))
    assert np.array_equal(m.dofs, np.array(g["dofs_"]))
    assert np.array_equal(m.nodesPeriodic, np.array(g["nodesPeriodic_"]))
    assert np.array_equal(m.dofsPeriodic, np.array(g["dofsPeriodic_"]))
    for k, v in g.items():
        if k.startswith("nodes") and k != "nodesPeriodic_":
            got = getattr(m, k[:-1])
            assert np.array_equal(np.atleast_1d(got), np.atleast_1d(v)), k


def test_hex8_mesh_golden(gf, fixtures):
    g = fixtures["MeshHex8_Regular_3_3_3"]
    m = gf.Mesh.Hex8.Regular(3, 3, 3)
    assert np.array_equal(m.coor, np.array(g["coor_"]))
    assert np.array_equal(m.conn, np.array(g["conn_"]))
    assert np.array_equal(m.dofs, np.array(g["dofs_"]))
    assert np.array_equal(m.nodesPeriodic, np.array(g["nodesPeriodic_"]))
    assert np.array_equal(m.dofsPeriodic, np.array(g["dofsPeriodic_"]))
    for k, v in g.items():
        if k.startswith("nodes") and k != "nodesPeriodic_":
            got = getattr(m, k[:-1])
            assert np.array_equal(np.atleast_1d(got), np.atleast_1d(v)), k


@pytest.mark.parametrize("n", [(4, 3, 5), (7, 2, 3), (12, 12, 12)])
def test_hex8_periodic_vs_oracle(gf, n):
    m = gf.Mesh.Hex8.Regular(*n)
    assert np.array_equal(m.nodesPeriodic, oracle.hex8_nodes_periodic(*n))
    c, conn = oracle.hex8_regular(*n)
    assert np.array_equal(m.conn, conn) and np.array_equal(m.coor, c)


def test_elem2node_coordination_golden(gf, fixtures):
    m = gf.Mesh.Quad4.Regular(3, 3)
    assert np.array_equal(gf.Mesh.coordination(m.conn), fixtures["Mesh_coordination_Regular_3_3"])
    assert gf.Mesh.elem2node(m.conn) == fixtures["Mesh_elem2node_Regular_3_3"]
    m = gf.Mesh.Quad4.Regular(4, 4)
    d = m.dofsPeriodic
    assert np.array_equal(gf.Mesh.coordination(d), np.array(fixtures["test_Mesh_coordination_dofs_4_4"]).ravel())
    assert gf.Mesh.node2dof(d) == fixtures["test_Mesh_node2dof_4_4"]
    assert gf.Mesh.nodaltyings(d) == fixtures["test_Mesh_nodaltyings_4_4"]
    real = np.sort(np.array(fixtures["test_Mesh_elem2node_periodic_4_4"]), axis=1)
    assert np.array_equal(np.array(gf.Mesh.elem2node(m.conn, d)), real)


@pytest.mark.parametrize("nx,ny", [(5, 4), (16, 9)])
def test_tyings_vs_oracle(gf, nx, ny):
    m = gf.Mesh.Quad4.Regular(nx, ny)
    ctrl = gf.Tyings.Control(m.coor, m.dofs)
    co, do, cd, cn = oracle.tyings_control(m.coor, m.dofs)
    assert np.array_equal(ctrl.coor, co) and np.array_equal(ctrl.dofs, do)
    assert np.array_equal(ctrl.controlDofs, cd) and np.array_equal(ctrl.controlNodes, cn)
    iip = np.concatenate((cd.ravel(), do[m.nodesOrigin]))
    ty = gf.Tyings.Periodic(ctrl.coor, ctrl.dofs, ctrl.controlDofs, m.nodesPeriodic, iip)
    ref = oracle.tyings_periodic(co, do, cd, oracle.quad4_nodes_periodic(nx, ny), iip)
    assert np.array_equal(ty.dofs, ref["dofs"]) and np.array_equal(ty.control, ref["control"])
    assert (ty.nnu, ty.nnp, ty.nni, ty.nnd) == (ref["nnu"], ref["nnp"], ref["nni"], ref["nnd"])
    for name in ("Cdi", "Cdu", "Cdp"):
        a, b = getattr(ty, name + "_csr"), ref[name + "_csr"]
        for x, y in zip(a, b):
            assert np.array_equal(x, y), name


@pytest.mark.parametrize("nx,ny,nz", [(4, 3, 5), (3, 3, 3)])
def test_tyings_vs_oracle_hex8(gf, nx, ny, nz):
    """Tyings::Control + Tyings::Periodic on a 3-D mesh (TyingsPeriodic.h:82-121, 238-343,
    361-392): reordered dofs, control dofs and nodes, sizes and the Cdi / Cdu / Cdp CSR arrays
    equal the oracle's, entry for entry."""
    m = gf.Mesh.Hex8.Regular(nx, ny, nz)
    ctrl = gf.Tyings.Control(m.coor, m.dofs)
    co, do, cd, cn = oracle.tyings_control(m.coor, m.dofs)
    assert np.array_equal(ctrl.coor, co) and np.array_equal(ctrl.dofs, do)
    assert np.array_equal(ctrl.controlDofs, cd) and np.array_equal(ctrl.controlNodes, cn)
    ties = m.nodesPeriodic
    assert np.array_equal(ties, oracle.hex8_nodes_periodic(nx, ny, nz))
    iip = np.concatenate((cd.ravel(), do[m.nodesOrigin]))
    ty = gf.Tyings.Periodic(ctrl.coor, ctrl.dofs, ctrl.controlDofs, ties, iip)
    ref = oracle.tyings_periodic(co, do, cd, ties, iip)
    assert np.array_equal(ty.dofs, ref["dofs"]) and np.array_equal(ty.control, ref["control"])
    assert (ty.nnu, ty.nnp, ty.nni, ty.nnd) == (ref["nnu"], ref["nnp"], ref["nni"], ref["nnd"])
    assert ty.nnd == 3 * len(ties) and ty.nnp == 12
    for name in ("Cdi", "Cdu", "Cdp"):
        a, b = getattr(ty, name + "_csr"), ref[name + "_csr"]
        for x, y in zip(a, b):
            assert np.array_equal(x, y), name


@pytest.mark.parametrize("dims", [(5, 4), (3, 4, 3)])
def test_tyings_csr_matches_dense(gf, dims):
    """The (rowptr, col, val, ncols) forms that sparse callers use expand to the dense Cdi / Cdu /
    Cdp properties (TyingsPeriodic.h:361-392), shapes included."""
    import scipy.sparse as sp

    m = gf.Mesh.Quad4.Regular(*dims) if len(dims) == 2 else gf.Mesh.Hex8.Regular(*dims)
    ctrl = gf.Tyings.Control(m.coor, m.dofs)
    iip = np.concatenate((ctrl.controlDofs.ravel(), ctrl.dofs[m.nodesOrigin]))
    ty = gf.Tyings.Periodic(ctrl.coor, ctrl.dofs, ctrl.controlDofs, m.nodesPeriodic, iip)
    for name, ncols in (("Cdi", ty.nni), ("Cdu", ty.nnu), ("Cdp", ty.nnp)):
        rp, cl, vl, nc = getattr(ty, name + "_csr")
        assert nc == ncols and rp.size - 1 == ty.nnd, name
        dense = sp.csr_matrix((vl, cl, rp), shape=(rp.size - 1, nc)).toarray()
        assert np.array_equal(dense, getattr(ty, name)), name


def test_tyings_hex8_affine_closed_form(gf):
    """u_d = u_i + (x_d - x_i) . ctrl (TyingsPeriodic.h:244-256) in 3-D, on the product's own Cdi:
    an affine field u = F x with the control node j holding row j of F is reproduced exactly on
    every dependent DOF, and Cdi = [Cdu | Cdp]."""
    nx, ny, nz = 4, 3, 5
    m = gf.Mesh.Hex8.Regular(nx, ny, nz)
    ctrl = gf.Tyings.Control(m.coor, m.dofs)
    iip = np.concatenate((ctrl.controlDofs.ravel(), ctrl.dofs[m.nodesOrigin]))
    ty = gf.Tyings.Periodic(ctrl.coor, ctrl.dofs, ctrl.controlDofs, m.nodesPeriodic, iip)
    F = np.array([[0.01, 0.03, -0.007], [-0.02, 0.005, 0.011], [0.004, -0.013, 0.021]])
    co = ctrl.coor
    nnode = m.coor.shape[0]
    u = np.zeros((co.shape[0], 3))
    u[:nnode] = m.coor @ F.T
    u[nnode:] = F
    d = ty.dofs
    ui = np.zeros(ty.nni)
    ud = np.zeros(ty.nnd)
    indep = d < ty.nni
    ui[d[indep]] = u[indep]
    ud[d[~indep] - ty.nni] = u[~indep]
    Cdi, Cdu, Cdp = ty.Cdi, ty.Cdu, ty.Cdp
    assert np.abs(Cdi @ ui - ud).max() < 1e-14
    assert np.array_equal(Cdi, np.hstack((Cdu, Cdp)))
    # a sign, index or transposition error in the dependent rows breaks the affine reproduction
    assert np.abs(Cdi @ (ui * (1 + (np.arange(ty.nni) % 2))) - ud).max() > 1e-3


def test_randn_matches_specification(gf):
    assert np.array_equal(gf.randn(3, 1001, 0.5, 2.0), oracle.randn(3, 1001, 0.5, 2.0))


def test_error_path_reports_message(gf):
    n = ctypes.c_int64(0)
    f = gf.lib().gf_mesh_nodeset
    f.restype = ctypes.c_int
    rc = f(ctypes.c_int64(3), ctypes.c_int64(3), ctypes.c_int64(0), b"nodesNowhere", None, ctypes.byref(n))
    assert rc == gf.GF_EINVAL
    assert b"unknown node set" in gf.lib().gf_last_error()


def test_warmstart_weights_file_is_plain_arrays():
    """The committed network weights load without unpickling and carry their training record."""
    import os

    path = os.path.join(ROOT, "goosefem_amd", "data", "warmstart_quad4_512.npz")
    with np.load(path, allow_pickle=False) as f:
        ch = [int(c) for c in f["channels"]]
        L = len(ch)
        nlayers = 2 * L + 2 * (L - 1) + 1
        assert all(f"w{i}" in f.files and f"b{i}" in f.files for i in range(nlayers))
        assert f["w0"].shape == (9 * ch[0] * 1,) and f[f"w{nlayers - 1}"].shape == (9 * 2 * ch[0],)
        assert int(f["meta_n"]) == 512 and float(f["meta_holdout_rel_error"]) < 0.1
